// golden.h — golden-section searches of the sigma update (host control logic).
//
// golden_sequential restates unified_golden_Q_diag (reference
// src/unified_operations.jl:371-449): minimise f(x) = a x + b/x + x^2 g(x) on
// [lo, hi] with the reference's bracket updates, stopping rule
// |hi - lo| <= tol * max(1, |lo|) and 200-step cap; answer (f1 < f2) ? x1 : x2.
//
// golden_speculative visits exactly the same points but evaluates g in
// batches: from a known state the next L steps can only reach the 2^L - 1
// points of a binary decision tree (node 0 follows the known decision; node
// k's children 2k+1 / 2k+2 follow "f1 > f2" true / false), and the points do
// not depend on f.  All tree points go to one batched evaluation, then the
// walk replays the reference's comparisons, re-deriving each state from the
// fully known previous one.
#pragma once
#include <math.h>

#include <vector>

namespace hpr {

struct GoldenState {
    double lo, hi, x1, x2, f1, f2;
};

inline void golden_step(const GoldenState &g, bool branch_a, GoldenState &o, double &new_point) {
    const double phi = (sqrt(5.0) - 1.0) / 2.0;
    o = g;
    if (branch_a) {  // f1 > f2
        o.lo = g.x1;
        o.x1 = g.x2;
        o.f1 = g.f2;
        o.x2 = o.lo + phi * (o.hi - o.lo);
        new_point = o.x2;
    } else {
        o.hi = g.x2;
        o.x2 = g.x1;
        o.f2 = g.f1;
        o.x1 = o.hi - phi * (o.hi - o.lo);
        new_point = o.x1;
    }
}

inline bool golden_continue(const GoldenState &g, int it, double tol) {
    const double lo_abs = fabs(g.lo);
    return fabs(g.hi - g.lo) > tol * (lo_abs > 1.0 ? lo_abs : 1.0) && it < 200;
}

// G: double(double x) -> g(x)
template <class G>
double golden_sequential(double a, double b, double lo, double hi, double tol, G g) {
    const double phi = (sqrt(5.0) - 1.0) / 2.0;
    auto f = [&](double x) { return a * x + b / x + x * x * g(x); };
    GoldenState S{lo, hi, hi - phi * (hi - lo), lo + phi * (hi - lo), 0, 0};
    S.f1 = f(S.x1);
    S.f2 = f(S.x2);
    int it = 0;
    while (golden_continue(S, it, tol)) {
        GoldenState n;
        double p;
        const bool A = S.f1 > S.f2;
        golden_step(S, A, n, p);
        if (A) n.f2 = f(p);
        else n.f1 = f(p);
        S = n;
        ++it;
    }
    return (S.f1 < S.f2) ? S.x1 : S.x2;
}

// The speculative search as a resumable walk, so that many independent searches
// (one per problem instance) can share each batched evaluation:
//   while (!w.done()) { pts = w.next_points(); gs = g(pts); w.consume(gs); }
class GoldenWalk {
public:
    GoldenWalk(double a, double b, double lo, double hi, double tol, int L)
        : a_(a), b_(b), tol_(tol), np_((1 << L) - 1), node_((size_t)((1 << L) - 1)), pts_((size_t)((1 << L) - 1)) {
        const double phi = (sqrt(5.0) - 1.0) / 2.0;
        S_ = GoldenState{lo, hi, hi - phi * (hi - lo), lo + phi * (hi - lo), 0, 0};
    }
    bool done() const { return started_ && !golden_continue(S_, it_, tol_); }
    // points whose g values the next consume() expects
    const std::vector<double> &next_points() {
        if (!started_) {
            init_ = {S_.x1, S_.x2};
            return init_;
        }
        golden_step(S_, S_.f1 > S_.f2, node_[0], pts_[0]);
        for (int k = 0; 2 * k + 2 < np_; ++k) {
            golden_step(node_[k], true, node_[2 * k + 1], pts_[2 * k + 1]);
            golden_step(node_[k], false, node_[2 * k + 2], pts_[2 * k + 2]);
        }
        return pts_;
    }
    void consume(const double *gs) {
        if (!started_) {
            S_.f1 = f_of(S_.x1, gs[0]);
            S_.f2 = f_of(S_.x2, gs[1]);
            started_ = true;
            return;
        }
        int k = -1;
        while (golden_continue(S_, it_, tol_)) {
            const bool A = S_.f1 > S_.f2;
            const int child = k < 0 ? 0 : (A ? 2 * k + 1 : 2 * k + 2);
            if (child >= np_) break;
            GoldenState n;
            double p;
            golden_step(S_, A, n, p);
            const double fv = f_of(p, gs[child]);
            if (A) n.f2 = fv;
            else n.f1 = fv;
            S_ = n;
            ++it_;
            k = child;
        }
    }
    double result() const { return (S_.f1 < S_.f2) ? S_.x1 : S_.x2; }

private:
    double f_of(double x, double gx) const { return a_ * x + b_ / x + x * x * gx; }
    double a_, b_, tol_;
    int np_;
    GoldenState S_{};
    int it_ = 0;
    bool started_ = false;
    std::vector<GoldenState> node_;
    std::vector<double> pts_, init_;
};

// Eval: void(const std::vector<double>& xs, std::vector<double>& gs) -> gs[k] = g(xs[k])
template <class Eval>
double golden_speculative(double a, double b, double lo, double hi, double tol, int L, Eval eval) {
    GoldenWalk w(a, b, lo, hi, tol, L);
    std::vector<double> gs;
    while (!w.done()) {
        const std::vector<double> &pts = w.next_points();
        eval(pts, gs);
        w.consume(gs.data());
    }
    return w.result();
}

}  // namespace hpr
