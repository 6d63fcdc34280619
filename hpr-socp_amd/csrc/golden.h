// golden.h — golden-section searches of the sigma update, shared by the host
// and the device: every function is __host__ __device__, compiled without
// floating-point contraction, and uses IEEE operations only, so a search gives
// the same bits wherever it runs.
//
// golden_sequential restates unified_golden_Q_diag (reference
// src/unified_operations.jl:371-449): minimise f(x) = a x + b/x + x^2 g(x) on
// [lo, hi] with the reference's bracket updates, stopping rule
// |hi - lo| <= tol * max(1, |lo|) and 200-step cap; answer (f1 < f2) ? x1 : x2.
//
// GoldenWalk visits exactly the same points but evaluates g in batches: from a
// known state the next L steps can only reach the 2^L - 1 points of a binary
// decision tree (node 0 follows the known decision; node k's children 2k+1 /
// 2k+2 follow "f1 > f2" true / false), and the points do not depend on f.  All
// tree points go to one batched evaluation, then the walk replays the
// reference's comparisons, re-deriving each state from the fully known previous
// one.  The walk is a plain struct so that many independent searches (one per
// problem instance) can live in device memory and share each evaluation pass.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <vector>

namespace hpr {

#define HPR_GOLD __host__ __device__ inline
#define HPR_GOLD_NOFMA _Pragma("clang fp contract(off)")

struct GoldenState {
    double lo, hi, x1, x2, f1, f2;
};

HPR_GOLD double golden_phi() {
    HPR_GOLD_NOFMA
    return (sqrt(5.0) - 1.0) / 2.0;
}

HPR_GOLD void golden_step(const GoldenState &g, bool branch_a, GoldenState &o, double &new_point) {
    HPR_GOLD_NOFMA
    const double phi = golden_phi();
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

HPR_GOLD bool golden_continue(const GoldenState &g, int it, double tol) {
    HPR_GOLD_NOFMA
    const double lo_abs = fabs(g.lo);
    return fabs(g.hi - g.lo) > tol * (lo_abs > 1.0 ? lo_abs : 1.0) && it < 200;
}

HPR_GOLD double golden_f(double a, double b, double x, double gx) {
    HPR_GOLD_NOFMA
    return a * x + b / x + x * x * gx;
}

// G: double(double x) -> g(x)
template <class G>
double golden_sequential(double a, double b, double lo, double hi, double tol, G g) {
    const double phi = golden_phi();
    GoldenState S{lo, hi, hi - phi * (hi - lo), lo + phi * (hi - lo), 0, 0};
    S.f1 = golden_f(a, b, S.x1, g(S.x1));
    S.f2 = golden_f(a, b, S.x2, g(S.x2));
    int it = 0;
    while (golden_continue(S, it, tol)) {
        GoldenState n;
        double p;
        const bool A = S.f1 > S.f2;
        golden_step(S, A, n, p);
        if (A) n.f2 = golden_f(a, b, p, g(p));
        else n.f1 = golden_f(a, b, p, g(p));
        S = n;
        ++it;
    }
    return (S.f1 < S.f2) ? S.x1 : S.x2;
}

constexpr int kGoldenMaxL = 6;                            // deepest lookahead
constexpr int kGoldenMaxPts = (1 << kGoldenMaxL) - 1;     // 63 points per pass
constexpr int kGoldenL = 5;                               // lookahead of the solver's searches
constexpr int kGoldenPasses = 1 + (200 + kGoldenL - 1) / kGoldenL;  // passes that finish any search

// One resumable search:
//   while (!gw_done(w)) { k = gw_next_points(w, xs); gs = g(xs[0..k)); gw_consume(w, gs); }
struct GoldenWalk {
    double a, b, tol;
    GoldenState S;
    GoldenState node[kGoldenMaxPts];
    int32_t np;       // points of a pass after the first: 2^L - 1
    int32_t it;       // golden steps taken
    int32_t started;  // f(x1), f(x2) known
    int32_t active;   // a search is running (device passes skip the others)
};

HPR_GOLD void gw_init(GoldenWalk &w, double a, double b, double lo, double hi, double tol, int L) {
    HPR_GOLD_NOFMA
    const double phi = golden_phi();
    w.a = a;
    w.b = b;
    w.tol = tol;
    w.S = GoldenState{lo, hi, hi - phi * (hi - lo), lo + phi * (hi - lo), 0, 0};
    w.np = (1 << L) - 1;
    w.it = 0;
    w.started = 0;
    w.active = 1;
}

HPR_GOLD bool gw_done(const GoldenWalk &w) { return w.started && !golden_continue(w.S, w.it, w.tol); }

// Writes the points whose g values the next gw_consume expects; returns their number.
HPR_GOLD int gw_next_points(GoldenWalk &w, double *out) {
    if (!w.started) {
        out[0] = w.S.x1;
        out[1] = w.S.x2;
        return 2;
    }
    golden_step(w.S, w.S.f1 > w.S.f2, w.node[0], out[0]);
    for (int k = 0; 2 * k + 2 < w.np; ++k) {
        golden_step(w.node[k], true, w.node[2 * k + 1], out[2 * k + 1]);
        golden_step(w.node[k], false, w.node[2 * k + 2], out[2 * k + 2]);
    }
    return w.np;
}

HPR_GOLD void gw_consume(GoldenWalk &w, const double *gs) {
    if (!w.started) {
        w.S.f1 = golden_f(w.a, w.b, w.S.x1, gs[0]);
        w.S.f2 = golden_f(w.a, w.b, w.S.x2, gs[1]);
        w.started = 1;
        return;
    }
    int k = -1;
    while (golden_continue(w.S, w.it, w.tol)) {
        const bool A = w.S.f1 > w.S.f2;
        const int child = k < 0 ? 0 : (A ? 2 * k + 1 : 2 * k + 2);
        if (child >= w.np) break;
        GoldenState n;
        double p;
        golden_step(w.S, A, n, p);
        const double fv = golden_f(w.a, w.b, p, gs[child]);
        if (A) n.f2 = fv;
        else n.f1 = fv;
        w.S = n;
        ++w.it;
        k = child;
    }
}

HPR_GOLD double gw_result(const GoldenWalk &w) { return (w.S.f1 < w.S.f2) ? w.S.x1 : w.S.x2; }

// Eval: void(const std::vector<double>& xs, std::vector<double>& gs) -> gs[k] = g(xs[k])
template <class Eval>
double golden_speculative(double a, double b, double lo, double hi, double tol, int L, Eval eval) {
    GoldenWalk w;
    gw_init(w, a, b, lo, hi, tol, L);
    std::vector<double> xs(kGoldenMaxPts), gs;
    while (!gw_done(w)) {
        const int k = gw_next_points(w, xs.data());
        eval(std::vector<double>(xs.begin(), xs.begin() + k), gs);
        gw_consume(w, gs.data());
    }
    return gw_result(w);
}

}  // namespace hpr
