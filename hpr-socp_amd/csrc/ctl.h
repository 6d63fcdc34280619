// ctl.h — the scalar control logic of the solve loop, shared by the host loop
// and the device controller (k_ctl_event) so that both take the same decisions
// bit for bit: the residual block of compute_residuals!, the M-norm of
// compute_M_norm!, check_restart, and update_sigma!'s proposal, blend,
// correction and clamp.  Every function is __host__ __device__, compiled
// without floating-point contraction, and uses only IEEE-exact operations
// (+ - * / sqrt, comparisons) plus ctl_exp / ctl_log below, which are written
// out here instead of taken from the host and device math libraries (those may
// differ in the last bit).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/hprsocp.h"

namespace hpr {
namespace ctl {

#define HPR_CTL __host__ __device__ inline
// no fused multiply-add inside these functions, on host or device
#define HPR_NOFMA _Pragma("clang fp contract(off)")

// Julia semantics: max/min propagate NaN, clamp = nested ifelse.
HPR_CTL double jmax(double a, double b) {
    HPR_NOFMA
    if (a != a || b != b) return NAN;
    return a > b ? a : (b > a ? b : (signbit(a) ? b : a));
}
HPR_CTL double jmin(double a, double b) {
    HPR_NOFMA
    if (a != a || b != b) return NAN;
    return a < b ? a : (b < a ? b : (signbit(a) ? a : b));
}
HPR_CTL double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

// ln 2 split in two parts (Cody & Waite argument reduction): kLn2Hi has 21
// trailing zero bits, so k * kLn2Hi is exact for |k| < 2^21.
constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInvLn2 = 1.44269504088896338700e+00;

// e^x: x = k ln 2 + r with |r| <= ln 2 / 2, e^r by its Taylor series to r^13
// (truncation below 5e-18 relative), scaled by 2^k.  Error about 1 ulp.
HPR_CTL double ctl_exp(double x) {
    HPR_NOFMA
    if (x != x) return x;
    if (x > 709.78) return INFINITY;
    if (x < -745.2) return 0.0;
    const double kd = floor(x * kInvLn2 + 0.5);
    const double r = (x - kd * kLn2Hi) - kd * kLn2Lo;
    double p = 1.0 / 6227020800.0;  // 1/13!
    const double inv[13] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040, 1.0 / 40320,
                            1.0 / 362880, 1.0 / 3628800, 1.0 / 39916800, 1.0 / 479001600};
    for (int j = 12; j >= 0; --j) p = inv[j] + r * p;
    return ldexp(p, (int)kd);
}

// ln x: x = m 2^k with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s) with
// s = (m - 1) / (m + 1), |s| <= 0.1716, by its odd series to s^25 (truncation
// below 1e-18 relative).  Error about 1 ulp.
HPR_CTL double ctl_log(double x) {
    HPR_NOFMA
    if (x != x || x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    int k = 0;
    double m = frexp(x, &k);  // x = m 2^k, m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        k -= 1;
    }
    const double f = m - 1.0;  // exact (Sterbenz)
    const double s = f / (2.0 + f);
    const double s2 = s * s;
    double t = 1.0 / 25.0;
    for (int j = 11; j >= 0; --j) t = 1.0 / (2 * j + 1) + s2 * t;
    const double kd = (double)k;
    return kd * kLn2Hi + (kd * kLn2Lo + 2.0 * s * t);
}

// ---- residual block, compute_residuals! (algorithm.jl:113-196) ----------------
// Per-instance constants of the unscaled problem.
struct InstConst {
    double norm_b_org, norm_c_org, obj_const;
    int32_t has_m, has_lin, has_con;
};

// rn: n-side reductions (<c, x>, <x, Qx>, <l|u, z>, |Rd|_inf, |ATy|_inf, |Qx|_inf),
// rm: m-side reductions (two dual-objective parts, |Rp_lin|, |Rp_soc|, |Ax|_inf)
// as the residual kernels leave them; bv: bounds violation (iteration 0).
HPR_CTL void residuals(const double *rn, const double *rm, double bv, double bs, double cs, const InstConst &H,
                       bool skip_q, bool first_iter, hprsocp_residuals &R) {
    HPR_NOFMA
    R.primal_obj = bs * cs * (rn[0] + 0.5 * rn[1]) + H.obj_const;
    R.dual_obj = bs * cs * (-0.5 * rn[1] + rn[2]) + H.obj_const;
    if (H.has_m) R.dual_obj += bs * cs * (rm[0] + rm[1]);
    R.rel_gap = fabs(R.primal_obj - R.dual_obj) / (1.0 + jmax(fabs(R.primal_obj), fabs(R.dual_obj)));
    const double qx_inf = skip_q ? 0.0 : rn[5];
    R.err_Rd = rn[3] / (1.0 + jmax(jmax(H.norm_c_org, rn[4]), qx_inf));
    if (H.has_m) {
        const double denom = 1.0 + jmax(H.norm_b_org, rm[4]);
        R.err_Rp = jmax(rm[2], rm[3]) / denom;
        R.err_Rp_linear = H.has_lin ? rm[2] / denom : 0.0;
        R.err_Rp_soc = H.has_con ? rm[3] / denom : 0.0;
    } else {
        R.err_Rp = R.err_Rp_linear = R.err_Rp_soc = 0.0;
    }
    if (first_iter) R.err_Rp = jmax(R.err_Rp, bv);
    R.kkt = jmax(jmax(R.err_Rp, R.err_Rd), R.rel_gap);
}

// ---- compute_M_norm! (algorithm.jl:1302-1359) from its 8 reduced dots --------
// d[0] |dx|^2, d[1] <dQw, dx>, d[2] <dQw, ATdy>, d[3] <ATdy, dx>,
// d[4] <ATdy fact_M, QATdy> (diagonal Q) or <ATdy, QATdy>, d[5] |dQw|^2 or <dw, dQw>, d[7] |dy|^2.
HPR_CTL double mnorm(const double *d, double sig, double lamA, double lamQ, bool q, bool hm, bool qdiag) {
    HPR_NOFMA
    double M1 = 0.0, M2 = (1.0 / sig) * d[0], M3 = 0.0;
    if (q) M2 -= 2.0 * d[1];
    if (hm) {
        M1 = sig * lamA * d[7];
        if (q) M1 -= 2.0 * sig * d[2];
        M2 += 2.0 * d[3];
        if (q) {
            if (qdiag) {
                M3 = d[4];
                M1 += sig * d[5];
            } else {
                M3 = (sig * sig) / (1.0 + sig * lamQ) * d[4];
                M1 += sig * lamQ * d[5];
            }
        }
    } else if (!qdiag && q) {
        M1 = sig * lamQ * d[5];
    }
    M2 += jmax(M1, 0.0);
    const double Mn = jmax(M2, 0.0) + jmax(M3, 0.0);
    return sqrt(Mn);
}

// ---- initialize_restart / check_restart (algorithm.jl:1886-1908, 1016-1099) ----
HPR_CTL void init_restart(hprsocp_restart_info &r) {
    HPR_NOFMA
    r.restart_flag = 0;
    r.first_restart = 1;
    r.last_gap = r.current_gap = r.save_gap = INFINITY;
    r.inner = r.step = r.sufficient = r.necessary = r.long_restarts = r.ratio = r.times = 0;
    r.weighted_norm = r.best_gap = r.best_kkt = INFINITY;
    r.best_sigma = 1.0;
    r.best_iter = 0;
    r.sigma_correction_active = 0;
    r.sigma_correction_hold = 0;
    r.sigma_correction_dir = 0;
}

HPR_CTL void check_restart(hprsocp_restart_info &r, int64_t iter, int64_t check_iter, double sigma, double kkt) {
    HPR_NOFMA
    r.restart_flag = 0;
    if (r.first_restart) {
        if (iter == check_iter) {
            r.first_restart = 0;
            r.restart_flag = 1;
            r.weighted_norm = r.current_gap;
            r.best_gap = r.current_gap;
            r.best_kkt = kkt;
            r.best_sigma = sigma;
        }
        return;
    }
    if (iter % check_iter != 0) return;
    if (r.current_gap < 0) r.current_gap = 1e-6;
    if (r.current_gap <= 0.36 * r.last_gap) {
        r.sufficient++;
        r.restart_flag = 1;
    }
    if (r.current_gap <= 0.8 * r.last_gap && r.current_gap > 1.00 * r.save_gap) {
        r.necessary++;
        r.restart_flag = 2;
    }
    const double fact = 0.15;
    if (((double)r.inner >= jmax(500.0, fact * (double)iter) ||
         (r.inner >= 10000 && r.current_gap <= 0.75 * r.last_gap)) &&
        r.current_gap > 0.9 * r.save_gap) {
        r.long_restarts++;
        r.restart_flag = 3;
    }
    // should_refresh_best_sigma, algorithm.jl:682-693
    const bool improved = kkt < r.best_kkt;
    const bool compat = r.current_gap < r.best_gap && kkt <= 1.01 * r.best_kkt;
    const bool periodic = r.best_iter >= 10 && kkt <= 1.2 * r.best_kkt;
    if (improved || compat || periodic) {
        r.best_gap = r.current_gap;
        r.best_kkt = kkt;
        r.best_sigma = sigma;
        r.best_iter = 0;
    }
    r.save_gap = r.current_gap;
}

// ---- update_sigma! (algorithm.jl:729-990) ------------------------------------
// Model terms a, b, c, d of the proposal from the reduced dots
// s[0] |x_bar - last_x|^2, s[1] <dQw, ATdy>, s[2] |dQw|^2, s[3] <dw, dQw>,
// s[4] <ATdy, QATdy>, s[7] |y_bar - last_y|^2.
HPR_CTL void sigma_terms(const double *s, bool skip_q, bool has_m, bool qd, double lamA, double lamQ, double &a,
                         double &b, double &c, double &dl) {
    HPR_NOFMA
    const double EPS = 1e-16;
    const double dd_x = s[0], dd_y = has_m ? s[7] : 0.0;
    a = 0.0;
    b = dd_x;
    c = 0.0;
    dl = 0.0;
    if (has_m) {
        if (!skip_q) a = lamA * dd_y - 2.0 * s[1];
        else a = lamA * dd_y;
    }
    if (!skip_q) {
        if (qd) {
            const double nq = sqrt(s[2]);
            a += nq * nq;
        } else {
            a += lamQ * s[3];
            if (has_m) {
                c = s[4];
                dl = lamQ;
            }
        }
    }
    a = jmax(a, EPS);
    b = jmax(b, EPS);
}

HPR_CTL double f_dev(double x, double a, double b, double c, double d) {
    HPR_NOFMA
    return a * x + b / x + c * x * x / (1 + d * x);
}

// golden (kernels.jl:6272-6305), at most 200 steps.
HPR_CTL double golden_scalar(double ap, double bp, double cp, double dp, double lo, double hi, double tol) {
    HPR_NOFMA
    const double phi = (sqrt(5.0) - 1.0) / 2.0;
    double a = lo, b = hi, c = b - phi * (b - a), d = a + phi * (b - a);
    double fc = f_dev(c, ap, bp, cp, dp), fd = f_dev(d, ap, bp, cp, dp);
    for (int i = 0; i < 200; ++i) {
        if (fd < fc) {
            a = c;
            c = d;
            fc = fd;
            d = a + phi * (b - a);
            fd = f_dev(d, ap, bp, cp, dp);
        } else {
            b = d;
            d = c;
            fd = fc;
            c = b - phi * (b - a);
            fc = f_dev(c, ap, bp, cp, dp);
        }
        if ((b - a) < tol) break;
    }
    return 0.5 * (a + b);
}

// The model proposal: sqrt(b / a), or the golden-section minimiser of the
// general-Q model; NaN for a diagonal Q with rows (the caller runs the search
// over the device vectors, unified_golden_Q_diag).
HPR_CTL double sigma_model(double a, double b, double c, double dl, bool skip_q, bool has_m, bool qd) {
    HPR_NOFMA
    if (skip_q || !has_m) return sqrt(b / a);
    if (qd) return NAN;
    return golden_scalar(a, b, c, dl, 1e-12, 1e12, 1e-13);
}

// Blend with the best sigma, residual-imbalance correction and clamp
// (algorithm.jl:836-986); updates the correction state and best_iter of r.
HPR_CTL double sigma_next(hprsocp_restart_info &r, const hprsocp_residuals &res, double sm, double sigma_old) {
    HPR_NOFMA
    const double SMIN = 1e-12, SMAX = 1e12, THETA_MAX = 1.0, IMB = 1e6, FLOOR = 2.0, THR = 1e-8, EPS = 1e-16;
    const double LOG100 = ctl_log(100.0);
    const double sref = r.best_sigma, ck = res.kkt, cg = r.current_gap, bk = r.best_kkt, bg = r.best_gap;
    const bool ref_ok = (ck <= 10.0 * bk) && (cg <= 10.0 * bg);
    double pr = sqrt((ck / bk) * (cg / bg));
    pr = jclamp(pr, 1.0, 10.0);
    double sbase = sm;
    if (ref_ok) {
        const double theta = jclamp(ctl_exp(-0.3 * (pr - 1.0)), 0.0, THETA_MAX);
        sbase = ctl_exp(theta * ctl_log(sm) + (1.0 - theta) * ctl_log(sref));
        sbase = jclamp(sbase, SMIN, SMAX);
    }
    const double ep = jmax(res.err_Rp, EPS), ed = jmax(res.err_Rd, EPS);
    const double es = jmin(ep, ed), el = jmax(ep, ed), ir = el / es;
    const bool both = el <= THR;
    const bool one = (es <= 0.5 * THR) && (el >= 2.0 * THR);
    const bool extreme = ir >= IMB;
    const bool trig = !both && (one || extreme);
    const bool keep = !both && ((es <= 0.8 * THR) || (ir >= 0.1 * IMB));
    const int cdir = ed >= ep ? 1 : -1;
    if (both) {
        r.sigma_correction_active = 0;
        r.sigma_correction_hold = -1;
        r.sigma_correction_dir = 0;
    } else if (trig || (r.sigma_correction_active && keep)) {
        r.sigma_correction_active = 1;
        r.sigma_correction_hold = 0;
        r.sigma_correction_dir = cdir;
    } else {
        r.sigma_correction_active = 0;
        r.sigma_correction_hold = -1;
        r.sigma_correction_dir = 0;
    }
    double beta = 0.0;
    if (!both) {
        if (one) beta = 0.5;
        if (extreme) beta = 1.0;
        else if (r.sigma_correction_active) beta = 0.5;
    }
    double kappa = 1.0;
    if (beta > 0.0) {
        double rk = ed / ep;
        if (r.sigma_correction_active) {
            if (r.sigma_correction_dir > 0) rk = jmax(rk, FLOOR);
            else if (r.sigma_correction_dir < 0) rk = jmin(rk, 1.0 / FLOOR);
        }
        kappa = ctl_exp(beta * jclamp(ctl_log(rk), -LOG100, LOG100));
    }
    double sn = sbase * kappa;
    if (r.sigma_correction_active) {
        if (r.sigma_correction_dir > 0) sn = jmax(sn, sigma_old * 1.5);
        else if (r.sigma_correction_dir < 0) sn = jmin(sn, sigma_old / 1.5);
    }
    sn = jclamp(sn, SMIN, SMAX);
    r.best_iter += 1;
    return sn;
}

// ---- the device controller -------------------------------------------------
// Per-instance state of the solve loop kept in device memory when the window-edge
// decisions run on the device (k_ctl_event): restart bookkeeping, residuals, the
// sigma / lambda of the deferred M-norm evaluations, milestones and the record
// of the termination.
struct CtlInst {
    hprsocp_restart_info R;
    hprsocp_residuals rs;      // residuals of the last event
    hprsocp_residuals rs_end;  // at termination
    double mn_sig[2], mn_lamA[2], mn_lamQ[2];
    double sigma_end;
    double sm;  // the sigma model between the two phases of a check with a golden-section search
    int64_t it4, it6, last_event, iter_end, restarts_end;
    int32_t pend[2], first4, first6, status;
    int32_t st_pend;  // termination status between the two phases
    int32_t gold;     // the sigma model comes from this instance's golden-section walk
};

}  // namespace ctl
}  // namespace hpr
