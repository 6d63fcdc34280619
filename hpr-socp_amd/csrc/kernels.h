// kernels.h — device-side data structures and launch wrappers of the gfx950
// HPR iteration.  All launches are asynchronous on the caller's stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hpr {

}  // namespace hpr

#include "hotpath.h"

namespace hpr {

// ---- reductions ---------------------------------------------------------
constexpr int kMaxRed = 8;
struct RedTerm {           // sum_i a[i]*b[i]*(c ? c[i] : 1) or max_i |a[i]| (op = 1)
    const double *a, *b, *c;
    int32_t op;            // 0 = dot, 1 = absmax(a), 2 = sumsq(a - b), 3 = max(a)
};
// Computes nterms reductions over [0, n) into out[0..nterms) (device memory),
// deterministic two-level tree.  partials needs nblocks*kMaxRed doubles.
void launch_reduce(hipStream_t st, int64_t n, const RedTerm *terms, int nterms, double *partials,
                   int nblocks, double *out);
int reduce_blocks(int64_t n);

// ---- elementwise helpers -------------------------------------------------
void launch_axpby(hipStream_t st, int64_t n, double a, const double *x, double b, const double *y,
                  double *out);                       // out = a x + b y
void launch_copy(hipStream_t st, int64_t n, const double *src, double *dst);
void launch_fill(hipStream_t st, int64_t n, double v, double *dst);
void launch_mul(hipStream_t st, int64_t n, const double *a, const double *b, double *out);  // out = a.*b
// precision conversion of the FP32 mirrors (round to nearest / exact widening)
void launch_demote(hipStream_t st, int64_t n, const double *src, float *dst);
void launch_promote(hipStream_t st, int64_t n, const float *src, double *dst);

// ---- residual kernels (compute_Rd!/compute_Rp!, kernels.jl:118-292) ---------
struct ResidNArgs {
    const double *c, *x_bar, *z_bar, *col_norm;
    double *qx, *atdy, *rd;       // qx may be null (noQ); atdy holds A^T y_bar on entry
    double c_scale;
};
// Writes Rd and converts qx/atdy to original units; reduces
// out[0]=<c,x_bar>, out[1]=<x_bar,Qx(scaled)>, out[2]=<z_bar,x_bar>, out[3]=max|Rd|,
// out[4]=max|atdy_org|, out[5]=max|qx_org|.
void launch_resid_n(hipStream_t st, int32_t n, const ResidNArgs &a, double *partials, int nblocks,
                    double *out);
struct ResidMArgs {
    const double *y_bar, *s, *rhs, *AL, *AU, *row_norm;
    double *ax, *rp;              // ax holds A x_bar (scaled) on entry
    const int32_t *row_cone;      // -1 for linear rows, cone id otherwise
    const int32_t *con_idx;       // cone starts (+ sentinel)
    const double *cone_t, *cone_ns; // per cone: t and ||xi|| in original units
    double b_scale;
    int32_t m1;
};
// Pass 1 converts ax to original units; cone stats computed by launch_cone_rp_stats;
// pass 2 writes Rp and reduces out[0]=<y_bar,s>, out[1]=<y_bar,rhs>, out[2]=max|Rp_lin|,
// out[3]=max|Rp_soc|, out[4]=max|Ax_org|.
void launch_ax_to_org(hipStream_t st, int32_t m, double *ax, const double *row_norm, double b_scale);
void launch_cone_rp_stats(hipStream_t st, int32_t ncones, const int32_t *con_idx, const double *ax_org,
                          const double *rhs, const double *row_norm, double b_scale, double *cone_t,
                          double *cone_ns);
void launch_resid_m(hipStream_t st, int32_t m, const ResidMArgs &a, double *partials, int nblocks,
                    double *out);
// compute_bounds_violation! (algorithm.jl:199-256): max violation into out[0].
void launch_bounds_violation(hipStream_t st, int32_t n, int32_t nvar, const int32_t *var_idx,
                             const double *x_bar, const double *l, const double *u,
                             const double *col_norm, double b_scale, double *dx, double *partials,
                             int nblocks, double *out);

// ---- golden section on diagonal Q (unified_golden_Q_diag) ----------------
// out[k] = sum_i atdy[i] * qatdy[i] / (1 + xs[k] * dq[i]) for k < npts
void launch_golden_eval(hipStream_t st, int32_t n, const double *atdy, const double *qatdy,
                        const double *dq, const double *xs, int npts, double *partials, int nblocks,
                        double *out);

// ---- power iteration ------------------------------------------------------
// q = z / ||z|| where nrm holds ||z||^2 (device scalar)
void launch_scale_by_rnorm(hipStream_t st, int64_t n, const double *z, const double *nrm2,
                           double *q, const int32_t *done);
// q = z - lam*q with lam from device scalar
void launch_power_residual(hipStream_t st, int64_t n, const double *z, const double *lam, double *q,
                           const int32_t *done);
// finishes one power step on device: reads sums {||q||^2 (new), ||z||^2, <q,z>} and updates state
struct PowerState {
    double lambda, err;
    int32_t done, iters;
};
void launch_power_check(hipStream_t st, const double *lam, const double *qq, const double *zz,
                        PowerState *ps, double tol, int32_t max_iter);

// ---- scaling (scaling.jl:785-1289, CPU semantics) -------------------------
// out[i] = sqrt(max_k |val| or sum_k |val|) over row i of M (plus row i of Q if given),
// 0 -> 1, NOT clipped.
void launch_row_stat(hipStream_t st, const CsrView &M, const CsrView *Q, int use_sum, double *out);
void launch_block_scale(hipStream_t st, int32_t ncones, const int32_t *idx, double *t, int32_t strategy,
                        int32_t phase, int32_t loc);
void launch_clip(hipStream_t st, int64_t n, double *v);
// val[k] = ((1/r[i]) * val[k]) * (1/c[col[k]])
void launch_scale_csr(hipStream_t st, int32_t nrows, const int32_t *rowptr, const int32_t *col,
                      double *val, const double *r, const double *c);
// vectors: c /= tcol; l,u *= tcol; col_norm *= tcol
void launch_scale_cols(hipStream_t st, int32_t n, const double *tcol, double *c, double *l, double *u,
                       double *col_norm);
// AL, AU, rhs /= trow; row_norm *= trow
void launch_scale_rows(hipStream_t st, int32_t m, const double *trow, double *AL, double *AU,
                       double *rhs, double *row_norm);
void launch_scale_scalar(hipStream_t st, int64_t n, double *v, double mul, double div);
// reductions for bc scaling/norms: finite-sumsq, finite-count, finite-absmax of v
void launch_finite_stats(hipStream_t st, int64_t n, const double *v, double *partials, int nblocks,
                         double *out /* [sumsq, count, absmax] */);
// max(|AL_nInf|,|AU_nInf|): out = [sumsq, absmax]
void launch_bound_stats(hipStream_t st, int32_t m, const double *AL, const double *AU, double *partials,
                        int nblocks, double *out);
// diag extraction + diagonality test: out = [count of rows with |rowsum - |diag|| > rtol*max]
void launch_q_diag(hipStream_t st, int32_t n, const int32_t *rowptr, const int32_t *col,
                   const double *val, double *diag, double *partials, int nblocks, double *out);
// diagonal Q factors, unified_update_Q_factors! (unified_operations.jl:468-482)
void launch_q_factors(hipStream_t st, int32_t n, const double *dq, double sigma, double *fact2,
                      double *fact, double *fact1, double *fact_m);
// initial z_bar and s (process_initial_points!, algorithm.jl:2656-2676)
void launch_init_zbar(hipStream_t st, int32_t n, const double *x_bar, const double *qx,
                      const double *aty, const double *c, double sigma, double *z_bar);
void launch_init_s(hipStream_t st, int32_t m, const double *ax, const double *y, const double *AL,
                   const double *AU, double f1, double *s);
// unscale for collect_results! (algorithm.jl:95-102): out = a * (v / d) or a * (v * d)
void launch_unscale(hipStream_t st, int64_t n, const double *v, const double *d, double a, int mul,
                    double *out);

}  // namespace hpr
namespace hpr {
void launch_scale_csr_t(hipStream_t st, int32_t nrows, const int32_t *rowptr, const int32_t *col,
                        double *val, const double *r, const double *c);
void launch_sym_q(hipStream_t st, int64_t nnz, const double *src, const int32_t *tpos, double *dst);
// dst[k] = src[perm[k]] (fills the slice-major copy of a matrix' values)
void launch_gather_perm(hipStream_t st, int64_t nnz, const double *src, const int32_t *perm, double *dst);
constexpr int kMaxGolden = 64;
// out = (v .* d) ./ s   (warm-start scaling, algorithm.jl:2628, 2646)
void launch_scale_vd(hipStream_t st, int64_t n, const double *v, const double *d, double s, double *out);
}  // namespace hpr
