// kernels.h — device-side data structures and launch wrappers of the gfx950
// HPR iteration.  All launches are asynchronous on the caller's stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hpr {

}  // namespace hpr

#include "hotpath.h"

namespace hpr {

// ---- instance segmentation ---------------------------------------------
// An index space [0, N) (columns or rows) cut per problem instance into chunks
// of at most kChunk elements; chunks of instance i are chunks[first[i] ..
// first[i+1]).  Reductions produce one result per instance; kernels that need
// per-instance scalars run one workgroup per chunk.
constexpr int kChunk = 2048;
struct SegChunk {
    int32_t start, end, inst, pad;
};
struct SegPlan {
    const SegChunk *chunks = nullptr;
    const int32_t *first = nullptr;  // ninst + 1
    int32_t nchunks = 0, ninst = 1;
};
// A per-instance scalar in device memory: value of instance i = p[i * stride].
struct InstVal {
    const double *p;
    int32_t stride;
};

// ---- reductions ---------------------------------------------------------
constexpr int kMaxRed = 8;
struct RedTerm {           // sum_i a[i]*b[i]*(c ? c[i] : 1) or max_i |a[i]| (op = 1)
    const double *a, *b, *c;
    int32_t op;            // 0 = dot, 1 = absmax(a), 2 = sumsq(a - b), 3 = max(a)
};
// Computes nterms reductions per instance into out[inst*ostride + k] (device
// memory); deterministic two-level tree (chunks, then the chunks of one
// instance in order).  partials needs sp.nchunks*kMaxRed doubles.
void launch_reduce(hipStream_t st, const SegPlan &sp, const RedTerm *terms, int nterms, double *partials,
                   double *out, int ostride);

// ---- elementwise helpers -------------------------------------------------
void launch_axpby(hipStream_t st, int64_t n, double a, const double *x, double b, const double *y,
                  double *out);                       // out = a x + b y
void launch_copy(hipStream_t st, int64_t n, const double *src, double *dst);
void launch_fill(hipStream_t st, int64_t n, double v, double *dst);
void launch_mul(hipStream_t st, int64_t n, const double *a, const double *b, double *out);  // out = a.*b
// precision conversion of the FP32 mirrors (round to nearest / exact widening)
void launch_demote(hipStream_t st, int64_t n, const double *src, float *dst);
void launch_promote(hipStream_t st, int64_t n, const float *src, double *dst);
// dst = (double)src on the chunks of the active instances whose updates run in FP32
void launch_promote_prec(hipStream_t st, const SegPlan &sp, const float *src, double *dst, const InstScal *isc);
// delta form, on the chunks of the active instances: d = (float)(v - anchor);
// v = anchor + (double)d; ra = xa + sigma ((img - qwa) - c) (qwa may be null);
// rya = A xa - rhs (rows >= m1, the SOC rows) - lambda_A sigma ya
void launch_delta_split(hipStream_t st, const SegPlan &sp, const double *v, const double *anchor, float *d,
                        const InstScal *isc);
void launch_delta_join(hipStream_t st, const SegPlan &sp, const double *anchor, const float *d, double *v,
                       const InstScal *isc);
void launch_delta_ra(hipStream_t st, const SegPlan &sp, const double *xa, const double *img, const double *qwa,
                     const double *c, const InstScal *isc, double *ra);
void launch_delta_rya(hipStream_t st, const SegPlan &sp, int64_t m1, const double *axa, const double *rhs,
                      const double *ya, const InstScal *isc, double *rya);

// ---- residual kernels (compute_Rd!/compute_Rp!, kernels.jl:118-292) ---------
struct ResidNArgs {
    const double *c, *x_bar, *z_bar, *col_norm;
    double *qx, *atdy, *rd;       // qx may be null (noQ); atdy holds A^T y_bar on entry
    const InstScal *isc;          // c_scale per instance
};
// Writes Rd and converts qx/atdy to original units; reduces per instance
// out[0]=<c,x_bar>, out[1]=<x_bar,Qx(scaled)>, out[2]=<z_bar,x_bar>, out[3]=max|Rd|,
// out[4]=max|atdy_org|, out[5]=max|qx_org|.
void launch_resid_n(hipStream_t st, const SegPlan &sp, const ResidNArgs &a, double *partials, double *out,
                    int ostride);
struct ResidMArgs {
    const double *y_bar, *s, *rhs, *AL, *AU, *row_norm;
    double *ax, *rp;              // ax holds A x_bar (original units) on entry
    const int32_t *row_cone;      // -1 for linear rows, cone id otherwise
    const int32_t *con_idx;       // cone starts (+ sentinel)
    const double *cone_t, *cone_ns; // per cone: t and ||xi|| in original units
    const InstScal *isc;          // b_scale per instance
    int32_t m1;
};
// ax to original units (per-instance b_scale)
void launch_ax_to_org(hipStream_t st, const SegPlan &sp, double *ax, const double *row_norm, const InstScal *isc);
// Constraint cones cut into pieces of at most kConePiece rows (one wavefront
// each), so a long cone is summed by many wavefronts: piece p covers rows
// [row[p], min(row[p] + kConePiece, end of its cone)) of cone[p]; the pieces of
// cone c are first[c] .. first[c+1]-1, and `multi` lists the cones with more
// than one piece (their partial sums go through part[]).
constexpr int kConePiece = 2048;
struct ConePieces {
    const int32_t *cone = nullptr, *row = nullptr, *first = nullptr, *multi = nullptr;
    int32_t npieces = 0, nmulti = 0;
    double *part = nullptr;
};
void launch_cone_rp_stats(hipStream_t st, int32_t ncones, const int32_t *con_idx, const int32_t *cone_inst,
                          const ConePieces &pc, const double *ax_org, const double *rhs, const double *row_norm,
                          const InstScal *isc, double *cone_t, double *cone_ns);
// writes Rp and reduces per instance out[0]=<y_bar,s>, out[1]=<y_bar,rhs>, out[2]=max|Rp_lin|,
// out[3]=max|Rp_soc|, out[4]=max|Ax_org|.
void launch_resid_m(hipStream_t st, const SegPlan &sp, const ResidMArgs &a, double *partials, double *out,
                    int ostride);
// compute_bounds_violation! (algorithm.jl:199-256): max violation per instance into out[0].
void launch_bounds_violation(hipStream_t st, const SegPlan &sp, int32_t nvar, const int32_t *var_idx,
                             const int32_t *cone_inst, const double *x_bar, const double *l, const double *u,
                             const double *col_norm, const InstScal *isc, double *dx, double *partials,
                             double *out, int ostride);

// ---- golden section on diagonal Q (unified_golden_Q_diag) ----------------
// per instance i: out[i*ostride + k] = sum_j atdy[j] * qatdy[j] / (1 + xs[i*kMaxGolden + k] * dq[j])
// for k < min(npts, np_inst[i]) (np_inst null: npts; instances with 0 points are skipped)
void launch_golden_eval(hipStream_t st, const SegPlan &sp, const double *atdy, const double *qatdy,
                        const double *dq, const double *xs, int npts, const int32_t *np_inst, double *partials,
                        double *out, int ostride);

// ---- power iteration ------------------------------------------------------
struct PowerState {
    double lambda, err;
    int32_t done, iters;
};
// q = z / ||z|| per instance, nrm2 = ||z||^2 per instance (InstVal)
void launch_scale_by_rnorm(hipStream_t st, const SegPlan &sp, const double *z, InstVal nrm2, double *q,
                           const PowerState *ps);
// q = z - lam*q per instance
void launch_power_residual(hipStream_t st, const SegPlan &sp, const double *z, InstVal lam, double *q,
                           const PowerState *ps);
// finishes one power step per instance: {||q||^2 (new), ||z||^2, <q,z>} -> state
void launch_power_check(hipStream_t st, int32_t ninst, InstVal lam, InstVal qq, InstVal zz, PowerState *ps,
                        double tol, int32_t max_iter);

// ---- scaling (scaling.jl:785-1289, CPU semantics) -------------------------
// out[i] = sqrt(max_k |val| or sum_k |val|) over row i of M (plus row i of Q if given),
// 0 -> 1, NOT clipped.
void launch_row_stat(hipStream_t st, const CsrView &M, const CsrView *Q, int use_sum, double *out);
// cone_cnt (per cone, may be null): cone count of the cone's instance
void launch_block_scale(hipStream_t st, int32_t ncones, const int32_t *idx, const int32_t *cone_cnt, double *t,
                        int32_t strategy, int32_t phase, int32_t loc);
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
// val[k] *= f(inst) over the entries of each instance's rows (a square operator on columns)
void launch_scale_rows_inst(hipStream_t st, const SegPlan &sp, const int32_t *rowptr, double *val, InstVal f);
// v[i] = v[i] / div(inst) when div(inst) != 1, else v[i] * mul(inst)
void launch_scale_scalar(hipStream_t st, const SegPlan &sp, double *v, InstVal mul, InstVal div);
// per-instance reductions for bc scaling/norms: finite-sumsq, finite-count, finite-absmax of v
void launch_finite_stats(hipStream_t st, const SegPlan &sp, const double *v, double *partials, double *out,
                         int ostride);
// max(|AL_nInf|,|AU_nInf|): out = [sumsq, absmax] per instance
void launch_bound_stats(hipStream_t st, const SegPlan &sp, const double *AL, const double *AU, double *partials,
                        double *out, int ostride);
// diag extraction + diagonality test per instance: out = [rows with |rowsum - |diag|| > rtol*max,
// off-diagonal entries, largest entry]
void launch_q_diag(hipStream_t st, const SegPlan &sp, const int32_t *rowptr, const int32_t *col,
                   const double *val, double *diag, double *partials, double *out, int ostride);
// diagonal Q factors with the instance's sigma, unified_update_Q_factors! (unified_operations.jl:468-482)
void launch_q_factors(hipStream_t st, const SegPlan &sp, const double *dq, const InstScal *isc, double *fact2,
                      double *fact, double *fact1, double *fact_m);
// initial z_bar and s (process_initial_points!, algorithm.jl:2656-2676)
void launch_init_zbar(hipStream_t st, const SegPlan &sp, const double *x_bar, const double *qx,
                      const double *aty, const double *c, const InstScal *isc, double *z_bar);
void launch_init_s(hipStream_t st, const SegPlan &sp, const double *ax, const double *y, const double *AL,
                   const double *AU, const InstScal *isc, double *s);
// unscale for collect_results! (algorithm.jl:95-102): out = a * (v / d) or a * (v * d), a per instance
void launch_unscale(hipStream_t st, const SegPlan &sp, const double *v, const double *d, InstVal a, int mul,
                    double *out);
// copy src -> dst on the chunks of instances with flag[inst] != 0
void launch_copy_flagged(hipStream_t st, const SegPlan &sp, const double *src, double *dst, const int32_t *flag);

}  // namespace hpr
namespace hpr {
void launch_scale_csr_t(hipStream_t st, int32_t nrows, const int32_t *rowptr, const int32_t *col,
                        double *val, const double *r, const double *c);
void launch_sym_q(hipStream_t st, int64_t nnz, const double *src, const int32_t *tpos, double *dst);
// dst[k] = src[perm[k]] (fills the slice-major copy of a matrix' values)
void launch_gather_perm(hipStream_t st, int64_t nnz, const double *src, const int32_t *perm, double *dst);
void launch_gather_perm(hipStream_t st, int64_t nnz, const float *src, const int32_t *perm, float *dst);
constexpr int kMaxGolden = 64;
// out = (v .* d) ./ s(inst)   (warm-start scaling, algorithm.jl:2628, 2646)
void launch_scale_vd(hipStream_t st, const SegPlan &sp, const double *v, const double *d, InstVal s, double *out);
}  // namespace hpr
