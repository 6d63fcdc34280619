// kernels.hip — FP64 kernels around the hot path: deterministic reductions,
// residuals / gap / bounds violation, the batched golden-section evaluation,
// power-iteration steps, scaling passes and small elementwise helpers.
// The fused SpMV + projection + Halpern kernels are in hotpath.hip.
#include "devutil.h"
#include "kernels.h"
#include "scale_rules.h"

#include <math.h>

#include <algorithm>

namespace hpr {

// device helpers (nmax, jclamp, wave/block reductions) live in devutil.h

// ---------------------------------------------------------------- reductions
struct RedTerms {
    RedTerm t[kMaxRed];
    int32_t n;
};

__device__ __forceinline__ double red_term(const RedTerm &r, int64_t i) {
    if (r.op == 0) return r.c ? (r.a[i] * r.b[i]) * r.c[i] : r.a[i] * r.b[i];
    if (r.op == 1) return fabs(r.a[i]);
    if (r.op == 3) return r.a[i];
    const double d = r.a[i] - r.b[i];
    return d * d;
}

#define CHUNK_LOOP(i, ch) for (int64_t i = (ch).start + (int64_t)threadIdx.x; i < (ch).end; i += kBlock)

__global__ __launch_bounds__(kBlock) void k_reduce1(SegPlan sp, RedTerms T, double *partials) {
    __shared__ double sh[4];
    const SegChunk ch = sp.chunks[blockIdx.x];
    double acc[kMaxRed];
#pragma unroll
    for (int k = 0; k < kMaxRed; ++k) acc[k] = (k < T.n && T.t[k].op == 3) ? -INFINITY : 0.0;
    CHUNK_LOOP(i, ch) {
#pragma unroll
        for (int k = 0; k < kMaxRed; ++k) {
            if (k < T.n) {
                const double v = red_term(T.t[k], i);
                acc[k] = (T.t[k].op == 1 || T.t[k].op == 3) ? nmax(acc[k], v) : acc[k] + v;
            }
        }
    }
    for (int k = 0; k < T.n; ++k) {
        const bool mx = T.t[k].op == 1 || T.t[k].op == 3;
        const double r = mx ? block_max(acc[k], sh) : block_sum(acc[k], sh);
        if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * kMaxRed + k] = r;
    }
}

// Second level: workgroup i reduces the partials of instance i's chunks in
// chunk order.  maxmask bit k: slot k is a max; bit 16+k: signed max (-inf identity).
__global__ __launch_bounds__(kBlock) void k_reduce2(SegPlan sp, int nterms, uint32_t maxmask, int pstride,
                                                    const double *partials, double *out, int ostride) {
    __shared__ double sh[4];
    const int inst = blockIdx.x;
    const int c0 = sp.first[inst], c1 = sp.first[inst + 1];
    for (int k = 0; k < nterms; ++k) {
        const bool mx = (maxmask >> k) & 1u;
        const bool neg = (maxmask >> (16 + k)) & 1u;
        double acc = neg ? -INFINITY : 0.0;
        for (int b = c0 + threadIdx.x; b < c1; b += kBlock) {
            const double v = partials[(int64_t)b * pstride + k];
            acc = mx ? nmax(acc, v) : acc + v;
        }
        const double r = mx ? block_max(acc, sh) : block_sum(acc, sh);
        if (threadIdx.x == 0) out[(int64_t)inst * ostride + k] = r;
    }
}

static void finish_reduce(hipStream_t st, const SegPlan &sp, int nterms, uint32_t maxmask, int pstride,
                          const double *partials, double *out, int ostride) {
    hipLaunchKernelGGL(k_reduce2, dim3(sp.ninst), dim3(kBlock), 0, st, sp, nterms, maxmask, pstride, partials, out,
                       ostride);
}

void launch_reduce(hipStream_t st, const SegPlan &sp, const RedTerm *terms, int nterms, double *partials,
                   double *out, int ostride) {
    RedTerms T{};
    uint32_t mask = 0;
    for (int k = 0; k < nterms; ++k) {
        T.t[k] = terms[k];
        if (terms[k].op == 1 || terms[k].op == 3) mask |= 1u << k;
        if (terms[k].op == 3) mask |= 1u << (16 + k);
    }
    T.n = nterms;
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_reduce1, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, T, partials);
    finish_reduce(st, sp, nterms, mask, kMaxRed, partials, out, ostride);
}

__device__ __forceinline__ double ival(InstVal v, int inst) { return v.p[(int64_t)inst * v.stride]; }

// ---------------------------------------------------------------- elementwise
#define GRID_STRIDE(i, n) \
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kBlock)

// One element per thread: each workgroup streams one contiguous block.  A
// 1 GiB copy runs at 6.2 TB/s this way on MI355X, 4.2-5.8 TB/s with a grid
// of 8-64 workgroups per CU striding over the array.
static inline int ew_blocks(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > (1 << 30)) b = 1 << 30;
    return (int)b;
}

__global__ void k_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out) {
    GRID_STRIDE(i, n) out[i] = a * x[i] + b * y[i];
}
__global__ void k_copy(int64_t n, const double *s, double *d) { GRID_STRIDE(i, n) d[i] = s[i]; }
__global__ void k_fill(int64_t n, double v, double *d) { GRID_STRIDE(i, n) d[i] = v; }
__global__ void k_mul(int64_t n, const double *a, const double *b, double *o) { GRID_STRIDE(i, n) o[i] = a[i] * b[i]; }

__global__ void k_demote(int64_t n, const double *s, float *d) { GRID_STRIDE(i, n) d[i] = (float)s[i]; }
__global__ void k_promote(int64_t n, const float *s, double *d) { GRID_STRIDE(i, n) d[i] = (double)s[i]; }

void launch_demote(hipStream_t st, int64_t n, const double *src, float *dst) {
    if (n > 0) hipLaunchKernelGGL(k_demote, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, src, dst);
}
void launch_promote(hipStream_t st, int64_t n, const float *src, double *dst) {
    if (n > 0) hipLaunchKernelGGL(k_promote, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, src, dst);
}
void launch_axpby(hipStream_t st, int64_t n, double a, const double *x, double b, const double *y, double *out) {
    if (n > 0) hipLaunchKernelGGL(k_axpby, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, a, x, b, y, out);
}
void launch_copy(hipStream_t st, int64_t n, const double *src, double *dst) {
    if (n > 0) hipLaunchKernelGGL(k_copy, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, src, dst);
}
void launch_fill(hipStream_t st, int64_t n, double v, double *dst) {
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, v, dst);
}
void launch_mul(hipStream_t st, int64_t n, const double *a, const double *b, double *out) {
    if (n > 0) hipLaunchKernelGGL(k_mul, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, a, b, out);
}

// ---------------------------------------------------------------- residuals
__global__ __launch_bounds__(kBlock) void k_resid_n(SegPlan sp, ResidNArgs a, double *partials) {
    __shared__ double sh[4];
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double cs = a.isc[ch.inst].c_scale;
    double s0 = 0, s1 = 0, s2 = 0, m3 = 0, m4 = 0, m5 = 0;
    CHUNK_LOOP(j, ch) {
        const double cn = a.col_norm[j];
        const double xb = a.x_bar[j], zb = a.z_bar[j];
        const double c_org = a.c[j] * cn * cs, at = a.atdy[j] * cn * cs, z = zb * cn * cs;
        double rd;
        s0 += a.c[j] * xb;
        s2 += zb * xb;
        if (a.qx) {
            const double qs = a.qx[j];
            const double qo = qs * cn * cs;
            s1 += xb * qs;
            rd = qo + c_org - at - z;
            a.qx[j] = qo;
            m5 = nmax(m5, fabs(qo));
        } else {
            rd = c_org - at - z;
        }
        a.rd[j] = rd;
        a.atdy[j] = at;
        m3 = nmax(m3, fabs(rd));
        m4 = nmax(m4, fabs(at));
    }
    const double vals[6] = {block_sum(s0, sh), block_sum(s1, sh), block_sum(s2, sh),
                            block_max(m3, sh), block_max(m4, sh), block_max(m5, sh)};
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) partials[(int64_t)blockIdx.x * kMaxRed + k] = vals[k];
}

void launch_resid_n(hipStream_t st, const SegPlan &sp, const ResidNArgs &a, double *partials, double *out,
                    int ostride) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_resid_n, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, a, partials);
    finish_reduce(st, sp, 6, 0x38u, kMaxRed, partials, out, ostride);
}

__global__ void k_ax_to_org(SegPlan sp, double *ax, const double *rn, const InstScal *isc) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double bs = isc[ch.inst].b_scale;
    CHUNK_LOOP(i, ch) ax[i] = ax[i] * rn[i] * bs;
}
void launch_ax_to_org(hipStream_t st, const SegPlan &sp, double *ax, const double *row_norm, const InstScal *isc) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_ax_to_org, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, ax, row_norm, isc);
}

// t and ||xi|| of (A x)_org - rhs_org per cone, one wavefront per cone piece;
// a cone of one piece is finished here, the others by k_cone_rp_fin.
__global__ __launch_bounds__(kBlock) void k_cone_rp_part(const int32_t *idx, const int32_t *cinst, ConePieces pc,
                                                         const double *ax, const double *rhs, const double *rn,
                                                         const InstScal *isc, double *ct, double *cns) {
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t p = wid; p < pc.npieces; p += nw) {
        const int c = pc.cone[p];
        const int s = idx[c], e = idx[c + 1];
        const int b = pc.row[p], pe = min(b + kConePiece, e);
        const double bs = isc[cinst ? cinst[c] : 0].b_scale;
        double sq = 0.0;
        for (int j = max(b, s + 1) + lane; j < pe; j += 64) {
            const double v = ax[j] - rhs[j] * rn[j] * bs;
            sq += v * v;
        }
        sq = wave_sum(sq);
        if (lane == 0) {
            if (pc.first[c + 1] - pc.first[c] == 1) {
                cns[c] = sqrt(sq);
                ct[c] = s < e ? ax[s] - rhs[s] * rn[s] * bs : 0.0;
            } else {
                pc.part[p] = sq;
            }
        }
    }
}
// the cones of several pieces: their partial sums in piece order
__global__ __launch_bounds__(kBlock) void k_cone_rp_fin(const int32_t *idx, const int32_t *cinst, ConePieces pc,
                                                        const double *ax, const double *rhs, const double *rn,
                                                        const InstScal *isc, double *ct, double *cns) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= pc.nmulti) return;
    const int c = pc.multi[k];
    const int s = idx[c];
    const double bs = isc[cinst ? cinst[c] : 0].b_scale;
    double sq = 0.0;
    for (int p = pc.first[c]; p < pc.first[c + 1]; ++p) sq += pc.part[p];
    cns[c] = sqrt(sq);
    ct[c] = ax[s] - rhs[s] * rn[s] * bs;
}
void launch_cone_rp_stats(hipStream_t st, int32_t ncones, const int32_t *con_idx, const int32_t *cone_inst,
                          const ConePieces &pc, const double *ax_org, const double *rhs, const double *row_norm,
                          const InstScal *isc, double *cone_t, double *cone_ns) {
    if (ncones <= 0 || pc.npieces <= 0) return;
    int blocks = (pc.npieces + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_cone_rp_part, dim3(blocks), dim3(kBlock), 0, st, con_idx, cone_inst, pc, ax_org, rhs,
                       row_norm, isc, cone_t, cone_ns);
    if (pc.nmulti > 0)
        hipLaunchKernelGGL(k_cone_rp_fin, dim3((pc.nmulti + kBlock - 1) / kBlock), dim3(kBlock), 0, st, con_idx,
                           cone_inst, pc, ax_org, rhs, row_norm, isc, cone_t, cone_ns);
}

__global__ __launch_bounds__(kBlock) void k_resid_m(SegPlan sp, ResidMArgs a, double *partials) {
    __shared__ double sh[4];
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double bs = a.isc[ch.inst].b_scale;
    double s0 = 0, s1 = 0, m2 = 0, m3 = 0, m4 = 0;
    CHUNK_LOOP(i, ch) {
        const double ax = a.ax[i];
        const double rn = a.row_norm[i];
        const double yb = a.y_bar[i];
        s0 += yb * a.s[i];
        s1 += yb * a.rhs[i];
        double rp;
        const int c = a.row_cone[i];
        if (c < 0) {
            const double al = a.AL[i] * rn * bs, au = a.AU[i] * rn * bs;
            rp = jclamp(ax, al, au) - ax;
        } else {
            const int s = a.con_idx[c];
            const double t = a.cone_t[c], ns = a.cone_ns[c];
            const double v = (i == s) ? t : ax - a.rhs[i] * rn * bs;
            if (ns <= -t) rp = v;
            else if (ns <= t) rp = 0.0;
            else if (i == s) rp = (ns + t) / 2 - t;
            else rp = ((1 + t / ns) / 2) * v - v;
        }
        a.rp[i] = rp;
        if (i < a.m1) m2 = nmax(m2, fabs(rp));
        else m3 = nmax(m3, fabs(rp));
        m4 = nmax(m4, fabs(ax));
    }
    const double vals[5] = {block_sum(s0, sh), block_sum(s1, sh), block_max(m2, sh), block_max(m3, sh),
                            block_max(m4, sh)};
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; ++k) partials[(int64_t)blockIdx.x * kMaxRed + k] = vals[k];
}

void launch_resid_m(hipStream_t st, const SegPlan &sp, const ResidMArgs &a, double *partials, double *out,
                    int ostride) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_resid_m, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, a, partials);
    finish_reduce(st, sp, 5, 0x1Cu, kMaxRed, partials, out, ostride);
}

// compute_bounds_violation!: box part over all coordinates, cone part per cone
__global__ void k_bv_box(SegPlan sp, const double *xb, const double *l, const double *u, const double *cn,
                         const InstScal *isc, double *dx) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double bs = isc[ch.inst].b_scale;
    CHUNK_LOOP(j, ch) {
        const double lo = nmax(l[j] - xb[j], 0.0), up = nmax(xb[j] - u[j], 0.0);
        dx[j] = (lo + up) * (bs / cn[j]);
    }
}
__global__ void k_bv_cone(int32_t nvar, const int32_t *idx, const int32_t *cinst, const double *xb,
                          const double *cn, const InstScal *isc, double *dx) {
    GRID_STRIDE(k, nvar) {
        const int s = idx[k], e = idx[k + 1];
        if (s >= e) continue;
        const double bs = isc[cinst ? cinst[k] : 0].b_scale;
        const double t = xb[s];
        double ns = 0.0;
        dx[s] = t;
        for (int j = s + 1; j < e; ++j) {
            dx[j] = xb[j];
            ns += xb[j] * xb[j];
        }
        ns = sqrt(ns);
        if (ns <= -t) {
        } else if (ns <= t) {
            for (int j = s; j < e; ++j) dx[j] = 0.0;
        } else {
            const double f = (1 + t / ns) / 2;
            dx[s] = (ns + t) / 2 - t;
            for (int j = s + 1; j < e; ++j) dx[j] = f * dx[j] - dx[j];
        }
        for (int j = s; j < e; ++j) dx[j] *= bs / cn[j];
    }
}
void launch_bounds_violation(hipStream_t st, const SegPlan &sp, int32_t nvar, const int32_t *var_idx,
                             const int32_t *cone_inst, const double *x_bar, const double *l, const double *u,
                             const double *col_norm, const InstScal *isc, double *dx, double *partials,
                             double *out, int ostride) {
    if (sp.nchunks > 0)
        hipLaunchKernelGGL(k_bv_box, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, x_bar, l, u, col_norm, isc, dx);
    if (nvar > 0)
        hipLaunchKernelGGL(k_bv_cone, dim3(ew_blocks(nvar)), dim3(kBlock), 0, st, nvar, var_idx, cone_inst, x_bar,
                           col_norm, isc, dx);
    RedTerm t{dx, nullptr, nullptr, 1};
    launch_reduce(st, sp, &t, 1, partials, out, ostride);
}

// ------------------------------------------------------------ golden section
__global__ __launch_bounds__(kBlock) void k_golden_eval(SegPlan sp, const double *atdy, const double *qatdy,
                                                        const double *dq, const double *xs, int npts_all,
                                                        const int32_t *np_inst, double *partials) {
    __shared__ double sh[4];
    __shared__ double xsh[kMaxGolden];
    const SegChunk ch = sp.chunks[blockIdx.x];
    const int npts = np_inst ? min(npts_all, (int)np_inst[ch.inst]) : npts_all;
    if (npts <= 0) return;  // no search on this instance (its slots are not read)
    if ((int)threadIdx.x < npts) xsh[threadIdx.x] = xs[(int64_t)ch.inst * kMaxGolden + threadIdx.x];
    __syncthreads();
    for (int k0 = 0; k0 < npts; k0 += kMaxRed) {
        double acc[kMaxRed];
#pragma unroll
        for (int k = 0; k < kMaxRed; ++k) acc[k] = 0.0;
        CHUNK_LOOP(i, ch) {
            const double c = atdy[i], d = qatdy[i], q = dq[i];
#pragma unroll
            for (int k = 0; k < kMaxRed; ++k)
                if (k0 + k < npts) acc[k] += c * (d / (1.0 + xsh[k0 + k] * q));
        }
        for (int k = 0; k < kMaxRed && k0 + k < npts; ++k) {
            const double r = block_sum(acc[k], sh);
            if (threadIdx.x == 0) partials[((int64_t)blockIdx.x * kMaxGolden) + k0 + k] = r;
        }
    }
}
void launch_golden_eval(hipStream_t st, const SegPlan &sp, const double *atdy, const double *qatdy,
                        const double *dq, const double *xs, int npts, const int32_t *np_inst, double *partials,
                        double *out, int ostride) {
    if (sp.nchunks > 0)
        hipLaunchKernelGGL(k_golden_eval, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, atdy, qatdy, dq, xs, npts,
                           np_inst, partials);
    finish_reduce(st, sp, npts, 0u, kMaxGolden, partials, out, ostride);
}

// ------------------------------------------------------------ power iteration
__global__ void k_scale_by_rnorm(SegPlan sp, const double *z, InstVal nrm2, double *q, const PowerState *ps) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    if (ps[ch.inst].done) return;
    const double nz = sqrt(ival(nrm2, ch.inst));
    CHUNK_LOOP(i, ch) q[i] = z[i] / nz;
}
__global__ void k_power_residual(SegPlan sp, const double *z, InstVal lam, double *q, const PowerState *ps) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    if (ps[ch.inst].done) return;
    const double l = ival(lam, ch.inst);
    CHUNK_LOOP(i, ch) q[i] = z[i] - l * q[i];
}
__global__ void k_power_check(int32_t ninst, InstVal lam, InstVal qq, InstVal zz, PowerState *ps, double tol,
                              int32_t max_iter) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ninst || ps[i].done) return;
    const double l = ival(lam, i);
    const double err = sqrt(ival(qq, i)) / (sqrt(ival(zz, i)) + l);
    ps[i].lambda = l;
    ps[i].err = err;
    ps[i].iters += 1;
    if (err < tol || ps[i].iters >= max_iter) ps[i].done = 1;
}
void launch_scale_by_rnorm(hipStream_t st, const SegPlan &sp, const double *z, InstVal nrm2, double *q,
                           const PowerState *ps) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_scale_by_rnorm, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, z, nrm2, q, ps);
}
void launch_power_residual(hipStream_t st, const SegPlan &sp, const double *z, InstVal lam, double *q,
                           const PowerState *ps) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_power_residual, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, z, lam, q, ps);
}
void launch_power_check(hipStream_t st, int32_t ninst, InstVal lam, InstVal qq, InstVal zz, PowerState *ps,
                        double tol, int32_t max_iter) {
    hipLaunchKernelGGL(k_power_check, dim3((ninst + kBlock - 1) / kBlock), dim3(kBlock), 0, st, ninst, lam, qq, zz,
                       ps, tol, max_iter);
}

// ------------------------------------------------------------ scaling
// one wavefront per row (rows of a batch can hold tens of thousands of entries)
__global__ __launch_bounds__(kBlock) void k_row_stat(CsrView M, CsrView Q, int has_q, int use_sum, double *out) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6; i < M.nrows; i += nw) {
        // four independent accumulators per lane keep four loads in flight on long rows
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, b = 0.0;
        const int e = M.rowptr[i + 1];
        int k = M.rowptr[i] + lane;
        for (; k + 192 < e; k += 256) {
            const double v0 = fabs(M.val[k]), v1 = fabs(M.val[k + 64]), v2 = fabs(M.val[k + 128]),
                         v3 = fabs(M.val[k + 192]);
            a0 = use_sum ? a0 + v0 : nmax(a0, v0);
            a1 = use_sum ? a1 + v1 : nmax(a1, v1);
            a2 = use_sum ? a2 + v2 : nmax(a2, v2);
            a3 = use_sum ? a3 + v3 : nmax(a3, v3);
        }
        for (; k < e; k += 64) {
            const double v = fabs(M.val[k]);
            a0 = use_sum ? a0 + v : nmax(a0, v);
        }
        double a = use_sum ? (a0 + a1) + (a2 + a3) : nmax(nmax(a0, a1), nmax(a2, a3));
        if (has_q)
            for (int k = Q.rowptr[i] + lane; k < Q.rowptr[i + 1]; k += 64) {
                const double v = fabs(Q.val[k]);
                b = use_sum ? b + v : nmax(b, v);
            }
        a = use_sum ? wave_sum(a) : wave_max(a);
        if (has_q) b = use_sum ? wave_sum(b) : wave_max(b);
        if (lane == 0) {
            double r = sqrt(has_q ? (use_sum ? a + b : nmax(a, b)) : a);
            if (r == 0.0) r = 1.0;
            out[i] = r;
        }
    }
}
static inline int wave_row_blocks(int64_t rows) {
    int64_t b = (rows + 3) / 4;
    if (b < 1) b = 1;
    if (b > 16384) b = 16384;
    return (int)b;
}
void launch_row_stat(hipStream_t st, const CsrView &M, const CsrView *Q, int use_sum, double *out) {
    if (M.nrows <= 0) return;
    CsrView q = Q ? *Q : CsrView{nullptr, nullptr, nullptr, 0};
    hipLaunchKernelGGL(k_row_stat, dim3(wave_row_blocks(M.nrows)), dim3(kBlock), 0, st, M, q, Q ? 1 : 0, use_sum,
                       out);
}

// One workgroup per cone: block max and sum of squares of the cone's per-row
// norms (lane-strided, then wave butterflies, then the waves in fixed order),
// the block rule, and the shared factor written back to every row.  Cones of
// 10^5 rows and more (one per robust-regression instance) would otherwise be
// reduced by a single thread.
__global__ __launch_bounds__(kBlock) void k_block_scale(int32_t ncones, const int32_t *idx, const int32_t *cnt,
                                                      double *t, int32_t strategy, int32_t phase, int32_t loc) {
    __shared__ double sh_m[kBlock / 64], sh_s[kBlock / 64];
    __shared__ double sv;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int64_t c = blockIdx.x; c < ncones; c += gridDim.x) {
        const int s = idx[c], e = idx[c + 1];
        if (s >= e) continue;  // uniform over the workgroup
        double bm = t[s], ss = 0.0;
        for (int j = s + (int)threadIdx.x; j < e; j += kBlock) {
            bm = nmax(bm, t[j]);
            ss += t[j] * t[j];
        }
        bm = wave_max(bm);
        ss = wave_sum(ss);
        if (lane == 0) {
            sh_m[wid] = bm;
            sh_s[wid] = ss;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = sh_m[0], q = sh_s[0];
            for (int w = 1; w < kBlock / 64; ++w) {
                m = nmax(m, sh_m[w]);
                q += sh_s[w];
            }
            const double rms = sqrt(q / (double)(e - s));
            sv = soc_block_value(e - s, m, rms, strategy, phase, loc, cnt ? cnt[c] : ncones);
        }
        __syncthreads();
        const double v = sv;
        for (int j = s + (int)threadIdx.x; j < e; j += kBlock) t[j] = v;
        __syncthreads();  // sh_m, sh_s and sv are reused by the next cone
    }
}
void launch_block_scale(hipStream_t st, int32_t ncones, const int32_t *idx, const int32_t *cone_cnt, double *t,
                        int32_t strategy, int32_t phase, int32_t loc) {
    if (ncones > 0)
        hipLaunchKernelGGL(k_block_scale, dim3((unsigned)std::min<int64_t>(ncones, 1 << 18)), dim3(kBlock), 0, st,
                           ncones, idx, cone_cnt, t, strategy, phase, loc);
}
__global__ void k_clip(int64_t n, double *v) {
    GRID_STRIDE(i, n) { const double x = v[i]; v[i] = x > 1e4 ? 1e4 : (x < 1e-4 ? 1e-4 : x); }
}
void launch_clip(hipStream_t st, int64_t n, double *v) {
    if (n > 0) hipLaunchKernelGGL(k_clip, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, v);
}
// val = ((1/r[i]) * val) * (1/c[j]) with (i, j) the entry's (row, col) of the
// operator being scaled, or its transpose when `trans` (keeps A and A^T bit-equal)
__global__ __launch_bounds__(kBlock) void k_scale_csr(int32_t nrows, const int32_t *rp, const int32_t *col,
                                                     double *val, const double *r, const double *c, int trans) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6; i < nrows; i += nw) {
        auto f = [&](int k, int j) {
            return trans ? ((1.0 / r[j]) * val[k]) * (1.0 / c[i]) : ((1.0 / r[i]) * val[k]) * (1.0 / c[j]);
        };
        const int e = rp[i + 1];
        int k = rp[i] + lane;
        for (; k + 192 < e; k += 256) {  // four entries in flight per lane
            const int j0 = col[k], j1 = col[k + 64], j2 = col[k + 128], j3 = col[k + 192];
            const double w0 = f(k, j0), w1 = f(k + 64, j1), w2 = f(k + 128, j2), w3 = f(k + 192, j3);
            val[k] = w0;
            val[k + 64] = w1;
            val[k + 128] = w2;
            val[k + 192] = w3;
        }
        for (; k < e; k += 64) val[k] = f(k, col[k]);
    }
}
void launch_scale_csr(hipStream_t st, int32_t nrows, const int32_t *rowptr, const int32_t *col,
                      double *val, const double *r, const double *c) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_scale_csr, dim3(wave_row_blocks(nrows)), dim3(kBlock), 0, st, nrows, rowptr, col, val, r,
                           c, 0);
}
void launch_scale_csr_t(hipStream_t st, int32_t nrows, const int32_t *rowptr, const int32_t *col,
                        double *val, const double *r, const double *c) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_scale_csr, dim3(wave_row_blocks(nrows)), dim3(kBlock), 0, st, nrows, rowptr, col, val, r,
                           c, 1);
}
__global__ void k_scale_cols(int32_t n, const double *t, double *c, double *l, double *u, double *cn) {
    GRID_STRIDE(j, n) {
        c[j] /= t[j];
        l[j] *= t[j];
        u[j] *= t[j];
        cn[j] *= t[j];
    }
}
void launch_scale_cols(hipStream_t st, int32_t n, const double *tcol, double *c, double *l, double *u,
                       double *col_norm) {
    if (n > 0) hipLaunchKernelGGL(k_scale_cols, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, tcol, c, l, u, col_norm);
}
__global__ void k_scale_rows(int32_t m, const double *t, double *AL, double *AU, double *rhs, double *rn) {
    GRID_STRIDE(i, m) {
        AL[i] /= t[i];
        AU[i] /= t[i];
        rhs[i] /= t[i];
        rn[i] *= t[i];
    }
}
void launch_scale_rows(hipStream_t st, int32_t m, const double *trow, double *AL, double *AU, double *rhs,
                       double *row_norm) {
    if (m > 0) hipLaunchKernelGGL(k_scale_rows, dim3(ew_blocks(m)), dim3(kBlock), 0, st, m, trow, AL, AU, rhs, row_norm);
}
__global__ void k_scale_scalar(SegPlan sp, double *v, InstVal mul, InstVal div) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double mu = ival(mul, ch.inst), dv = ival(div, ch.inst);
    CHUNK_LOOP(i, ch) v[i] = dv != 1.0 ? v[i] / dv : v[i] * mu;
}
void launch_scale_scalar(hipStream_t st, const SegPlan &sp, double *v, InstVal mul, InstVal div) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_scale_scalar, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, v, mul, div);
}
__global__ __launch_bounds__(kBlock) void k_finite_stats(SegPlan sp, const double *v, double *partials) {
    __shared__ double sh[4];
    const SegChunk ch = sp.chunks[blockIdx.x];
    double ss = 0, cnt = 0, mx = 0;
    CHUNK_LOOP(i, ch) {
        const double x = v[i];
        if (isfinite(x)) {
            ss += x * x;
            cnt += 1.0;
            mx = nmax(mx, fabs(x));
        }
    }
    const double a = block_sum(ss, sh), b = block_sum(cnt, sh), c = block_max(mx, sh);
    if (threadIdx.x == 0) {
        partials[(int64_t)blockIdx.x * kMaxRed + 0] = a;
        partials[(int64_t)blockIdx.x * kMaxRed + 1] = b;
        partials[(int64_t)blockIdx.x * kMaxRed + 2] = c;
    }
}
void launch_finite_stats(hipStream_t st, const SegPlan &sp, const double *v, double *partials, double *out,
                         int ostride) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_finite_stats, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, v, partials);
    finish_reduce(st, sp, 3, 0x4u, kMaxRed, partials, out, ostride);
}
__global__ __launch_bounds__(kBlock) void k_bound_stats(SegPlan sp, const double *AL, const double *AU,
                                                        double *partials) {
    __shared__ double sh[4];
    const SegChunk ch = sp.chunks[blockIdx.x];
    double ss = 0, mx = 0;
    CHUNK_LOOP(i, ch) {
        const double a = AL[i] == -INFINITY ? 0.0 : AL[i];
        const double b = AU[i] == INFINITY ? 0.0 : AU[i];
        const double v = nmax(fabs(a), fabs(b));
        ss += v * v;
        mx = nmax(mx, v);
    }
    const double a = block_sum(ss, sh), c = block_max(mx, sh);
    if (threadIdx.x == 0) {
        partials[(int64_t)blockIdx.x * kMaxRed + 0] = a;
        partials[(int64_t)blockIdx.x * kMaxRed + 1] = c;
    }
}
void launch_bound_stats(hipStream_t st, const SegPlan &sp, const double *AL, const double *AU, double *partials,
                        double *out, int ostride) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_bound_stats, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, AL, AU, partials);
    finish_reduce(st, sp, 2, 0x2u, kMaxRed, partials, out, ostride);
}
// check_Q_diagonal: diag_Q and count of coordinates where colsum|Q| is not ≈ |diag|
// (Q is symmetric here, so row sums stand for column sums); out[1] = off-diagonal entries
__global__ __launch_bounds__(kBlock) void k_q_diag(SegPlan sp, const int32_t *rp, const int32_t *col,
                                                   const double *val, double *diag, double *partials) {
    __shared__ double sh[4];
    const SegChunk ch = sp.chunks[blockIdx.x];
    double bad = 0, off = 0, mx = -INFINITY;
    CHUNK_LOOP(i, ch) {
        double d = 0.0, rs = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            rs += fabs(val[k]);
            mx = nmax(mx, val[k]);
            if (col[k] == i) d += val[k];
            else off += 1.0;
        }
        diag[i] = d;
        const double a = rs, b = fabs(d);
        const double rtol = 1.4901161193847656e-08;
        if (!(a == b || fabs(a - b) <= rtol * fmax(a, b))) bad += 1.0;
    }
    const double a = block_sum(bad, sh), b = block_sum(off, sh), c = block_max(mx, sh);
    if (threadIdx.x == 0) {
        partials[(int64_t)blockIdx.x * kMaxRed + 0] = a;
        partials[(int64_t)blockIdx.x * kMaxRed + 1] = b;
        partials[(int64_t)blockIdx.x * kMaxRed + 2] = c;
    }
}
void launch_q_diag(hipStream_t st, const SegPlan &sp, const int32_t *rowptr, const int32_t *col, const double *val,
                   double *diag, double *partials, double *out, int ostride) {
    if (sp.nchunks > 0)
        hipLaunchKernelGGL(k_q_diag, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, rowptr, col, val, diag, partials);
    finish_reduce(st, sp, 3, 0x4u | (1u << 18), kMaxRed, partials, out, ostride);
}
// v[k] *= f(inst) for every entry k of the rows (columns space) of each instance
__global__ void k_scale_rows_inst(SegPlan sp, const int32_t *rp, double *val, InstVal f) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double fv = ival(f, ch.inst);
    CHUNK_LOOP(i, ch) for (int k = rp[i]; k < rp[i + 1]; ++k) val[k] = val[k] * fv;
}
void launch_scale_rows_inst(hipStream_t st, const SegPlan &sp, const int32_t *rowptr, double *val, InstVal f) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_scale_rows_inst, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, rowptr, val, f);
}
__global__ void k_q_factors(SegPlan sp, const double *dq, const InstScal *isc, double *f2, double *f, double *f1,
                            double *fm) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double s = isc[ch.inst].sigma;
    CHUNK_LOOP(i, ch) {
        const double a = 1.0 / (1.0 + s * dq[i]);
        f2[i] = a;
        f[i] = s * a;
        f1[i] = s * dq[i] * a;
        fm[i] = (s * s) * a;
    }
}
void launch_q_factors(hipStream_t st, const SegPlan &sp, const double *dq, const InstScal *isc, double *fact2,
                      double *fact, double *fact1, double *fact_m) {
    if (sp.nchunks > 0)
        hipLaunchKernelGGL(k_q_factors, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, dq, isc, fact2, fact, fact1, fact_m);
}
__global__ void k_init_zbar(SegPlan sp, const double *xb, const double *qx, const double *aty, const double *c,
                            const InstScal *isc, double *zb) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double sigma = isc[ch.inst].sigma;
    CHUNK_LOOP(j, ch) {
        const double tmp = qx ? (-qx[j] + aty[j]) - c[j] : aty[j] - c[j];
        const double zr = xb[j] + sigma * tmp;
        zb[j] = (xb[j] - zr) / sigma;
    }
}
void launch_init_zbar(hipStream_t st, const SegPlan &sp, const double *x_bar, const double *qx, const double *aty,
                      const double *c, const InstScal *isc, double *z_bar) {
    if (sp.nchunks > 0)
        hipLaunchKernelGGL(k_init_zbar, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, x_bar, qx, aty, c, isc, z_bar);
}
__global__ void k_init_s(SegPlan sp, const double *ax, const double *y, const double *AL, const double *AU,
                         const InstScal *isc, double *s) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double f1 = isc[ch.inst].lambda_A * isc[ch.inst].sigma;
    CHUNK_LOOP(i, ch) s[i] = jclamp(ax[i] - f1 * y[i], AL[i], AU[i]);
}
void launch_init_s(hipStream_t st, const SegPlan &sp, const double *ax, const double *y, const double *AL,
                   const double *AU, const InstScal *isc, double *s) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_init_s, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, ax, y, AL, AU, isc, s);
}
__global__ void k_scale_vd(SegPlan sp, const double *v, const double *d, InstVal s, double *out) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double sv = ival(s, ch.inst);
    CHUNK_LOOP(i, ch) out[i] = (v[i] * d[i]) / sv;
}
void launch_scale_vd(hipStream_t st, const SegPlan &sp, const double *v, const double *d, InstVal s, double *out) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_scale_vd, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, v, d, s, out);
}
__global__ void k_unscale(SegPlan sp, const double *v, const double *d, InstVal a, int mul, double *out) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const double av = ival(a, ch.inst);
    CHUNK_LOOP(i, ch) out[i] = mul ? av * (v[i] * d[i]) : av * (v[i] / d[i]);
}
void launch_unscale(hipStream_t st, const SegPlan &sp, const double *v, const double *d, InstVal a, int mul,
                    double *out) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_unscale, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, v, d, a, mul, out);
}
__global__ void k_promote_prec(SegPlan sp, const float *src, double *dst, const InstScal *isc) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const InstScal s = isc[ch.inst];
    if (!s.active || !s.fp32) return;
    CHUNK_LOOP(i, ch) dst[i] = (double)src[i];
}
void launch_promote_prec(hipStream_t st, const SegPlan &sp, const float *src, double *dst, const InstScal *isc) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_promote_prec, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, src, dst, isc);
}

__global__ void k_delta_split(SegPlan sp, const double *v, const double *a, float *d, const InstScal *isc) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    if (!isc[ch.inst].active) return;
    CHUNK_LOOP(i, ch) d[i] = (float)(v[i] - a[i]);
}
void launch_delta_split(hipStream_t st, const SegPlan &sp, const double *v, const double *anchor, float *d,
                        const InstScal *isc) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_delta_split, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, v, anchor, d, isc);
}
__global__ void k_delta_join(SegPlan sp, const double *a, const float *d, double *v, const InstScal *isc) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    if (!isc[ch.inst].active) return;
    CHUNK_LOOP(i, ch) v[i] = a[i] + (double)d[i];
}
void launch_delta_join(hipStream_t st, const SegPlan &sp, const double *anchor, const float *d, double *v,
                       const InstScal *isc) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_delta_join, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, anchor, d, v, isc);
}
__global__ void k_delta_ra(SegPlan sp, const double *xa, const double *img, const double *qwa, const double *c,
                           const InstScal *isc, double *ra) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const InstScal s = isc[ch.inst];
    if (!s.active) return;
    CHUNK_LOOP(i, ch) ra[i] = xa[i] + s.sigma * ((-(qwa ? qwa[i] : 0.0) + img[i]) - c[i]);
}
void launch_delta_ra(hipStream_t st, const SegPlan &sp, const double *xa, const double *img, const double *qwa,
                     const double *c, const InstScal *isc, double *ra) {
    if (sp.nchunks > 0)
        hipLaunchKernelGGL(k_delta_ra, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, xa, img, qwa, c, isc, ra);
}
__global__ void k_delta_rya(SegPlan sp, int64_t m1, const double *axa, const double *rhs, const double *ya,
                            const InstScal *isc, double *rya) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    const InstScal s = isc[ch.inst];
    if (!s.active) return;
    const double f1 = s.lambda_A * s.sigma;
    CHUNK_LOOP(i, ch) rya[i] = (i >= m1 ? axa[i] - rhs[i] : axa[i]) - f1 * ya[i];
}
void launch_delta_rya(hipStream_t st, const SegPlan &sp, int64_t m1, const double *axa, const double *rhs,
                      const double *ya, const InstScal *isc, double *rya) {
    if (sp.nchunks > 0)
        hipLaunchKernelGGL(k_delta_rya, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, m1, axa, rhs, ya, isc, rya);
}

__global__ void k_copy_flagged(SegPlan sp, const double *src, double *dst, const int32_t *flag) {
    const SegChunk ch = sp.chunks[blockIdx.x];
    if (!flag[ch.inst]) return;
    CHUNK_LOOP(i, ch) dst[i] = src[i];
}
void launch_copy_flagged(hipStream_t st, const SegPlan &sp, const double *src, double *dst, const int32_t *flag) {
    if (sp.nchunks > 0) hipLaunchKernelGGL(k_copy_flagged, dim3(sp.nchunks), dim3(kBlock), 0, st, sp, src, dst, flag);
}

template <typename T>
__global__ void k_gather_perm(int64_t nnz, const T *src, const int32_t *perm, T *dst) {
    GRID_STRIDE(k, nnz) dst[k] = src[perm[k]];
}
void launch_gather_perm(hipStream_t st, int64_t nnz, const double *src, const int32_t *perm, double *dst) {
    if (nnz > 0)
        hipLaunchKernelGGL(k_gather_perm<double>, dim3(ew_blocks(nnz)), dim3(kBlock), 0, st, nnz, src, perm, dst);
}
void launch_gather_perm(hipStream_t st, int64_t nnz, const float *src, const int32_t *perm, float *dst) {
    if (nnz > 0)
        hipLaunchKernelGGL(k_gather_perm<float>, dim3(ew_blocks(nnz)), dim3(kBlock), 0, st, nnz, src, perm, dst);
}

// symmetrize Q values in place through the transpose-partner map: v = (v + v[tpos]) / 2
__global__ void k_sym_q(int64_t nnz, const double *src, const int32_t *tpos, double *dst) {
    GRID_STRIDE(k, nnz) dst[k] = (src[k] + src[tpos[k]]) / 2.0;
}
void launch_sym_q(hipStream_t st, int64_t nnz, const double *src, const int32_t *tpos, double *dst) {
    if (nnz > 0) hipLaunchKernelGGL(k_sym_q, dim3(ew_blocks(nnz)), dim3(kBlock), 0, st, nnz, src, tpos, dst);
}

}  // namespace hpr
