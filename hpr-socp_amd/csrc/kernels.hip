// kernels.hip — gfx950 kernels of the HPR iteration (FP64 iterates).
//
// Hot path: one workgroup (256 threads = 4 wave64) per Tile.  Phase A runs the
// CSR SpMV rows of the tile with `g` lanes per row (shuffle reduction inside
// the lane group, or the whole workgroup for very long rows) into LDS.  Phase B
// is the fused epilogue: box clamp + reflection + Halpern averaging for
// LINEAR tiles; for CONES tiles a lane group of `g2` lanes owns one cone,
// reduces ||xi||^2 with shuffles and applies the SOC projection in registers;
// HUGE tiles write raw values and a partial ||xi||^2 and a second kernel
// finishes the projection once all partials of the cone exist.
#include "kernels.h"
#include "scale_rules.h"

#include <math.h>

namespace hpr {

// NaN-propagating helpers (Julia's max/min/clamp semantics on the CPU path).
__device__ __forceinline__ double nmax(double a, double b) { return (a > b || a != a) ? a : b; }
__device__ __forceinline__ double jclamp(double z, double lo, double hi) {
    return (z != z) ? z : fmin(fmax(z, lo), hi);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o, 64));
    return v;
}
// Deterministic block reductions; result valid in every thread.
__device__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    return r;
}
__device__ double block_max(double v, double *sh) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    return nmax(nmax(sh[0], sh[1]), nmax(sh[2], sh[3]));
}

// ---------------------------------------------------------------- SpMV phase
// CSR-stream: coalesced loads of the tile's nonzeros, independent gathers, one
// product per LDS slot, then a sequential per-row sum (CSR order, products
// rounded before summation).
__device__ void tile_spmv_stream(const CsrView &M, const double *__restrict__ v, const Tile &t,
                                 double *acc, double *prod, int32_t *rp) {
    const int tid = threadIdx.x;
    const int nrows = t.re - t.rb;
    for (int r = tid; r <= nrows; r += kBlock) rp[r] = M.rowptr[t.rb + r];
    __syncthreads();
    const int k0 = rp[0];
    const int nnz = rp[nrows] - k0;
    const int32_t *__restrict__ col = M.col + k0;
    const double *__restrict__ val = M.val + k0;
#pragma unroll 4
    for (int k = tid; k < nnz; k += kBlock) {
        const double p = val[k] * v[col[k]];
        prod[k] = p;
    }
    __syncthreads();
    if (tid < nrows) {
        const int a = rp[tid] - k0, b = rp[tid + 1] - k0;
        double s = 0.0;
        for (int k = a; k < b; ++k) s += prod[k];
        acc[tid] = s;
    }
}

// Vector path: g lanes per row (shuffle reduction), or the whole workgroup for
// very long rows.
__device__ void tile_spmv(const CsrView &M, const double *__restrict__ v, const Tile &t,
                          double *acc, double *sh) {
    const int tid = threadIdx.x;
    const int nrows = t.re - t.rb;
    const int G = t.g;
    if (G <= 64) {
        const int lane = tid & (G - 1), grp = tid / G, ngrp = kBlock / G;
        for (int r = grp; r < nrows; r += ngrp) {
            const int row = t.rb + r;
            const int s = M.rowptr[row], e = M.rowptr[row + 1];
            double sum = 0.0;
            for (int k = s + lane; k < e; k += G) sum += M.val[k] * v[M.col[k]];
            for (int o = G >> 1; o > 0; o >>= 1) sum += __shfl_xor(sum, o, G);
            if (lane == 0) acc[r] = sum;
        }
    } else {
        for (int r = 0; r < nrows; ++r) {
            const int row = t.rb + r;
            const int s = M.rowptr[row], e = M.rowptr[row + 1];
            double sum = 0.0;
            for (int k = s + tid; k < e; k += kBlock) sum += M.val[k] * v[M.col[k]];
            sum = block_sum(sum, sh);
            if (tid == 0) acc[r] = sum;
        }
    }
}

// Projection scalars of one cone from its head value t and ||xi||.
struct ConeCase {
    int cas;        // 0 polar (-> 0), 1 inside (keep), 2 boundary projection
    double pt, alpha;
};
__device__ __forceinline__ ConeCase cone_case(double t, double ns) {
    ConeCase c;
    c.pt = 0.0;
    c.alpha = 0.0;
    if (ns <= -t) c.cas = 0;
    else if (ns <= t) c.cas = 1;
    else {
        c.cas = 2;
        c.pt = (ns + t) / 2.0;
        c.alpha = c.pt / ns;
    }
    return c;
}
__device__ __forceinline__ double cone_apply(const ConeCase &c, bool head, double raw) {
    return c.cas == 0 ? 0.0 : (c.cas == 1 ? raw : (head ? c.pt : c.alpha * raw));
}

// ------------------------------------------------------------ epilogue ops
struct OpNoQPrimal {
    NoQPrimalArgs a;
    __device__ double raw(int i, double acc, bool) const { return a.x[i] + a.sigma * (acc - a.c[i]); }
    __device__ double lin_proj(int i, double r) const { return jclamp(r, a.l[i], a.u[i]); }
    __device__ void finish(int i, double raw, double p) const {
        const double xi = a.x[i];
        const double xh = 2.0 * p - xi;
        if (a.full) {
            a.x_bar[i] = p;
            a.z_bar[i] = (p - raw) / a.sigma;
            a.dx[i] = p - xi;
        }
        a.x[i] = a.h2 * xh + a.h1 * a.last_x[i];
        a.x_hat[i] = xh;
    }
};

struct OpDual {
    DualArgs a;
    __device__ double raw(int i, double acc, bool cone) const {
        return cone ? (acc - a.rhs[i]) - a.f1 * a.y[i] : acc - a.f1 * a.y[i];
    }
    __device__ double lin_proj(int i, double r) const { return jclamp(r, a.AL[i], a.AU[i]); }
    __device__ void finish(int i, double raw, double p) const {
        const double yi = a.y[i];
        const double yb = a.f2 * (p - raw);
        if (a.full) {
            a.s[i] = p;
            a.dy[i] = yb - yi;
        }
        if (a.full || a.write_ybar) a.y_bar[i] = yb;
        a.y[i] = a.h2 * (2.0 * yb - yi) + a.h1 * a.last_y[i];
    }
};

struct OpQPPrimal {
    QPPrimalArgs a;
    __device__ double qw(int i) const { return a.diag ? a.dq[i] * a.w[i] : a.qw[i]; }
    __device__ double raw(int i, double, bool) const {
        return a.x[i] + a.sigma * ((-qw(i) + a.aty[i]) - a.c[i]);
    }
    __device__ double lin_proj(int i, double r) const { return jclamp(r, a.l[i], a.u[i]); }
    __device__ void finish(int i, double raw, double p) const {
        const double xi = a.x[i], wi = a.w[i];
        const double xh = 2.0 * p - xi;
        const double wb = a.vec_fact ? a.fact1[i] * wi + a.fact2[i] * xh : a.f1s * wi + a.f2s * xh;
        if (a.full) {
            a.x_bar[i] = p;
            a.z_bar[i] = (p - raw) / a.sigma;
            a.dx[i] = p - xi;
        }
        a.x[i] = a.h2 * xh + a.h1 * a.last_x[i];
        a.x_hat[i] = xh;
        a.w_bar[i] = wb;
        if (a.diag) a.tempv[i] = xh + a.sigma * (a.dq[i] * wi - a.dq[i] * wb);
    }
};

struct OpW2 {
    W2Args a;
    __device__ double raw(int, double acc, bool) const { return acc; }
    __device__ double lin_proj(int, double r) const { return r; }
    __device__ void finish(int i, double rb, double) const {
        const double wi = a.w[i], atyi = a.aty[i];
        const double f = a.diag ? a.fact[i] : a.fs;
        const double wbn = a.w_bar[i] + f * (rb - atyi);
        a.w[i] = a.h1 * a.last_w[i] + a.h2 * (2.0 * wbn - wi);
        a.aty[i] = a.h1 * a.last_aty[i] + a.h2 * (2.0 * rb - atyi);
        if (a.full) {
            a.w_bar[i] = wbn;
            a.dw[i] = wbn - wi;
            a.atdy[i] = rb - atyi;
            a.aty_bar[i] = rb;
        }
    }
};

struct OpStore {
    double *out;
    const double *diag, *v2;
    __device__ double raw(int, double acc, bool) const { return acc; }
    __device__ double lin_proj(int, double r) const { return r; }
    __device__ void finish(int i, double r, double) const { out[i] = diag ? r + diag[i] * v2[i] : r; }
};

struct OpTempv {
    TempvArgs a;
    __device__ double raw(int, double acc, bool) const { return acc; }
    __device__ double lin_proj(int, double r) const { return r; }
    __device__ void finish(int i, double r, double) const {
        const double qwb = a.dq ? r + a.dq[i] * a.wbar[i] : r;
        a.tempv[i] = a.x_hat[i] + a.sigma * (a.qw[i] - qwb);
    }
};

// ------------------------------------------------------------ XCD-sliced SpMV
// Block b: slice p = b % P (the XCD it is dispatched to), row tile b / P.  The
// slice's nonzeros of the tile are streamed (non-temporal: read once) in chunks
// of kStreamNnz products through LDS; every row sums its products in CSR order.
__global__ __launch_bounds__(kBlock) void k_spmv_slices(CsrView M, const double *__restrict__ x) {
    __shared__ double prod[kStreamNnz];
    __shared__ int32_t rps[kSliceRows + 1];
    const int P = M.nslices;
    const int p = blockIdx.x % P;
    const SliceTile T = M.s_tiles[blockIdx.x / P];
    const int tid = threadIdx.x;
    const int nr = T.re - T.rb;
    const int32_t *rp = M.s_rowptr + (int64_t)p * (M.nrows + 1);
    for (int r = tid; r <= nr; r += kBlock) rps[r] = rp[T.rb + r];
    __syncthreads();
    const int k0 = rps[0], k1 = rps[nr];
    constexpr int RPT = kSliceRows / kBlock;  // rows per thread
    double acc[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) acc[j] = 0.0;
    for (int c0 = k0; c0 < k1; c0 += kStreamNnz) {
        const int c1 = min(c0 + kStreamNnz, k1);
#pragma unroll 4
        for (int k = c0 + tid; k < c1; k += kBlock) {
            const double v = __builtin_nontemporal_load(M.s_val + k);
            const int c = __builtin_nontemporal_load(M.s_col + k);
            prod[k - c0] = v * x[c];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int r = tid + j * kBlock;
            if (r < nr) {
                const int a = max(rps[r], c0), e = min(rps[r + 1], c1);
                double s = acc[j];
                for (int k = a; k < e; ++k) s += prod[k - c0];
                acc[j] = s;
            }
        }
        __syncthreads();
    }
    double *out = M.s_part + (int64_t)p * M.nrows + T.rb;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = tid + j * kBlock;
        if (r < nr) out[r] = acc[j];
    }
}

static void launch_slices(hipStream_t st, const CsrView &M, const double *x) {
    if (M.s_ntiles > 0)
        hipLaunchKernelGGL(k_spmv_slices, dim3(M.s_ntiles * M.nslices), dim3(kBlock), 0, st, M, x);
}

__global__ void k_gather_perm(int64_t nnz, const double *src, const int32_t *perm, double *dst);

// ------------------------------------------------------------ tile kernel
// SRC: 0 = no SpMV, 1 = CSR SpMV inside the tile, 2 = sum of the XCD-slice partials
enum : int { SRC_NONE = 0, SRC_CSR = 1, SRC_SLICES = 2 };

template <class Op, int SRC, bool CONES>
__global__ __launch_bounds__(kBlock) void k_tile(const Tile *__restrict__ tiles, CsrView M,
                                                 const double *__restrict__ vin,
                                                 const int32_t *__restrict__ cone_idx, HugeInfo huge,
                                                 Op op, double *__restrict__ raw_buf,
                                                 double *__restrict__ partials) {
    __shared__ double acc[kTileRows];
    __shared__ double sh[4];
    __shared__ double prod[SRC == SRC_CSR ? kStreamNnz : 1];
    __shared__ int32_t rps[SRC == SRC_CSR ? kTileRows + 1 : 1];
    const Tile t = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int nrows = t.re - t.rb;
    if constexpr (SRC == SRC_CSR) {
        if (t.g == 0) tile_spmv_stream(M, vin, t, acc, prod, rps);
        else tile_spmv(M, vin, t, acc, sh);
    } else if constexpr (SRC == SRC_SLICES) {
        if (tid < nrows) {
            const double *pp = M.s_part + t.rb + tid;
            double a = 0.0;
            for (int p = 0; p < M.nslices; ++p) a += pp[(int64_t)p * M.nrows];
            acc[tid] = a;
        }
    } else {
        if (tid < nrows) acc[tid] = 0.0;
    }
    __syncthreads();
    if (!CONES || t.kind == TILE_LINEAR) {
        if (tid < nrows) {
            const int i = t.rb + tid;
            const double r = op.raw(i, acc[tid], false);
            op.finish(i, r, op.lin_proj(i, r));
        }
        return;
    }
    if (t.kind == TILE_CONES) {
        const int G2 = t.g2;
        const int lane = tid & (G2 - 1), grp = tid / G2, ngrp = kBlock / G2;
        for (int c = t.c0 + grp; c < t.c1; c += ngrp) {
            const int s = cone_idx[c], e = cone_idx[c + 1];
            double sq = 0.0, head = 0.0;
            for (int i = s + lane; i < e; i += G2) {
                const double r = op.raw(i, acc[i - t.rb], true);
                acc[i - t.rb] = r;
                if (i == s) head = r;
                else sq += r * r;
            }
            for (int o = G2 >> 1; o > 0; o >>= 1) sq += __shfl_xor(sq, o, G2);
            head = __shfl(head, 0, G2);
            const ConeCase cc = cone_case(head, sqrt(sq));
            for (int i = s + lane; i < e; i += G2) {
                const double r = acc[i - t.rb];
                op.finish(i, r, cone_apply(cc, i == s, r));
            }
        }
        return;
    }
    // HUGE segment, phase 1: raw values and a partial ||xi||^2
    double sq = 0.0;
    if (tid < nrows) {
        const int i = t.rb + tid;
        const double r = op.raw(i, acc[tid], true);
        raw_buf[i] = r;
        if (i != huge.start[t.c0]) sq = r * r;
    }
    sq = block_sum(sq, sh);
    if (tid == 0) partials[t.c1] = sq;
}

// HUGE segment, phase 2: the cone's ||xi|| from all partials, then the projection.
template <class Op>
__global__ __launch_bounds__(kBlock) void k_huge_finish(const Tile *__restrict__ tiles, HugeInfo huge,
                                                        Op op, const double *__restrict__ raw_buf,
                                                        const double *__restrict__ partials) {
    __shared__ double sh[4];
    const Tile t = tiles[blockIdx.x];
    const int h = t.c0;
    const int s = huge.start[h];
    const int slot0 = huge.slot0[h], nslot = huge.nslot[h];
    double sum = 0.0;
    for (int k = threadIdx.x; k < nslot; k += kBlock) sum += partials[slot0 + k];
    sum = block_sum(sum, sh);
    const ConeCase cc = cone_case(raw_buf[s], sqrt(sum));
    const int nrows = t.re - t.rb;
    if ((int)threadIdx.x < nrows) {
        const int i = t.rb + threadIdx.x;
        const double r = raw_buf[i];
        op.finish(i, r, cone_apply(cc, i == s, r));
    }
}

// Fused SpMV + epilogue over a plan.  With XCD slicing the SpMV runs first as
// k_spmv_slices and the tile kernel sums the slice partials.
template <class Op, int SRC, bool CONES>
static void run_tiles(hipStream_t st, const DevPlan &plan, const CsrView &M, const double *v,
                      const int32_t *cone_idx, const Op &op, double *raw_buf, double *partials) {
    if (plan.ntiles == 0) return;
    if (SRC == SRC_CSR && M.nslices > 1) {
        launch_slices(st, M, v);
        hipLaunchKernelGGL((k_tile<Op, SRC_SLICES, CONES>), dim3(plan.ntiles), dim3(kBlock), 0, st, plan.tiles,
                           M, v, cone_idx, plan.huge, op, raw_buf, partials);
    } else {
        hipLaunchKernelGGL((k_tile<Op, SRC, CONES>), dim3(plan.ntiles), dim3(kBlock), 0, st, plan.tiles, M, v,
                           cone_idx, plan.huge, op, raw_buf, partials);
    }
    if (CONES) {
        const int nh = plan.ntiles - plan.first_huge;
        if (nh > 0)
            hipLaunchKernelGGL((k_huge_finish<Op>), dim3(nh), dim3(kBlock), 0, st, plan.tiles + plan.first_huge,
                               plan.huge, op, raw_buf, partials);
    }
}

void launch_noq_primal(hipStream_t st, const DevPlan &plan, const CsrView &AT, const int32_t *var_idx,
                       const NoQPrimalArgs &a, double *raw_buf, double *partials) {
    run_tiles<OpNoQPrimal, SRC_CSR, true>(st, plan, AT, a.y, var_idx, OpNoQPrimal{a}, raw_buf, partials);
}
void launch_dual(hipStream_t st, const DevPlan &plan, const CsrView &A, const int32_t *con_idx,
                 const DualArgs &a, double *raw_buf, double *partials) {
    run_tiles<OpDual, SRC_CSR, true>(st, plan, A, a.v, con_idx, OpDual{a}, raw_buf, partials);
}
void launch_qp_primal(hipStream_t st, const DevPlan &plan, const int32_t *var_idx,
                      const QPPrimalArgs &a, double *raw_buf, double *partials) {
    CsrView none{nullptr, nullptr, nullptr, 0};
    run_tiles<OpQPPrimal, SRC_NONE, true>(st, plan, none, nullptr, var_idx, OpQPPrimal{a}, raw_buf, partials);
}
void launch_w2(hipStream_t st, const DevPlan &plan, const CsrView &AT, const W2Args &a) {
    run_tiles<OpW2, SRC_CSR, false>(st, plan, AT, a.ybar, nullptr, OpW2{a}, nullptr, nullptr);
}
void launch_spmv(hipStream_t st, const DevPlan &plan, const CsrView &M, const double *v, double *out,
                 const double *diag, const double *v2) {
    run_tiles<OpStore, SRC_CSR, false>(st, plan, M, v, nullptr, OpStore{out, diag, v2}, nullptr, nullptr);
}
void launch_spmv_tempv(hipStream_t st, const DevPlan &plan, const CsrView &M, const double *wbar,
                       const TempvArgs &a) {
    run_tiles<OpTempv, SRC_CSR, false>(st, plan, M, wbar, nullptr, OpTempv{a}, nullptr, nullptr);
}

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

__global__ __launch_bounds__(kBlock) void k_reduce1(int64_t n, RedTerms T, double *partials) {
    __shared__ double sh[4];
    double acc[kMaxRed];
#pragma unroll
    for (int k = 0; k < kMaxRed; ++k) acc[k] = (k < T.n && T.t[k].op == 3) ? -INFINITY : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
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

// ops: per slot 0 = sum, 1 = max (packed bits in `maxmask`)
__global__ __launch_bounds__(kBlock) void k_reduce2(int nblocks, int nterms, uint32_t maxmask,
                                                    const double *partials, double *out) {
    __shared__ double sh[4];
    for (int k = 0; k < nterms; ++k) {
        const bool mx = (maxmask >> k) & 1u;
        const bool neg = (maxmask >> (16 + k)) & 1u;  // signed max: -inf identity
        double acc = neg ? -INFINITY : 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) {
            const double v = partials[(int64_t)b * kMaxRed + k];
            acc = mx ? nmax(acc, v) : acc + v;
        }
        const double r = mx ? block_max(acc, sh) : block_sum(acc, sh);
        if (threadIdx.x == 0) out[k] = r;
    }
}

int reduce_blocks(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (int)b;
}

void launch_reduce(hipStream_t st, int64_t n, const RedTerm *terms, int nterms, double *partials,
                   int nblocks, double *out) {
    RedTerms T{};
    uint32_t mask = 0;
    for (int k = 0; k < nterms; ++k) {
        T.t[k] = terms[k];
        if (terms[k].op == 1 || terms[k].op == 3) mask |= 1u << k;
        if (terms[k].op == 3) mask |= 1u << (16 + k);
    }
    T.n = nterms;
    hipLaunchKernelGGL(k_reduce1, dim3(nblocks), dim3(kBlock), 0, st, n, T, partials);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, st, nblocks, nterms, mask, partials, out);
}

// ---------------------------------------------------------------- elementwise
#define GRID_STRIDE(i, n) \
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kBlock)

static inline int ew_blocks(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (int)b;
}

__global__ void k_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out) {
    GRID_STRIDE(i, n) out[i] = a * x[i] + b * y[i];
}
__global__ void k_copy(int64_t n, const double *s, double *d) { GRID_STRIDE(i, n) d[i] = s[i]; }
__global__ void k_fill(int64_t n, double v, double *d) { GRID_STRIDE(i, n) d[i] = v; }
__global__ void k_mul(int64_t n, const double *a, const double *b, double *o) { GRID_STRIDE(i, n) o[i] = a[i] * b[i]; }

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
__global__ __launch_bounds__(kBlock) void k_resid_n(int32_t n, ResidNArgs a, double *partials) {
    __shared__ double sh[4];
    double s0 = 0, s1 = 0, s2 = 0, m3 = 0, m4 = 0, m5 = 0;
    GRID_STRIDE(j, n) {
        const double cn = a.col_norm[j], cs = a.c_scale;
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

void launch_resid_n(hipStream_t st, int32_t n, const ResidNArgs &a, double *partials, int nblocks,
                    double *out) {
    hipLaunchKernelGGL(k_resid_n, dim3(nblocks), dim3(kBlock), 0, st, n, a, partials);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, st, nblocks, 6, 0x38u, partials, out);
}

__global__ void k_ax_to_org(int32_t m, double *ax, const double *rn, double bs) {
    GRID_STRIDE(i, m) ax[i] = ax[i] * rn[i] * bs;
}
void launch_ax_to_org(hipStream_t st, int32_t m, double *ax, const double *row_norm, double b_scale) {
    if (m > 0) hipLaunchKernelGGL(k_ax_to_org, dim3(ew_blocks(m)), dim3(kBlock), 0, st, m, ax, row_norm, b_scale);
}

// one wavefront per cone: t and ||xi|| of (A x)_org - rhs_org
__global__ __launch_bounds__(kBlock) void k_cone_rp_stats(int32_t ncones, const int32_t *idx,
                                                          const double *ax, const double *rhs,
                                                          const double *rn, double bs, double *ct,
                                                          double *cns) {
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * kBlock) >> 6;
    for (int64_t c = wid; c < ncones; c += nw) {
        const int s = idx[c], e = idx[c + 1];
        double sq = 0.0;
        for (int j = s + 1 + lane; j < e; j += 64) {
            const double v = ax[j] - rhs[j] * rn[j] * bs;
            sq += v * v;
        }
        sq = wave_sum(sq);
        if (lane == 0) {
            cns[c] = sqrt(sq);
            ct[c] = s < e ? ax[s] - rhs[s] * rn[s] * bs : 0.0;
        }
    }
}
void launch_cone_rp_stats(hipStream_t st, int32_t ncones, const int32_t *con_idx, const double *ax_org,
                          const double *rhs, const double *row_norm, double b_scale, double *cone_t,
                          double *cone_ns) {
    if (ncones <= 0) return;
    int blocks = (int)((ncones + 3) / 4);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_cone_rp_stats, dim3(blocks), dim3(kBlock), 0, st, ncones, con_idx, ax_org, rhs,
                       row_norm, b_scale, cone_t, cone_ns);
}


__global__ __launch_bounds__(kBlock) void k_resid_m(int32_t m, ResidMArgs a, double *partials) {
    __shared__ double sh[4];
    double s0 = 0, s1 = 0, m2 = 0, m3 = 0, m4 = 0;
    GRID_STRIDE(i, m) {
        const double ax = a.ax[i];
        const double rn = a.row_norm[i], bs = a.b_scale;
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

void launch_resid_m(hipStream_t st, int32_t m, const ResidMArgs &a, double *partials, int nblocks,
                    double *out) {
    hipLaunchKernelGGL(k_resid_m, dim3(nblocks), dim3(kBlock), 0, st, m, a, partials);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, st, nblocks, 5, 0x1Cu, partials, out);
}

// compute_bounds_violation!: box part over all coordinates, cone part per cone
__global__ void k_bv_box(int32_t n, const double *xb, const double *l, const double *u, const double *cn,
                         double bs, double *dx) {
    GRID_STRIDE(j, n) {
        const double lo = nmax(l[j] - xb[j], 0.0), up = nmax(xb[j] - u[j], 0.0);
        dx[j] = (lo + up) * (bs / cn[j]);
    }
}
__global__ void k_bv_cone(int32_t nvar, const int32_t *idx, const double *xb, const double *cn, double bs,
                          double *dx) {
    GRID_STRIDE(k, nvar) {
        const int s = idx[k], e = idx[k + 1];
        if (s >= e) continue;
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
void launch_bounds_violation(hipStream_t st, int32_t n, int32_t nvar, const int32_t *var_idx,
                             const double *x_bar, const double *l, const double *u,
                             const double *col_norm, double b_scale, double *dx, double *partials,
                             int nblocks, double *out) {
    if (n > 0)
        hipLaunchKernelGGL(k_bv_box, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, x_bar, l, u, col_norm, b_scale, dx);
    if (nvar > 0)
        hipLaunchKernelGGL(k_bv_cone, dim3(ew_blocks(nvar)), dim3(kBlock), 0, st, nvar, var_idx, x_bar, col_norm,
                           b_scale, dx);
    RedTerm t{dx, nullptr, nullptr, 1};
    launch_reduce(st, n, &t, 1, partials, nblocks, out);
}

// ------------------------------------------------------------ golden section
constexpr int kMaxGoldenPts = 64;
__global__ __launch_bounds__(kBlock) void k_golden_eval(int32_t n, const double *atdy, const double *qatdy,
                                                        const double *dq, const double *xs, int npts,
                                                        double *partials) {
    __shared__ double sh[4];
    __shared__ double xsh[kMaxGoldenPts];
    if ((int)threadIdx.x < npts) xsh[threadIdx.x] = xs[threadIdx.x];
    __syncthreads();
    for (int k0 = 0; k0 < npts; k0 += kMaxRed) {
        double acc[kMaxRed];
#pragma unroll
        for (int k = 0; k < kMaxRed; ++k) acc[k] = 0.0;
        GRID_STRIDE(i, n) {
            const double c = atdy[i], d = qatdy[i], q = dq[i];
#pragma unroll
            for (int k = 0; k < kMaxRed; ++k)
                if (k0 + k < npts) acc[k] += c * (d / (1.0 + xsh[k0 + k] * q));
        }
        for (int k = 0; k < kMaxRed && k0 + k < npts; ++k) {
            const double r = block_sum(acc[k], sh);
            if (threadIdx.x == 0) partials[((int64_t)blockIdx.x * kMaxGoldenPts) + k0 + k] = r;
        }
    }
}
__global__ __launch_bounds__(kBlock) void k_golden_sum(int nblocks, int npts, const double *partials,
                                                       double *out) {
    __shared__ double sh[4];
    for (int k = 0; k < npts; ++k) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) acc += partials[(int64_t)b * kMaxGoldenPts + k];
        const double r = block_sum(acc, sh);
        if (threadIdx.x == 0) out[k] = r;
    }
}
void launch_golden_eval(hipStream_t st, int32_t n, const double *atdy, const double *qatdy,
                        const double *dq, const double *xs, int npts, double *partials, int nblocks,
                        double *out) {
    hipLaunchKernelGGL(k_golden_eval, dim3(nblocks), dim3(kBlock), 0, st, n, atdy, qatdy, dq, xs, npts, partials);
    hipLaunchKernelGGL(k_golden_sum, dim3(1), dim3(kBlock), 0, st, nblocks, npts, partials, out);
}

// ------------------------------------------------------------ power iteration
__global__ void k_scale_by_rnorm(int64_t n, const double *z, const double *nrm2, double *q, const int32_t *done) {
    if (*done) return;
    const double nz = sqrt(*nrm2);
    GRID_STRIDE(i, n) q[i] = z[i] / nz;
}
__global__ void k_power_residual(int64_t n, const double *z, const double *lam, double *q, const int32_t *done) {
    if (*done) return;
    const double l = *lam;
    GRID_STRIDE(i, n) q[i] = z[i] - l * q[i];
}
__global__ void k_power_check(const double *lam, const double *qq, const double *zz, PowerState *ps, double tol,
                              int32_t max_iter) {
    if (ps->done) return;
    const double l = *lam;
    const double err = sqrt(*qq) / (sqrt(*zz) + l);
    ps->lambda = l;
    ps->err = err;
    ps->iters += 1;
    if (err < tol || ps->iters >= max_iter) ps->done = 1;
}
void launch_scale_by_rnorm(hipStream_t st, int64_t n, const double *z, const double *nrm2, double *q,
                           const int32_t *done) {
    hipLaunchKernelGGL(k_scale_by_rnorm, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, z, nrm2, q, done);
}
void launch_power_residual(hipStream_t st, int64_t n, const double *z, const double *lam, double *q,
                           const int32_t *done) {
    hipLaunchKernelGGL(k_power_residual, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, z, lam, q, done);
}
void launch_power_check(hipStream_t st, const double *lam, const double *qq, const double *zz,
                        PowerState *ps, double tol, int32_t max_iter) {
    hipLaunchKernelGGL(k_power_check, dim3(1), dim3(1), 0, st, lam, qq, zz, ps, tol, max_iter);
}

// ------------------------------------------------------------ scaling
__global__ void k_row_stat(CsrView M, CsrView Q, int has_q, int use_sum, double *out) {
    GRID_STRIDE(i, M.nrows) {
        double a = 0.0, b = 0.0;
        for (int k = M.rowptr[i]; k < M.rowptr[i + 1]; ++k) {
            const double v = fabs(M.val[k]);
            a = use_sum ? a + v : nmax(a, v);
        }
        if (has_q)
            for (int k = Q.rowptr[i]; k < Q.rowptr[i + 1]; ++k) {
                const double v = fabs(Q.val[k]);
                b = use_sum ? b + v : nmax(b, v);
            }
        double r = sqrt(has_q ? (use_sum ? a + b : nmax(a, b)) : a);
        if (r == 0.0) r = 1.0;
        out[i] = r;
    }
}
void launch_row_stat(hipStream_t st, const CsrView &M, const CsrView *Q, int use_sum, double *out) {
    if (M.nrows <= 0) return;
    CsrView q = Q ? *Q : CsrView{nullptr, nullptr, nullptr, 0};
    hipLaunchKernelGGL(k_row_stat, dim3(ew_blocks(M.nrows)), dim3(kBlock), 0, st, M, q, Q ? 1 : 0, use_sum, out);
}

__global__ void k_block_scale(int32_t ncones, const int32_t *idx, double *t, int32_t strategy, int32_t phase,
                              int32_t loc) {
    GRID_STRIDE(c, ncones) {
        const int s = idx[c], e = idx[c + 1];
        if (s >= e) continue;
        double bm = t[s], ss = 0.0;
        for (int j = s; j < e; ++j) {
            bm = nmax(bm, t[j]);
            ss += t[j] * t[j];
        }
        const double rms = sqrt(ss / (double)(e - s));
        const double v = soc_block_value(e - s, bm, rms, strategy, phase, loc, ncones);
        for (int j = s; j < e; ++j) t[j] = v;
    }
}
void launch_block_scale(hipStream_t st, int32_t ncones, const int32_t *idx, double *t, int32_t strategy,
                        int32_t phase, int32_t loc) {
    if (ncones > 0)
        hipLaunchKernelGGL(k_block_scale, dim3(ew_blocks(ncones)), dim3(kBlock), 0, st, ncones, idx, t, strategy,
                           phase, loc);
}
__global__ void k_clip(int64_t n, double *v) {
    GRID_STRIDE(i, n) { const double x = v[i]; v[i] = x > 1e4 ? 1e4 : (x < 1e-4 ? 1e-4 : x); }
}
void launch_clip(hipStream_t st, int64_t n, double *v) {
    if (n > 0) hipLaunchKernelGGL(k_clip, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, v);
}
// val = ((1/r[i]) * val) * (1/c[j]) with (i, j) the entry's (row, col) of the
// operator being scaled, or its transpose when `trans` (keeps A and A^T bit-equal)
__global__ void k_scale_csr(int32_t nrows, const int32_t *rp, const int32_t *col, double *val, const double *r,
                            const double *c, int trans) {
    GRID_STRIDE(i, nrows) {
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = col[k];
            val[k] = trans ? ((1.0 / r[j]) * val[k]) * (1.0 / c[i]) : ((1.0 / r[i]) * val[k]) * (1.0 / c[j]);
        }
    }
}
void launch_scale_csr(hipStream_t st, int32_t nrows, const int32_t *rowptr, const int32_t *col,
                      double *val, const double *r, const double *c) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_scale_csr, dim3(ew_blocks(nrows)), dim3(kBlock), 0, st, nrows, rowptr, col, val, r, c, 0);
}
void launch_scale_csr_t(hipStream_t st, int32_t nrows, const int32_t *rowptr, const int32_t *col,
                        double *val, const double *r, const double *c) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_scale_csr, dim3(ew_blocks(nrows)), dim3(kBlock), 0, st, nrows, rowptr, col, val, r, c, 1);
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
__global__ void k_scale_scalar(int64_t n, double *v, double mul, double div) {
    GRID_STRIDE(i, n) v[i] = div != 1.0 ? v[i] / div : v[i] * mul;
}
void launch_scale_scalar(hipStream_t st, int64_t n, double *v, double mul, double div) {
    if (n > 0) hipLaunchKernelGGL(k_scale_scalar, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, v, mul, div);
}
__global__ __launch_bounds__(kBlock) void k_finite_stats(int64_t n, const double *v, double *partials) {
    __shared__ double sh[4];
    double ss = 0, cnt = 0, mx = 0;
    GRID_STRIDE(i, n) {
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
void launch_finite_stats(hipStream_t st, int64_t n, const double *v, double *partials, int nblocks, double *out) {
    hipLaunchKernelGGL(k_finite_stats, dim3(nblocks), dim3(kBlock), 0, st, n, v, partials);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, st, nblocks, 3, 0x4u, partials, out);
}
__global__ __launch_bounds__(kBlock) void k_bound_stats(int32_t m, const double *AL, const double *AU,
                                                        double *partials) {
    __shared__ double sh[4];
    double ss = 0, mx = 0;
    GRID_STRIDE(i, m) {
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
void launch_bound_stats(hipStream_t st, int32_t m, const double *AL, const double *AU, double *partials,
                        int nblocks, double *out) {
    hipLaunchKernelGGL(k_bound_stats, dim3(nblocks), dim3(kBlock), 0, st, m, AL, AU, partials);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, st, nblocks, 2, 0x2u, partials, out);
}
// check_Q_diagonal: diag_Q and count of coordinates where colsum|Q| is not ≈ |diag|
// (Q is symmetric here, so row sums stand for column sums); out[1] = off-diagonal entries
__global__ __launch_bounds__(kBlock) void k_q_diag(int32_t n, const int32_t *rp, const int32_t *col,
                                                   const double *val, double *diag, double *partials) {
    __shared__ double sh[4];
    double bad = 0, off = 0;
    GRID_STRIDE(i, n) {
        double d = 0.0, rs = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            rs += fabs(val[k]);
            if (col[k] == i) d += val[k];
            else off += 1.0;
        }
        diag[i] = d;
        const double a = rs, b = fabs(d);
        const double rtol = 1.4901161193847656e-08;
        if (!(a == b || fabs(a - b) <= rtol * fmax(a, b))) bad += 1.0;
    }
    const double a = block_sum(bad, sh), b = block_sum(off, sh);
    if (threadIdx.x == 0) {
        partials[(int64_t)blockIdx.x * kMaxRed + 0] = a;
        partials[(int64_t)blockIdx.x * kMaxRed + 1] = b;
    }
}
void launch_q_diag(hipStream_t st, int32_t n, const int32_t *rowptr, const int32_t *col, const double *val,
                   double *diag, double *partials, int nblocks, double *out) {
    hipLaunchKernelGGL(k_q_diag, dim3(nblocks), dim3(kBlock), 0, st, n, rowptr, col, val, diag, partials);
    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, st, nblocks, 2, 0x0u, partials, out);
}
__global__ void k_q_factors(int32_t n, const double *dq, double s, double *f2, double *f, double *f1, double *fm) {
    GRID_STRIDE(i, n) {
        const double a = 1.0 / (1.0 + s * dq[i]);
        f2[i] = a;
        f[i] = s * a;
        f1[i] = s * dq[i] * a;
        fm[i] = (s * s) * a;
    }
}
void launch_q_factors(hipStream_t st, int32_t n, const double *dq, double sigma, double *fact2, double *fact,
                      double *fact1, double *fact_m) {
    if (n > 0) hipLaunchKernelGGL(k_q_factors, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, dq, sigma, fact2, fact, fact1, fact_m);
}
__global__ void k_init_zbar(int32_t n, const double *xb, const double *qx, const double *aty, const double *c,
                            double sigma, double *zb) {
    GRID_STRIDE(j, n) {
        const double tmp = qx ? (-qx[j] + aty[j]) - c[j] : aty[j] - c[j];
        const double zr = xb[j] + sigma * tmp;
        zb[j] = (xb[j] - zr) / sigma;
    }
}
void launch_init_zbar(hipStream_t st, int32_t n, const double *x_bar, const double *qx, const double *aty,
                      const double *c, double sigma, double *z_bar) {
    if (n > 0) hipLaunchKernelGGL(k_init_zbar, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, x_bar, qx, aty, c, sigma, z_bar);
}
__global__ void k_init_s(int32_t m, const double *ax, const double *y, const double *AL, const double *AU, double f1,
                         double *s) {
    GRID_STRIDE(i, m) s[i] = jclamp(ax[i] - f1 * y[i], AL[i], AU[i]);
}
void launch_init_s(hipStream_t st, int32_t m, const double *ax, const double *y, const double *AL, const double *AU,
                   double f1, double *s) {
    if (m > 0) hipLaunchKernelGGL(k_init_s, dim3(ew_blocks(m)), dim3(kBlock), 0, st, m, ax, y, AL, AU, f1, s);
}
__global__ void k_scale_vd(int64_t n, const double *v, const double *d, double s, double *out) {
    GRID_STRIDE(i, n) out[i] = (v[i] * d[i]) / s;
}
void launch_scale_vd(hipStream_t st, int64_t n, const double *v, const double *d, double s, double *out) {
    if (n > 0) hipLaunchKernelGGL(k_scale_vd, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, v, d, s, out);
}
__global__ void k_unscale(int64_t n, const double *v, const double *d, double a, int mul, double *out) {
    GRID_STRIDE(i, n) out[i] = mul ? a * (v[i] * d[i]) : a * (v[i] / d[i]);
}
void launch_unscale(hipStream_t st, int64_t n, const double *v, const double *d, double a, int mul, double *out) {
    if (n > 0) hipLaunchKernelGGL(k_unscale, dim3(ew_blocks(n)), dim3(kBlock), 0, st, n, v, d, a, mul, out);
}

__global__ void k_gather_perm(int64_t nnz, const double *src, const int32_t *perm, double *dst) {
    GRID_STRIDE(k, nnz) dst[k] = src[perm[k]];
}
void launch_gather_perm(hipStream_t st, int64_t nnz, const double *src, const int32_t *perm, double *dst) {
    if (nnz > 0) hipLaunchKernelGGL(k_gather_perm, dim3(ew_blocks(nnz)), dim3(kBlock), 0, st, nnz, src, perm, dst);
}

// symmetrize Q values in place through the transpose-partner map: v = (v + v[tpos]) / 2
__global__ void k_sym_q(int64_t nnz, const double *src, const int32_t *tpos, double *dst) {
    GRID_STRIDE(k, nnz) dst[k] = (src[k] + src[tpos[k]]) / 2.0;
}
void launch_sym_q(hipStream_t st, int64_t nnz, const double *src, const int32_t *tpos, double *dst) {
    if (nnz > 0) hipLaunchKernelGGL(k_sym_q, dim3(ew_blocks(nnz)), dim3(kBlock), 0, st, nnz, src, tpos, dst);
}

}  // namespace hpr
