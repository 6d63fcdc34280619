// hotpath.hip — the fused HPR kernels (SpMV + projection + reflection + Halpern),
// templated on the value type T (double: FP64 method; float: FP32 bulk).
//
// One workgroup (256 threads = 4 wave64) per Tile.  Phase A produces the tile's
// row sums in LDS: CSR-stream (coalesced loads, independent gathers, products
// in LDS, per-row sums in CSR order), lane groups of g lanes per row for long
// rows, or — when the operator is XCD-sliced — the fixed-order sum of the slice
// partials written by k_spmv_slices.  Phase B is the epilogue: box clamp and
// Halpern averaging for LINEAR tiles; for CONES tiles a lane group of g2 lanes
// owns one cone, reduces |xi|^2 in FP64 with shuffles and applies the SOC
// projection in registers; HUGE tiles store raw values and a partial |xi|^2,
// and k_huge_finish projects once all partials of the cone exist.
#include <algorithm>
#include <type_traits>

#include "devutil.h"
#include "hotpath.h"

namespace hpr {

// ---------------------------------------------------------------- SpMV phase
// s + prod[a] + prod[a+1] + ... + prod[e-1], added in index order.  The LDS
// reads are issued four at a time ahead of their adds, so a row waits for LDS
// once per four entries (once for a row of at most three) instead of once per
// entry; the adds keep the sequential order, so the sum is the plain loop's.
template <typename T>
__device__ __forceinline__ T lds_row_sum(const T *prod, int a, int e, T s) {
    int k = a;
    for (; k + 4 <= e; k += 4) {
        const T p0 = prod[k], p1 = prod[k + 1], p2 = prod[k + 2], p3 = prod[k + 3];
        s += p0;
        s += p1;
        s += p2;
        s += p3;
    }
    const int r = e - k;
    if (r > 0) {
        const T p0 = prod[k];
        const T p1 = r > 1 ? prod[k + 1] : T(0);
        const T p2 = r > 2 ? prod[k + 2] : T(0);
        s += p0;
        if (r > 1) s += p1;
        if (r > 2) s += p2;
    }
    return s;
}

// The row sums of a thread's RPT rows over one chunk (row j: chunk entries
// [max(ra[j], c0), min(re[j], c1)) at prod[k - c0]), rows taken four at a time:
// the first two products of all four rows are read before any add (short rows,
// the common case, then cost one LDS wait per four rows), then each row goes
// on in order.  Every row's adds keep the sequential order.
template <typename T, int RPT>
__device__ __forceinline__ void lds_rows_sum(const T *prod, const int32_t *ra, const int32_t *re, int c0, int c1,
                                             T *acc) {
    static_assert(RPT % 4 == 0, "rows in groups of four");
#pragma unroll
    for (int g = 0; g < RPT; g += 4) {
        T q0[4], q1[4];
        int a[4], e[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            a[j] = max(ra[g + j], c0) - c0;
            e[j] = min(re[g + j], c1) - c0;
            q0[j] = a[j] < e[j] ? prod[a[j]] : T(0);
            q1[j] = a[j] + 1 < e[j] ? prod[a[j] + 1] : T(0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T sj = acc[g + j];
            if (a[j] < e[j]) sj += q0[j];
            if (a[j] + 1 < e[j]) sj += q1[j];
            acc[g + j] = lds_row_sum(prod, a[j] + 2, e[j], sj);
        }
    }
}

// acc: the tile's row sums, stored as A (T, or double for the delta form's FP32 SpMV)
template <typename T, typename A>
__device__ void tile_spmv_stream(const CsrT<T> &M, const T *__restrict__ v, const Tile &t, A *acc, T *prod,
                                 int32_t *rp) {
    const int tid = threadIdx.x;
    const int nrows = t.re - t.rb;
    for (int r = tid; r <= nrows; r += kBlock) rp[r] = M.rowptr[t.rb + r];
    __syncthreads();
    const int k0 = rp[0];
    const int nnz = rp[nrows] - k0;
    const int32_t *__restrict__ col = M.col + k0;
    const T *__restrict__ val = M.val + k0;
#pragma unroll 4
    for (int k = tid; k < nnz; k += kBlock) prod[k] = val[k] * v[col[k]];
    __syncthreads();
    if (tid < nrows) {
        acc[tid] = lds_row_sum(prod, rp[tid] - k0, rp[tid + 1] - k0, T(0));
    }
}

// 16-byte vectors of T with only T's alignment (rows start anywhere; global
// loads on CDNA need element alignment only).
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4), aligned(4)));
    static constexpr int W = 4;
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2), aligned(8)));
    static constexpr int W = 2;
};

// Lane-group SpMV: g lanes per row (1..64), or the whole workgroup per row
// (256).  Each lane keeps four independent partial sums so that four gathers
// are in flight per lane (long rows are otherwise bound by the latency of one
// dependent load per step); the partials are combined in a fixed order.  Rows
// with consecutive columns (M.c0) read the matrix and the input without
// column indices, in 16-byte vectors: lane l takes vectors l, l + G, ... of
// the row, four in flight, and the row's last n mod W entries go to the first
// lanes; the order is fixed for a given g.
template <typename T>
__device__ __forceinline__ T row_dot(const CsrT<T> &M, const T *__restrict__ v, int row, int s, int e, int lane,
                                     int G) {
    T a0 = T(0), a1 = T(0), a2 = T(0), a3 = T(0);
    int k = s + lane;
    if (M.c0) {
        using VT = typename Vec16<T>::type;
        constexpr int W = Vec16<T>::W;
        const VT *__restrict__ pw = reinterpret_cast<const VT *>(M.val + s);
        const VT *__restrict__ px = reinterpret_cast<const VT *>(v + M.c0[row]);
        const int nv = (e - s) / W;
        int q = lane;
        for (; q + 3 * G < nv; q += 4 * G) {  // four vector pairs in flight
            const VT w0 = pw[q], x0 = px[q], w1 = pw[q + G], x1 = px[q + G];
            const VT w2 = pw[q + 2 * G], x2 = px[q + 2 * G], w3 = pw[q + 3 * G], x3 = px[q + 3 * G];
#pragma unroll
            for (int c = 0; c < W; c += 2) {
                a0 += w0[c] * x0[c];
                a1 += w0[c + 1] * x0[c + 1];
                a2 += w1[c] * x1[c];
                a3 += w1[c + 1] * x1[c + 1];
                a0 += w2[c] * x2[c];
                a1 += w2[c + 1] * x2[c + 1];
                a2 += w3[c] * x3[c];
                a3 += w3[c + 1] * x3[c + 1];
            }
        }
        for (; q < nv; q += G) {
            const VT w0 = pw[q], x0 = px[q];
#pragma unroll
            for (int c = 0; c < W; c += 2) {
                a0 += w0[c] * x0[c];
                a1 += w0[c + 1] * x0[c + 1];
            }
        }
        for (int t = s + nv * W + lane; t < e; t += G) a0 += M.val[t] * v[M.c0[row] + (t - s)];  // < W left
        return (a0 + a1) + (a2 + a3);
    }
    for (; k + 3 * G < e; k += 4 * G) {
        const int c0 = M.col[k], c1 = M.col[k + G], c2 = M.col[k + 2 * G], c3 = M.col[k + 3 * G];
        const T w0 = M.val[k], w1 = M.val[k + G], w2 = M.val[k + 2 * G], w3 = M.val[k + 3 * G];
        a0 += w0 * v[c0];
        a1 += w1 * v[c1];
        a2 += w2 * v[c2];
        a3 += w3 * v[c3];
    }
    for (; k < e; k += G) a0 += M.val[k] * v[M.col[k]];
    return (a0 + a1) + (a2 + a3);
}

template <typename T, typename A>
__device__ void tile_spmv(const CsrT<T> &M, const T *__restrict__ v, const Tile &t, A *acc, double *sh) {
    const int tid = threadIdx.x;
    const int nrows = t.re - t.rb;
    const int G = t.g;
    if (G <= 64) {
        const int lane = tid & (G - 1), grp = tid / G, ngrp = kBlock / G;
        for (int r = grp; r < nrows; r += ngrp) {
            const int row = t.rb + r;
            T sum = row_dot(M, v, row, M.rowptr[row], M.rowptr[row + 1], lane, G);
            for (int o = G >> 1; o > 0; o >>= 1) sum += __shfl_xor(sum, o, G);
            if (lane == 0) acc[r] = sum;
        }
    } else {
        for (int r = 0; r < nrows; ++r) {
            const int row = t.rb + r;
            T sum = row_dot(M, v, row, M.rowptr[row], M.rowptr[row + 1], tid, kBlock);
            // workgroup sum in T (wave butterflies, fixed order)
            for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
            __syncthreads();
            if ((tid & 63) == 0) sh[tid >> 6] = (double)sum;
            __syncthreads();
            if (tid == 0) acc[r] = (T)sh[0] + (T)sh[1] + (T)sh[2] + (T)sh[3];  // summed in T
        }
    }
}

// Projection scalars of one cone from its head value t and |xi| (FP64).
struct ConeCase {
    int cas;  // 0 polar (-> 0), 1 inside (keep), 2 boundary projection
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
template <typename T>
__device__ __forceinline__ T cone_apply(const ConeCase &c, bool head, T raw) {
    return c.cas == 0 ? T(0) : (c.cas == 1 ? raw : (head ? (T)c.pt : (T)(c.alpha * (double)raw)));
}

// ------------------------------------------------------------ epilogue ops
// bind(inst) loads the instance's scalars (rounded to T); false = instance
// retired (its tiles write nothing).
__device__ __forceinline__ void halpern_weights(const InstScal &s, int64_t k, double &h1, double &h2) {
    h1 = 1.0 / (s.inner0 + (double)k + 2.0);
    h2 = 1.0 - h1;
}

template <typename T>
struct OpNoQPrimal {
    using V = T;  // raw / projected values
    NoQPrimalArgsT<T> a;
    T sigma, h1, h2;
    __device__ bool bind(int inst) {
        const InstScal s = a.isc[inst];
        double d1, d2;
        halpern_weights(s, a.kstep, d1, d2);
        sigma = (T)s.sigma;
        h1 = (T)d1;
        h2 = (T)d2;
        return inst_runs<T>(s);
    }
    __device__ T raw(int i, T acc, bool) const { return a.x[i] + sigma * (acc - a.c[i]); }
    __device__ T lin_proj(int i, T r) const { return jclamp(r, a.l[i], a.u[i]); }
    __device__ void finish(int i, T raw, T p) const {
        const T xi = a.x[i];
        const T xh = T(2) * p - xi;
        if (a.full) {
            a.x_bar[i] = p;
            a.z_bar[i] = (p - raw) / sigma;
            a.dx[i] = p - xi;
        }
        a.x[i] = h2 * xh + h1 * a.last_x[i];
        a.x_hat[i] = xh;
    }
};

template <typename T>
struct OpDual {
    using V = T;  // raw / projected values
    DualArgsT<T> a;
    T f1, f2, h1, h2;
    __device__ bool bind(int inst) {
        const InstScal s = a.isc[inst];
        double d1, d2;
        halpern_weights(s, a.kstep, d1, d2);
        const double g = s.lambda_A * s.sigma;
        f1 = (T)g;
        f2 = (T)(1.0 / g);
        h1 = (T)d1;
        h2 = (T)d2;
        return inst_runs<T>(s);
    }
    __device__ T raw(int i, T acc, bool cone) const {
        return cone ? (acc - a.rhs[i]) - f1 * a.y[i] : acc - f1 * a.y[i];
    }
    __device__ T lin_proj(int i, T r) const { return jclamp(r, a.AL[i], a.AU[i]); }
    __device__ void finish(int i, T raw, T p) const {
        T sv[kStash];
        stash(i, sv);
        finish_with(i, raw, p, sv);
    }
    // cone rows: y and last_y are read with the raw value, before the cone's
    // reduction, and handed to the lane that finishes the row
    static constexpr int kStash = 2;
    __device__ void stash(int i, T *sv) const {
        sv[0] = a.y[i];
        sv[1] = a.last_y[i];
    }
    __device__ void finish_with(int i, T raw, T p, const T *sv) const {
        const T yi = sv[0], lyi = sv[1];
        const T yb = f2 * (p - raw);
        if (a.full) {
            a.s[i] = p;
            a.dy[i] = yb - yi;
        }
        if (a.full || a.write_ybar) a.y_bar[i] = yb;
        a.y[i] = h2 * (T(2) * yb - yi) + h1 * lyi;
    }
};

template <typename T>
struct OpQPPrimal {
    using V = T;  // raw / projected values
    QPPrimalArgsT<T> a;
    T sigma, h1, h2, f1s, f2s;
    __device__ bool bind(int inst) {
        const InstScal s = a.isc[inst];
        double d1, d2;
        halpern_weights(s, a.kstep, d1, d2);
        const double q2 = 1.0 / (1.0 + s.sigma * s.lambda_Q);
        sigma = (T)s.sigma;
        h1 = (T)d1;
        h2 = (T)d2;
        f2s = (T)q2;
        f1s = (T)(1.0 - q2);
        return inst_runs<T>(s);
    }
    __device__ T qw(int i) const { return a.diag ? a.dq[i] * a.w[i] : a.qw[i]; }
    __device__ T raw(int i, T, bool) const { return a.x[i] + sigma * ((-qw(i) + a.aty[i]) - a.c[i]); }
    __device__ T lin_proj(int i, T r) const { return jclamp(r, a.l[i], a.u[i]); }
    __device__ void finish(int i, T raw, T p) const {
        T sv[kStash];
        stash(i, sv);
        finish_with(i, raw, p, sv);
    }
    // cone rows: x, w, last_x and the per-coordinate Q factors are read with the
    // raw value, before the cone's reduction
    static constexpr int kStash = 5;
    __device__ void stash(int i, T *sv) const {
        sv[0] = a.x[i];
        sv[1] = a.w[i];
        sv[2] = a.last_x[i];
        sv[3] = a.vec_fact ? a.fact1[i] : T(0);
        sv[4] = a.vec_fact ? a.fact2[i] : T(0);
    }
    __device__ void finish_with(int i, T raw, T p, const T *sv) const {
        const T xi = sv[0], wi = sv[1];
        const T xh = T(2) * p - xi;
        const T wb = a.vec_fact ? sv[3] * wi + sv[4] * xh : f1s * wi + f2s * xh;
        if (a.full) {
            a.x_bar[i] = p;
            a.z_bar[i] = (p - raw) / sigma;
            a.dx[i] = p - xi;
        }
        a.x[i] = h2 * xh + h1 * sv[2];
        a.x_hat[i] = xh;
        a.w_bar[i] = wb;
        if (a.diag) a.tempv[i] = xh + sigma * (a.dq[i] * wi - a.dq[i] * wb);
    }
};

// OpQPPrimal behind the F^T product of a factored Q = F^T F + D: the tile's
// F^T row sum plus D w is Q w, formed as the stand-alone product forms it
// (OpStore), stored for the tempv step and used in place of the stored Q w.
template <typename T>
struct OpQPPrimalFq : OpQPPrimal<T> {
    T *qw_out;
    const T *fdiag;
    __device__ T raw(int i, T acc, bool) const {
        const T q = fdiag ? acc + fdiag[i] * this->a.w[i] : acc;
        qw_out[i] = q;
        return this->a.x[i] + this->sigma * ((-q + this->a.aty[i]) - this->a.c[i]);
    }
};

template <typename T>
struct OpW2 {
    using V = T;  // raw / projected values
    W2ArgsT<T> a;
    T fs, h1, h2;
    __device__ bool bind(int inst) {
        const InstScal s = a.isc[inst];
        double d1, d2;
        halpern_weights(s, a.kstep, d1, d2);
        fs = (T)(s.sigma / (1.0 + s.sigma * s.lambda_Q));
        h1 = (T)d1;
        h2 = (T)d2;
        return inst_runs<T>(s);
    }
    __device__ T raw(int, T acc, bool) const { return acc; }
    __device__ T lin_proj(int, T r) const { return r; }
    __device__ void finish(int i, T rb, T) const {
        const T wi = a.w[i], atyi = a.aty[i];
        const T f = a.diag ? a.fact[i] : fs;
        const T wbn = a.w_bar[i] + f * (rb - atyi);
        // null anchors read as zero (the delta form iterates on differences from them)
        const T lw = a.last_w ? a.last_w[i] : T(0), la = a.last_aty ? a.last_aty[i] : T(0);
        a.w[i] = h1 * lw + h2 * (T(2) * wbn - wi);
        a.aty[i] = h1 * la + h2 * (T(2) * rb - atyi);
        if (a.full) {
            a.w_bar[i] = wbn;
            a.dw[i] = wbn - wi;
            a.atdy[i] = rb - atyi;
            a.aty_bar[i] = rb;
        }
    }
};

// plain SpMV output (all instances, no scalars)
template <typename T>
struct OpStore {
    using V = T;
    T *out;
    const T *diag, *v2;
    __device__ bool bind(int) { return true; }
    __device__ T raw(int, T acc, bool) const { return acc; }
    __device__ T lin_proj(int, T r) const { return r; }
    __device__ void finish(int i, T r, T) const { out[i] = diag ? r + diag[i] * v2[i] : r; }
};

template <typename T>
struct OpTempv {
    using V = T;  // raw / projected values
    TempvArgsT<T> a;
    T sigma;
    __device__ bool bind(int inst) {
        const InstScal s = a.isc[inst];
        sigma = (T)s.sigma;
        return inst_runs<T>(s);
    }
    __device__ T raw(int, T acc, bool) const { return acc; }
    __device__ T lin_proj(int, T r) const { return r; }
    __device__ void finish(int i, T r, T) const {
        const T qwb = a.dq ? r + a.dq[i] * a.wbar[i] : r;
        a.tempv[i] = a.x_hat[i] + sigma * (a.qw[i] - qwb);
    }
};

// ------------------------------------------------------------ XCD-sliced SpMV
// Block b: slice p = b % P (the XCD it is dispatched to), slice tile b / P.  The
// slice's nonzeros of the tile are streamed (non-temporal: read once) in chunks
// of kStreamNnz products through LDS; every row sums its products in CSR order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_spmv_slices(CsrT<T> M, const T *__restrict__ x) {
    __shared__ T prod[kStreamNnz];
    __shared__ int32_t rps[kSliceRows + 1];
    const int P = M.nslices;
    const int p = blockIdx.x % P;
    const int tix = blockIdx.x / P;
    const SliceTile S = M.s_tiles[tix];
    const int tid = threadIdx.x;
    const int nr = S.re - S.rb;
    const int32_t *base = M.s_base + (int64_t)p * (M.s_ntiles + 1) + tix;
    const int32_t k0 = base[0];
    const int32_t tnnz = base[1] - k0;
    const uint16_t *off = M.s_off + (int64_t)p * M.nrows + S.rb;
    for (int r = tid; r < nr; r += kBlock) rps[r] = off[r];
    if (tid == 0) rps[nr] = tnnz;
    __syncthreads();
    constexpr int RPT = kSliceRows / kBlock;  // rows per thread
    T acc[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) acc[j] = T(0);
    const T *__restrict__ val = M.s_val + k0;
    const int32_t *__restrict__ col = M.s_col + k0;
    for (int c0 = 0; c0 < tnnz; c0 += kStreamNnz) {
        const int c1 = min(c0 + kStreamNnz, tnnz);
#pragma unroll 4
        for (int k = c0 + tid; k < c1; k += kBlock) {
            const T v = __builtin_nontemporal_load(val + k);
            const int c = __builtin_nontemporal_load(col + k);
            prod[k - c0] = v * x[c];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int r = tid + j * kBlock;
            if (r < nr) {
                acc[j] = lds_row_sum(prod, max(rps[r], c0) - c0, min(rps[r + 1], c1) - c0, acc[j]);
            }
        }
        __syncthreads();
    }
    T *out = M.s_part + (int64_t)p * M.nrows + S.rb;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = tid + j * kBlock;
        if (r < nr) out[r] = acc[j];
    }
}

// Column-sorted form of k_spmv_slices.  Each chunk of the slice tile's entries
// is stored in column order, so consecutive lanes gather neighbouring columns
// and one gather instruction touches far fewer cache lines than with the
// row-major order (the gathers, not the stream, bound the row-major kernel on
// random sparsity).  A product is written to LDS at its row-major position,
// and every row sums its products in the row-major order, so the partial sums
// are bit-identical to k_spmv_slices.  Row ranges live in registers: thread t
// owns rows t, t + kSortBlock, ...
template <typename T, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_spmv_sorted(CsrT<T> M, const T *__restrict__ x) {
    using Spec = SortSpec<T>;
    static_assert(Spec::rows % BLOCK == 0, "rows of a slice tile per thread");
    __shared__ T prod[Spec::chunk];
    const int P = M.nslices;
    const int p = blockIdx.x % P;
    const int tix = blockIdx.x / P;
    const SliceTile S = M.s_tiles[tix];
    const int tid = threadIdx.x;
    const int nr = S.re - S.rb;
    const int32_t *base = M.s_base + (int64_t)p * (M.s_ntiles + 1) + tix;
    const int32_t k0 = base[0];
    const int32_t tnnz = base[1] - k0;
    const uint16_t *off = M.s_off + (int64_t)p * M.nrows + S.rb;
    const T *__restrict__ xs = x + (int64_t)M.s_ncols * p / P;
    constexpr int RPT = Spec::rows / BLOCK;
    T acc[RPT];
    int32_t ra[RPT], re[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = tid + j * BLOCK;
        acc[j] = T(0);
        ra[j] = r < nr ? (int32_t)off[r] : 0;
        re[j] = r < nr ? (r + 1 < nr ? (int32_t)off[r + 1] : tnnz) : 0;
    }
    const T *__restrict__ val = M.s_val + k0;
    const uint32_t *__restrict__ pk = (const uint32_t *)M.s_col + k0;
    // chunks of the layout (at most Spec::chunk entries, Spec::rows rows)
    const int pbits = M.s_pos_bits, chunk = 1 << pbits;
    const uint32_t pmask = (uint32_t)chunk - 1u;
    for (int c0 = 0; c0 < tnnz; c0 += chunk) {
        const int c1 = min(c0 + chunk, tnnz);
#pragma unroll 4
        for (int k = c0 + tid; k < c1; k += BLOCK) {
            const uint32_t w = __builtin_nontemporal_load(pk + k);
            const T v = __builtin_nontemporal_load(val + k);
            prod[w & pmask] = v * xs[w >> pbits];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            acc[j] = lds_row_sum(prod, max(ra[j], c0) - c0, min(re[j], c1) - c0, acc[j]);
        }
        __syncthreads();
    }
    T *out = M.s_part + (int64_t)p * M.nrows + S.rb;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = tid + j * BLOCK;
        if (r < nr) out[r] = acc[j];
    }
}

// Slice-fused form of k_spmv_sorted (CsrT::s_fused): one workgroup per slice
// tile walks the slices in order 0..P-1 and carries the running row sums in
// registers, so one finished sum per row reaches memory instead of P partials
// that the tile kernel reads back.  Each slice's products are summed exactly as
// in k_spmv_sorted (from zero, in row-major order) and the slice sums are added
// in slice order starting from zero, as the tile kernel adds the partials: the
// row sums are bit-identical to the unfused form.  The XCDs no longer own one
// slice each; locality rests on the workgroups resident on an XCD walking the
// slices in step (they start together and carry about equal entries per
// slice), so its L2 holds about one or two slices of the input at a time.
template <typename T, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_spmv_sorted_fused(CsrT<T> M, const T *__restrict__ x) {
    using Spec = SortSpec<T>;
    static_assert(Spec::rows % BLOCK == 0, "rows of a slice tile per thread");
    __shared__ T prod[Spec::chunk];
    const int P = M.nslices;
    const int tix = blockIdx.x;
    const SliceTile S = M.s_tiles[tix];
    const int tid = threadIdx.x;
    const int nr = S.re - S.rb;
    constexpr int RPT = Spec::rows / BLOCK;
    T tot[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) tot[j] = T(0);
    const int pbits = M.s_pos_bits, chunk = 1 << pbits;
    const uint32_t pmask = (uint32_t)chunk - 1u;
    for (int p = 0; p < P; ++p) {
        const int32_t *base = M.s_base + (int64_t)p * (M.s_ntiles + 1) + tix;
        const int32_t k0 = base[0];
        const int32_t tnnz = base[1] - k0;
        const uint16_t *off = M.s_off + (int64_t)p * M.nrows + S.rb;
        const T *__restrict__ xs = x + (int64_t)M.s_ncols * p / P;
        T acc[RPT];
        int32_t ra[RPT], re[RPT];
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int r = tid + j * BLOCK;
            acc[j] = T(0);
            ra[j] = r < nr ? (int32_t)off[r] : 0;
            re[j] = r < nr ? (r + 1 < nr ? (int32_t)off[r + 1] : tnnz) : 0;
        }
        const T *__restrict__ val = M.s_val + k0;
        const uint32_t *__restrict__ pk = (const uint32_t *)M.s_col + k0;
        for (int c0 = 0; c0 < tnnz; c0 += chunk) {
            const int c1 = min(c0 + chunk, tnnz);
            // FP64: eight gathers in flight per thread (measured faster than four
            // here, while the unfused kernel and the FP32 chunks keep four)
            if constexpr (sizeof(T) == 8) {
#pragma unroll 8
                for (int k = c0 + tid; k < c1; k += BLOCK) {
                    const uint32_t w = __builtin_nontemporal_load(pk + k);
                    const T v = __builtin_nontemporal_load(val + k);
                    prod[w & pmask] = v * xs[w >> pbits];
                }
            } else {
#pragma unroll 4
                for (int k = c0 + tid; k < c1; k += BLOCK) {
                    const uint32_t w = __builtin_nontemporal_load(pk + k);
                    const T v = __builtin_nontemporal_load(val + k);
                    prod[w & pmask] = v * xs[w >> pbits];
                }
            }
            __syncthreads();
            lds_rows_sum<T, RPT>(prod, ra, re, c0, c1, acc);
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < RPT; ++j) tot[j] += acc[j];
    }
    T *out = M.s_part + S.rb;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = tid + j * BLOCK;
        if (r < nr) out[r] = tot[j];
    }
}

// ------------------------------------------------------------ tile kernel
// SRC: 0 = no SpMV, 1 = CSR SpMV inside the tile, 2 = sum of the XCD-slice partials
enum : int { SRC_NONE = 0, SRC_CSR = 1, SRC_SLICES = 2 };

// Op::kStash > 0: the epilogue can take kStash operands of a cone row read
// early (stash / finish_with) instead of reading them after the cone's reduction.
template <class Op, class = void>
struct op_stash_n : std::integral_constant<int, 0> {};
template <class Op>
struct op_stash_n<Op, std::void_t<decltype(Op::kStash)>> : std::integral_constant<int, Op::kStash> {};

// T: SpMV arithmetic; Op::V: raw and projected values (T, or double for the delta form).
template <typename T, class Op, int SRC, bool CONES>
__global__ __launch_bounds__(kBlock) void k_tile(const Tile *__restrict__ tiles, CsrT<T> M, const T *__restrict__ vin,
                                                 const int32_t *__restrict__ cone_idx, HugeInfo huge, Op op_in,
                                                 typename Op::V *__restrict__ raw_buf,
                                                 double *__restrict__ partials) {
    using V = typename Op::V;
    __shared__ V acc[kTileRows];
    // cone rows' stashed epilogue operands (one-row-per-thread sources only)
    constexpr int kStashN = CONES && SRC != SRC_CSR ? op_stash_n<Op>::value : 0;
    __shared__ V stv[kStashN > 0 ? kStashN : 1][kStashN > 0 ? kTileRows : 1];
    __shared__ double sh[4];
    __shared__ T prod[SRC == SRC_CSR ? kStreamNnz : 1];
    __shared__ int32_t rps[SRC == SRC_CSR ? kTileRows + 1 : 1];
    const Tile t = tiles[blockIdx.x];
    Op op = op_in;
    if (!op.bind(t.inst)) return;  // retired instance (uniform over the workgroup)
    const int tid = threadIdx.x;
    const int nrows = t.re - t.rb;
    // linear rows: raw value, box projection and update, one row per thread
    auto linear_rows = [&](V a) {
        if (tid < nrows) {
            const int i = t.rb + tid;
            const V r = op.raw(i, a, false);
            op.finish(i, r, op.lin_proj(i, r));
        }
    };
    // HUGE segment, phase 1: raw values and a partial |xi|^2
    auto huge_phase1 = [&](V a) {
        double sq = 0.0;
        if (tid < nrows) {
            const int i = t.rb + tid;
            const V r = op.raw(i, a, true);
            raw_buf[i] = r;
            if (i != huge.start[t.c0]) sq = (double)r * (double)r;
        }
        sq = block_sum(sq, sh);
        if (tid == 0) partials[t.c1] = sq;
    };
    // raw_in_acc: acc holds the rows' raw values (not their products) for the cone loop
    bool raw_in_acc = false;
    if constexpr (SRC == SRC_CSR) {
        if (t.g == 0) tile_spmv_stream(M, vin, t, acc, prod, rps);
        else tile_spmv(M, vin, t, acc, sh);
        __syncthreads();
        if (!CONES || t.kind == TILE_LINEAR) {
            linear_rows(tid < nrows ? acc[tid] : V(0));
            return;
        }
        if (t.kind == TILE_HUGE) {
            huge_phase1(tid < nrows ? acc[tid] : V(0));
            return;
        }
    } else {
        // one row per thread: the row's product stays in a register, and a row
        // that no other thread reads (linear rows, HUGE segments) takes no LDS
        // round trip and no barrier, so its loads of the product and of the
        // epilogue's operands go out together
        V a = V(0);
        if constexpr (SRC == SRC_SLICES) {
            if (tid < nrows) {
                const T *pp = M.s_part + t.rb + tid;
                const int np = M.s_fused ? 1 : M.nslices;  // fused: s_part holds the finished sums
                T sum = T(0);
                for (int p = 0; p < np; ++p) sum += pp[(int64_t)p * M.nrows];
                a = sum;
            }
        }
        if (!CONES || t.kind == TILE_LINEAR) {
            linear_rows(a);
            return;
        }
        if (t.kind == TILE_HUGE) {
            huge_phase1(a);
            return;
        }
        // cone rows: raw values formed here, read by the cones' lane groups
        if (tid < nrows) {
            acc[tid] = op.raw(t.rb + tid, a, true);
            if constexpr (kStashN > 0) {
                V sv[kStashN > 0 ? kStashN : 1];
                op.stash(t.rb + tid, sv);
#pragma unroll
                for (int k = 0; k < kStashN; ++k) stv[k][tid] = sv[k];
            }
        }
        raw_in_acc = true;
        __syncthreads();
    }
    // CONES tile: one lane group of g2 lanes per cone
    const int G2 = t.g2;
    const int lane = tid & (G2 - 1), grp = tid / G2, ngrp = kBlock / G2;
    for (int c = t.c0 + grp; c < t.c1; c += ngrp) {
        const int s = cone_idx[c], e = cone_idx[c + 1];
        double sq = 0.0, head = 0.0;
        for (int i = s + lane; i < e; i += G2) {
            const V r = raw_in_acc ? acc[i - t.rb] : op.raw(i, acc[i - t.rb], true);
            acc[i - t.rb] = r;
            if (i == s) head = (double)r;
            else sq += (double)r * (double)r;
        }
        for (int o = G2 >> 1; o > 0; o >>= 1) sq += __shfl_xor(sq, o, G2);
        head = __shfl(head, 0, G2);
        const ConeCase cc = cone_case(head, sqrt(sq));
        for (int i = s + lane; i < e; i += G2) {
            const V r = acc[i - t.rb];
            if constexpr (kStashN > 0) {
                V sv[kStashN > 0 ? kStashN : 1];
#pragma unroll
                for (int k = 0; k < kStashN; ++k) sv[k] = stv[k][i - t.rb];
                op.finish_with(i, r, cone_apply(cc, i == s, r), sv);
            } else {
                op.finish(i, r, cone_apply(cc, i == s, r));
            }
        }
    }
}

// HUGE segment, phase 2: the cone's |xi| from all partials, then the projection.
template <class Op>
__global__ __launch_bounds__(kBlock) void k_huge_finish(const Tile *__restrict__ tiles, HugeInfo huge, Op op_in,
                                                        const typename Op::V *__restrict__ raw_buf,
                                                        const double *__restrict__ partials) {
    using V = typename Op::V;
    __shared__ double sh[4];
    const Tile t = tiles[blockIdx.x];
    Op op = op_in;
    if (!op.bind(t.inst)) return;
    const int h = t.c0;
    const int s = huge.start[h];
    const int slot0 = huge.slot0[h], nslot = huge.nslot[h];
    double sum = 0.0;
    for (int k = threadIdx.x; k < nslot; k += kBlock) sum += partials[slot0 + k];
    sum = block_sum(sum, sh);
    const ConeCase cc = cone_case((double)raw_buf[s], sqrt(sum));
    const int nrows = t.re - t.rb;
    if ((int)threadIdx.x < nrows) {
        const int i = t.rb + threadIdx.x;
        const V r = raw_buf[i];
        op.finish(i, r, cone_apply(cc, i == s, r));
    }
}

// Fused SpMV + epilogue over a plan.  With XCD slicing the SpMV runs first as
// k_spmv_slices and the tile kernel sums the slice partials.
template <typename T, class Op, int SRC, bool CONES>
static void run_tiles(hipStream_t st, const DevPlan &plan, const CsrT<T> &M, const T *v, const int32_t *cone_idx,
                      const Op &op, typename Op::V *raw_buf, double *partials) {
    if (plan.ntiles == 0) return;
    if (SRC == SRC_CSR && M.nslices > 1) {
        if (M.s_ntiles > 0 && M.s_sorted && M.s_fused && M.s_block == 1024)
            hipLaunchKernelGGL((k_spmv_sorted_fused<T, 1024>), dim3(M.s_ntiles), dim3(1024), 0, st, M, v);
        else if (M.s_ntiles > 0 && M.s_sorted && M.s_fused)
            hipLaunchKernelGGL((k_spmv_sorted_fused<T, kSortBlock>), dim3(M.s_ntiles), dim3(kSortBlock), 0, st, M, v);
        else if (M.s_ntiles > 0 && M.s_sorted && M.s_block == 1024)
            hipLaunchKernelGGL((k_spmv_sorted<T, 1024>), dim3(M.s_ntiles * M.nslices), dim3(1024), 0, st, M, v);
        else if (M.s_ntiles > 0 && M.s_sorted)
            hipLaunchKernelGGL((k_spmv_sorted<T, kSortBlock>), dim3(M.s_ntiles * M.nslices), dim3(kSortBlock), 0, st,
                               M, v);
        else if (M.s_ntiles > 0)
            hipLaunchKernelGGL((k_spmv_slices<T>), dim3(M.s_ntiles * M.nslices), dim3(kBlock), 0, st, M, v);
        hipLaunchKernelGGL((k_tile<T, Op, SRC_SLICES, CONES>), dim3(plan.ntiles), dim3(kBlock), 0, st, plan.tiles, M,
                           v, cone_idx, plan.huge, op, raw_buf, partials);
    } else {
        hipLaunchKernelGGL((k_tile<T, Op, SRC, CONES>), dim3(plan.ntiles), dim3(kBlock), 0, st, plan.tiles, M, v,
                           cone_idx, plan.huge, op, raw_buf, partials);
    }
    if (CONES) {
        const int nh = plan.ntiles - plan.first_huge;
        if (nh > 0)
            hipLaunchKernelGGL((k_huge_finish<Op>), dim3(nh), dim3(kBlock), 0, st, plan.tiles + plan.first_huge,
                               plan.huge, op, raw_buf, partials);
    }
}

template <typename T>
void launch_noq_primal(hipStream_t st, const DevPlan &plan, const CsrT<T> &AT, const int32_t *var_idx,
                       const NoQPrimalArgsT<T> &a, T *raw_buf, double *partials) {
    run_tiles<T, OpNoQPrimal<T>, SRC_CSR, true>(st, plan, AT, a.y, var_idx, OpNoQPrimal<T>{a}, raw_buf, partials);
}
template <typename T>
void launch_dual(hipStream_t st, const DevPlan &plan, const CsrT<T> &A, const int32_t *con_idx,
                 const DualArgsT<T> &a, T *raw_buf, double *partials) {
    run_tiles<T, OpDual<T>, SRC_CSR, true>(st, plan, A, a.v, con_idx, OpDual<T>{a}, raw_buf, partials);
}
template <typename T>
void launch_qp_primal(hipStream_t st, const DevPlan &plan, const int32_t *var_idx, const QPPrimalArgsT<T> &a,
                      T *raw_buf, double *partials) {
    CsrT<T> none{};
    run_tiles<T, OpQPPrimal<T>, SRC_NONE, true>(st, plan, none, nullptr, var_idx, OpQPPrimal<T>{a}, raw_buf,
                                                partials);
}
template <typename T>
void launch_qp_primal_fq(hipStream_t st, const DevPlan &planFT, const CsrT<T> &FT, const T *Ftmp, const T *fdiag,
                         T *qw, const QPPrimalArgsT<T> &a) {
    OpQPPrimalFq<T> op{};
    op.a = a;
    op.qw_out = qw;
    op.fdiag = fdiag;
    run_tiles<T, OpQPPrimalFq<T>, SRC_CSR, false>(st, planFT, FT, Ftmp, nullptr, op, nullptr, nullptr);
}
template <typename T>
void launch_w2(hipStream_t st, const DevPlan &plan, const CsrT<T> &AT, const W2ArgsT<T> &a) {
    run_tiles<T, OpW2<T>, SRC_CSR, false>(st, plan, AT, a.ybar, nullptr, OpW2<T>{a}, nullptr, nullptr);
}
template <typename T>
void launch_spmv(hipStream_t st, const DevPlan &plan, const CsrT<T> &M, const T *v, T *out, const T *diag,
                 const T *v2) {
    run_tiles<T, OpStore<T>, SRC_CSR, false>(st, plan, M, v, nullptr, OpStore<T>{out, diag, v2}, nullptr, nullptr);
}
template <typename T>
void launch_spmv_tempv(hipStream_t st, const DevPlan &plan, const CsrT<T> &M, const T *wbar,
                       const TempvArgsT<T> &a) {
    run_tiles<T, OpTempv<T>, SRC_CSR, false>(st, plan, M, wbar, nullptr, OpTempv<T>{a}, nullptr, nullptr);
}

// ------------------------------------------------------------ delta form
// State = FP64 anchor of the restart cycle + FP32 delta (DPrimalArgs).  The
// raw values are ra + (FP32 terms) and rya + (FP32 terms), formed and
// projected in FP64; the reflections and Halpern steps are written as FP32
// differences from the anchors, in which the anchor terms cancel:
//   dx'  = h2 (x_hat - last_x),  dy' = h2 (2 (y_bar - last_y) - dy).
template <bool QP>
struct OpDPrimal {
    using V = double;
    DPrimalArgs a;
    double sigma, h2, f1s, f2s;
    __device__ bool bind(int inst) {
        const InstScal s = a.isc[inst];
        double d1, d2;
        halpern_weights(s, a.kstep, d1, d2);
        sigma = s.sigma;
        h2 = d2;
        f2s = 1.0 / (1.0 + s.sigma * s.lambda_Q);
        f1s = 1.0 - f2s;
        return inst_runs<float>(s);
    }
    __device__ double raw(int i, double acc, bool) const {
        if constexpr (QP) {
            const double qd = a.sdiag ? a.dq[i] * (double)a.dw[i] : (double)a.dqw[i];
            return a.ra[i] + ((double)a.dx[i] + sigma * ((double)a.daty[i] - qd));
        } else {
            return a.ra[i] + ((double)a.dx[i] + sigma * acc);
        }
    }
    __device__ double lin_proj(int i, double r) const { return jclamp(r, a.l[i], a.u[i]); }
    __device__ void finish(int i, double, double p) const {
        const double xa = a.xa[i], dxi = (double)a.dx[i];
        const double dxh = 2.0 * (p - xa) - dxi;  // x_hat - last_x, x_hat = 2p - x
        a.dx[i] = (float)(h2 * dxh);
        a.dxh[i] = (float)dxh;
        if constexpr (QP) {
            const double wa = a.wa[i], dwi = (double)a.dw[i];
            const double xh = 2.0 * p - (xa + dxi), w = wa + dwi;
            const double wb = a.vec_fact ? a.fact1[i] * w + a.fact2[i] * xh : f1s * w + f2s * xh;
            const double dwb = wb - wa;
            a.dwb[i] = (float)dwb;
            // tempv = x_hat + sigma dq (w - w_bar), as a difference from last_x
            if (a.sdiag) a.dtv[i] = (float)(dxh + sigma * (a.dq[i] * (dwi - dwb)));
        }
    }
};

// OpDPrimal<true> behind the F^T product of Q dw (as OpQPPrimalFq): the FP32
// row sum plus D dw is Q dw, rounded to FP32 as the stand-alone product stores it.
struct OpDPrimalFq : OpDPrimal<true> {
    float *dqw_out;
    const float *fdiag;
    __device__ double raw(int i, double acc, bool) const {
        const float s = (float)acc;
        const float q = fdiag ? s + fdiag[i] * a.dw[i] : s;
        dqw_out[i] = q;
        return a.ra[i] + ((double)a.dx[i] + sigma * ((double)a.daty[i] - (double)q));
    }
};

struct OpDDual {
    using V = double;
    DDualArgs a;
    double f1, f2, h2;
    __device__ bool bind(int inst) {
        const InstScal s = a.isc[inst];
        double d1, d2;
        halpern_weights(s, a.kstep, d1, d2);
        f1 = s.lambda_A * s.sigma;
        f2 = 1.0 / f1;
        h2 = d2;
        return inst_runs<float>(s);
    }
    // rya holds A last_x - rhs (SOC rows) - f1 last_y
    __device__ double raw(int i, double acc, bool) const { return a.rya[i] + (acc - f1 * (double)a.dy[i]); }
    __device__ double lin_proj(int i, double r) const { return jclamp(r, a.AL[i], a.AU[i]); }
    __device__ void finish(int i, double raw, double p) const {
        double sv[kStash];
        stash(i, sv);
        finish_with(i, raw, p, sv);
    }
    // cone rows: last_y and dy read before the cone's reduction (as OpDual)
    static constexpr int kStash = 2;
    __device__ void stash(int i, double *sv) const {
        sv[0] = a.ya[i];
        sv[1] = (double)a.dy[i];
    }
    __device__ void finish_with(int i, double raw, double p, const double *sv) const {
        const double yai = sv[0], dyi = sv[1];
        const double dyb = f2 * (p - raw) - yai;  // y_bar - last_y
        a.dy[i] = (float)(h2 * (2.0 * dyb - dyi));
        a.dyb[i] = (float)dyb;
    }
};

void launch_dprimal(hipStream_t st, const DevPlan &plan, const CsrT<float> &AT, const int32_t *var_idx,
                    const DPrimalArgs &a, bool qp, double *raw_buf, double *partials) {
    if (qp) {
        CsrT<float> none{};
        run_tiles<float, OpDPrimal<true>, SRC_NONE, true>(st, plan, none, nullptr, var_idx, OpDPrimal<true>{a},
                                                         raw_buf, partials);
    } else {
        run_tiles<float, OpDPrimal<false>, SRC_CSR, true>(st, plan, AT, a.dy, var_idx, OpDPrimal<false>{a},
                                                          raw_buf, partials);
    }
}
void launch_dprimal_fq(hipStream_t st, const DevPlan &planFT, const CsrT<float> &FT, const float *Ftmp,
                       const float *fdiag, float *dqw, const DPrimalArgs &a) {
    OpDPrimalFq op{};
    op.a = a;
    op.dqw_out = dqw;
    op.fdiag = fdiag;
    run_tiles<float, OpDPrimalFq, SRC_CSR, false>(st, planFT, FT, Ftmp, nullptr, op, nullptr, nullptr);
}
void launch_ddual(hipStream_t st, const DevPlan &plan, const CsrT<float> &A, const int32_t *con_idx,
                  const float *v, const DDualArgs &a, double *raw_buf, double *partials) {
    run_tiles<float, OpDDual, SRC_CSR, true>(st, plan, A, v, con_idx, OpDDual{a}, raw_buf, partials);
}

// ------------------------------------------------------ gather ceiling probe
__device__ __forceinline__ uint64_t probe_hash(uint64_t z) {  // SplitMix64 finaliser
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Entry k of chunk c = k / C, j = k % C.  Random order: a uniform column,
// position j.  Sorted order: the columns of a sorted uniform sample of C
// points of the slice — each wavefront's 64 entries start at j0 * width / C
// and advance by exponential gaps of mean width / C (an inclusive shuffle scan
// over the wavefront; every wavefront of the launch is whole because n and the
// grid stride are multiples of 64) — with positions an odd-multiplier
// permutation of j.
template <typename T>
__global__ void k_probe_fill(int64_t n, int32_t width, int sorted, uint32_t *pk, T *val) {
    const int C = sorted ? SortSpec<T>::chunk : kStreamNnz;
    const int bits = sorted ? SortSpec<T>::pos_bits : kProbePosBits;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t j = (uint32_t)(k % C);
        const uint64_t r = probe_hash((uint64_t)k);
        uint32_t col, pos;
        if (sorted) {
            const double step = (double)width / C;
            const double u = (double)((r >> 11) + 1) * 0x1.0p-53;  // (0, 1]
            const double gap = -log(u) * step;
            double sc = gap;
            const int lane = (int)(threadIdx.x & 63);
            for (int o = 1; o < 64; o <<= 1) {
                const double t = __shfl_up(sc, o, 64);
                if (lane >= o) sc += t;
            }
            const int64_t c = (int64_t)((double)(j & ~63u) * step + (sc - gap));
            col = (uint32_t)(c < width ? c : width - 1);
            pos = (j * 40503u) & (uint32_t)(C - 1);
        } else {
            col = (uint32_t)(r % (uint64_t)width);
            pos = j;
        }
        pk[k] = (col << bits) | pos;
        val[k] = (T)1;
    }
}

template <typename T, int C, int NT, int BITS>
__global__ __launch_bounds__(NT) void k_probe_gather(int32_t slices, int32_t width, const uint32_t *__restrict__ pk,
                                                     const T *__restrict__ val, const T *__restrict__ x,
                                                     T *__restrict__ out) {
    __shared__ T prod[C];
    const int64_t b = blockIdx.x;
    const T *__restrict__ xs = x + (int64_t)(b % slices) * width;
    const int64_t base = b * (int64_t)C;
    constexpr uint32_t mask = (1u << BITS) - 1;
#pragma unroll 4
    for (int k = threadIdx.x; k < C; k += NT) {
        const uint32_t w = __builtin_nontemporal_load(pk + base + k);
        const T v = __builtin_nontemporal_load(val + base + k);
        prod[w & mask] = v * xs[w >> BITS];
    }
    __syncthreads();
    T s = 0;
    for (int k = 0; k < C / NT; ++k) s += prod[threadIdx.x * (C / NT) + k];
    out[b * NT + threadIdx.x] = s;
}

template <typename T>
void launch_probe_fill(hipStream_t st, int64_t nch, int32_t width, int sorted, uint32_t *pk, T *val) {
    const int64_t n = nch * (sorted ? SortSpec<T>::chunk : kStreamNnz);
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(k_probe_fill<T>, dim3((unsigned)blocks), dim3(kBlock), 0, st, n, width, sorted, pk, val);
}

template <typename T>
void launch_probe_gather(hipStream_t st, int64_t nch, int32_t slices, int32_t width, int sorted, const uint32_t *pk,
                         const T *val, const T *x, T *out) {
    if (nch <= 0) return;
    if (sorted)
        hipLaunchKernelGGL((k_probe_gather<T, SortSpec<T>::chunk, kSortBlock, SortSpec<T>::pos_bits>),
                           dim3((unsigned)nch), dim3(kSortBlock), 0, st, slices, width, pk, val, x, out);
    else
        hipLaunchKernelGGL((k_probe_gather<T, kStreamNnz, kBlock, kProbePosBits>), dim3((unsigned)nch), dim3(kBlock),
                           0, st, slices, width, pk, val, x, out);
}

#define HPR_INSTANTIATE(T)                                                                                        \
    template void launch_noq_primal<T>(hipStream_t, const DevPlan &, const CsrT<T> &, const int32_t *,          \
                                       const NoQPrimalArgsT<T> &, T *, double *);                                \
    template void launch_dual<T>(hipStream_t, const DevPlan &, const CsrT<T> &, const int32_t *,                \
                                 const DualArgsT<T> &, T *, double *);                                           \
    template void launch_qp_primal<T>(hipStream_t, const DevPlan &, const int32_t *, const QPPrimalArgsT<T> &,  \
                                      T *, double *);                                                            \
    template void launch_w2<T>(hipStream_t, const DevPlan &, const CsrT<T> &, const W2ArgsT<T> &);              \
    template void launch_qp_primal_fq<T>(hipStream_t, const DevPlan &, const CsrT<T> &, const T *, const T *, T *, \
                                         const QPPrimalArgsT<T> &);                                              \
    template void launch_spmv<T>(hipStream_t, const DevPlan &, const CsrT<T> &, const T *, T *, const T *,      \
                                 const T *);                                                                     \
    template void launch_spmv_tempv<T>(hipStream_t, const DevPlan &, const CsrT<T> &, const T *,                \
                                       const TempvArgsT<T> &);                                                   \
    template void launch_probe_fill<T>(hipStream_t, int64_t, int32_t, int, uint32_t *, T *);                     \
    template void launch_probe_gather<T>(hipStream_t, int64_t, int32_t, int32_t, int, const uint32_t *,          \
                                         const T *, const T *, T *);
HPR_INSTANTIATE(double)
HPR_INSTANTIATE(float)

}  // namespace hpr
