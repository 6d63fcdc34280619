// hotpath.h — data structures and launch wrappers of the fused HPR kernels,
// templated on the value type T of matrices and iterates (double for the FP64
// method, float for the FP32 bulk of the mixed-precision form).  Cone norms
// (the t^2 - |xi|^2 test) are always accumulated in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hpr {

constexpr int kBlock = 256;       // threads per workgroup (4 wavefronts of 64)
constexpr int kTileRows = 256;    // max rows of one tile (one row per thread in epilogues)
constexpr int kStreamNnz = 2048;  // LDS product buffer of the CSR-stream SpMV paths
constexpr int kSliceRows = 1024;  // rows of one slice tile
constexpr int kSliceTileNnz = 60000;  // cap of one slice tile (per-slice offsets fit uint16)
// Column-sorted slice tiles: a chunk of <= SortSpec<T>::chunk consecutive
// (row-major) entries of one slice tile is stored sorted by column, each entry
// packed as (column - slice start) << pos_bits | (row-major position in the
// chunk), so that the lanes of one gather instruction hit few distinct cache
// lines.  A chunk's products fill 64 KB of LDS in either precision: FP32
// chunks hold twice the entries, so their columns lie twice as dense.
template <typename T>
struct SortSpec {
    static constexpr int chunk = 65536 / (int)sizeof(T);       // entries of one chunk (power of two)
    static constexpr int pos_bits = sizeof(T) == 8 ? 13 : 14;  // log2(chunk)
    static constexpr int col_bits = 32 - pos_bits;             // slice width limit 2^col_bits
    static constexpr int rows = chunk / 2;                     // rows of one sorted slice tile
};
constexpr int kSortBlock = 512;    // threads of the sorted-slice kernel (default; 1024 where it times faster)
constexpr int kSortChunk = SortSpec<double>::chunk;       // the FP64 layout (gather probe)
constexpr int kSortPosBits = SortSpec<double>::pos_bits;

enum TileKind : int32_t { TILE_LINEAR = 0, TILE_CONES = 1, TILE_HUGE = 2 };

// One workgroup's unit of work: rows [rb, re) of an operator.
//   LINEAR : independent rows (box coordinates / linear constraint rows)
//   CONES  : whole cones c0..c1-1 (each with at most kTileRows rows)
//   HUGE   : a row segment of a long cone; c0 = huge-cone id, c1 = partial slot
// g  : SpMV phase: 0 = CSR-stream (the tile's <= kStreamNnz products are
//      gathered by the whole workgroup into LDS, then summed per row in CSR
//      order); otherwise lanes per row (1..64, or 256 = whole workgroup)
// g2 : lanes per cone of the projection phase (CONES tiles)
// inst: the problem instance the tile's rows belong to (tiles never straddle
//      instances; 0 for a single problem)
struct Tile {
    int32_t rb, re, kind, g, c0, c1, g2, inst;
};

// Per-instance scalars of the iteration, read by every tile of the instance.
// h1 = 1/(inner0 + kstep + 2) with kstep the iterations since inner0 was set.
// active = 0 retires the instance; fp32 selects the precision its updates run
// in (the float kernels update only fp32 instances, the double kernels only
// the others).
struct InstScal {
    double sigma, lambda_A, lambda_Q, b_scale, c_scale, inner0;
    int32_t active, fp32;
};

// Whether the update kernels of value type T work on this instance.
template <typename T>
__host__ __device__ inline bool inst_runs(const InstScal &s) {
    return s.active != 0 && (s.fp32 != 0) == (sizeof(T) == sizeof(float));
}

// Long cones split over several tiles (two-phase projection).
struct HugeInfo {
    const int32_t *start, *end, *slot0, *nslot;  // per huge cone
};

// Rows handled by one workgroup of the XCD-sliced SpMV (all slices share them).
struct SliceTile {
    int32_t rb, re;
};

template <typename T>
struct CsrT {
    const int32_t *rowptr = nullptr, *col = nullptr;
    const T *val = nullptr;
    int32_t nrows = 0;
    // When every row's columns are consecutive (dense row blocks), c0[r] is the
    // first column of row r and the lane-group SpMV reads v[c0[r] + (k - rowptr[r])]
    // without loading col[k]; null otherwise.
    const int32_t *c0 = nullptr;
    // XCD column slicing (nslices > 1): the columns are cut into nslices equal
    // ranges and the entries regrouped slice-major.  Workgroup b of the slice
    // kernel works on slice b % nslices — the XCD it is dispatched to — so each
    // XCD gathers from one L2-resident slice of the input vector; per-row
    // partial sums land in s_part[slice * nrows + row].  Row offsets are
    // stored per slice as uint16 relative to the slice tile's base:
    //   entries of (slice p, row r) start at s_base[p*(s_ntiles+1) + t] + s_off[p*nrows + r]
    // with t the slice tile holding r; the tile's end is s_base[... + t + 1].
    // With s_sorted = 1 the entries of each chunk of 2^s_pos_bits entries are in
    // column order and s_col holds packed words (see SortSpec; an FP32 copy may
    // keep the FP64 layout's smaller chunks); s_ncols gives the slice starts
    // (ncols * p / nslices).
    int32_t nslices = 1;
    int32_t s_sorted = 0;
    int32_t s_pos_bits = 0;
    int32_t s_block = kSortBlock;  // threads of the sorted-slice kernel: 512 or 1024 (tuned per operator)
    // s_fused = 1 (sorted layout only): one workgroup per slice tile sums all
    // slices in order and s_part[row] holds the finished row sums.
    int32_t s_fused = 0;
    int32_t s_ncols = 0;
    const int32_t *s_base = nullptr;
    const uint16_t *s_off = nullptr;
    const int32_t *s_col = nullptr;
    const T *s_val = nullptr;
    const SliceTile *s_tiles = nullptr;
    int32_t s_ntiles = 0;
    T *s_part = nullptr;
};
using CsrView = CsrT<double>;

// Tile plan on the device for one operator (A or A^T).
struct DevPlan {
    Tile *tiles = nullptr;
    int32_t ntiles = 0;       // all tiles
    int32_t first_huge = 0;   // tiles[first_huge .. ntiles) are HUGE segments
    int32_t nslots = 0;
    int32_t nhuge = 0;
    HugeInfo huge{};
    int32_t *huge_buf = nullptr;  // backing store of huge.*
};

// ---- hot-path launches ------------------------------------------------------
// Every argument block carries the per-instance scalars (isc, indexed by the
// tile's instance) and the step count since InstScal::inner0 was set.
template <typename T>
struct NoQPrimalArgsT {  // unified_update_zx (noQ), reference kernels.jl:5955-6063, 1775-3790
    const T *y;              // SpMV input (A^T y)
    T *x, *x_hat, *x_bar, *z_bar, *dx;
    const T *last_x, *c, *l, *u;
    const InstScal *isc;
    int64_t kstep;
    int32_t full;
};
template <typename T>
struct DualArgsT {  // unified_update_y (linear + SOC rows), kernels.jl:4354-5368
    const T *v;              // SpMV input (A x_hat or A tempv)
    T *y, *y_bar, *s, *dy;
    const T *last_y, *AL, *AU, *rhs;
    const InstScal *isc;
    int64_t kstep;
    int32_t full, write_ybar;
};
template <typename T>
struct QPPrimalArgsT {  // zxw1 + tempv, kernels.jl:379-1650, 4268-4299
    T *x, *x_hat, *x_bar, *z_bar, *dx, *w_bar, *tempv;
    const T *w, *last_x, *c, *l, *u, *aty, *qw, *dq, *fact1, *fact2;
    const InstScal *isc;
    int64_t kstep;
    int32_t full;
    int32_t diag;                 // Q structurally diagonal: Qw = dq .* w, tempv written here
    int32_t vec_fact;             // Q_is_diag: per-coordinate factors fact1/fact2
};
template <typename T>
struct W2ArgsT {  // update_w2, kernels.jl:5384-5497
    const T *ybar;           // SpMV input (A^T y_bar)
    T *w, *w_bar, *aty, *dw, *atdy, *aty_bar;
    const T *last_w, *last_aty, *fact;
    const InstScal *isc;
    int64_t kstep;
    int32_t full, diag;
};
template <typename T>
struct TempvArgsT {  // tempv = x_hat + sigma*(Qw - Q w_bar) for non-diagonal Q
    const T *x_hat, *qw;
    T *tempv;
    const T *dq;             // factored Q diagonal (may be null)
    const T *wbar;           // needed with dq
    const InstScal *isc;
};
// Delta form (precision FP32_DELTA): the state is the FP64 Halpern anchor of
// the restart cycle plus an FP32 delta, x = last_x + dx, y = last_y + dy,
// w = last_w + dw, aty = last_aty + daty.  The SpMV inputs are deltas too
// (x_hat - last_x, tempv - last_x, y_bar - last_y, w_bar - last_w), so every
// SpMV runs in FP32 on vectors that shrink as the cycle converges; the images
// of the anchors are folded into two FP64 vectors computed once per cycle:
//   ra  = last_x + sigma (last_aty - Q last_w - c)        (QP)
//       = last_x + sigma (A^T last_y - c)                 (noQ)
//   rya = A last_x - rhs (SOC rows) - lambda_A sigma last_y
// so that the raw values the projections see are ra + (FP32 terms) and
// rya + (FP32 terms), formed in FP64.
struct DPrimalArgs {
    float *dx, *dxh, *dwb, *dtv;   // dx in/out; x_hat - last_x; w_bar - last_w; tempv - last_x (sdiag)
    const float *dy;               // noQ: SpMV input (A^T dy)
    const float *dw, *daty, *dqw;  // QP: w, aty deltas; Q dw (unless sdiag)
    const double *ra, *xa, *wa, *l, *u, *dq, *fact1, *fact2;
    const InstScal *isc;
    int64_t kstep;
    int32_t sdiag, vec_fact;
};
struct DDualArgs {
    float *dy, *dyb;               // dy in/out; y_bar - last_y
    const double *rya, *ya, *AL, *AU;
    const InstScal *isc;
    int64_t kstep;
};

using NoQPrimalArgs = NoQPrimalArgsT<double>;
using DualArgs = DualArgsT<double>;
using QPPrimalArgs = QPPrimalArgsT<double>;
using W2Args = W2ArgsT<double>;
using TempvArgs = TempvArgsT<double>;

// raw_buf: one T per row (HUGE cones); partials: one double per huge slot
template <typename T>
void launch_noq_primal(hipStream_t st, const DevPlan &plan, const CsrT<T> &AT, const int32_t *var_idx,
                       const NoQPrimalArgsT<T> &a, T *raw_buf, double *partials);
template <typename T>
void launch_dual(hipStream_t st, const DevPlan &plan, const CsrT<T> &A, const int32_t *con_idx,
                 const DualArgsT<T> &a, T *raw_buf, double *partials);
template <typename T>
void launch_qp_primal(hipStream_t st, const DevPlan &plan, const int32_t *var_idx,
                      const QPPrimalArgsT<T> &a, T *raw_buf, double *partials);
// The QP primal step fused behind the F^T product of a factored Q = F^T F + D
// (problems without variable cones): planFT / FT / Ftmp = F w as for
// launch_spmv; Q w = F^T Ftmp + fdiag .* w is formed per variable, stored to
// qw and used at once.  Same values as launch_spmv followed by launch_qp_primal.
template <typename T>
void launch_qp_primal_fq(hipStream_t st, const DevPlan &planFT, const CsrT<T> &FT, const T *Ftmp, const T *fdiag,
                         T *qw, const QPPrimalArgsT<T> &a);
template <typename T>
void launch_w2(hipStream_t st, const DevPlan &plan, const CsrT<T> &AT, const W2ArgsT<T> &a);
// out = M v (+ diag .* v2 when diag != null); plan must be built for M.
template <typename T>
void launch_spmv(hipStream_t st, const DevPlan &plan, const CsrT<T> &M, const T *v, T *out,
                 const T *diag = nullptr, const T *v2 = nullptr);
template <typename T>
void launch_spmv_tempv(hipStream_t st, const DevPlan &plan, const CsrT<T> &M, const T *wbar,
                       const TempvArgsT<T> &a);
// Delta form: primal (QP: no SpMV; noQ: A^T dy) and dual (A v, v = the FP32
// delta of x_hat or tempv); raw values and projections in FP64 (raw_buf FP64).
void launch_dprimal(hipStream_t st, const DevPlan &plan, const CsrT<float> &AT, const int32_t *var_idx,
                    const DPrimalArgs &a, bool qp, double *raw_buf, double *partials);
void launch_ddual(hipStream_t st, const DevPlan &plan, const CsrT<float> &A, const int32_t *con_idx,
                  const float *v, const DDualArgs &a, double *raw_buf, double *partials);
// Delta-form QP primal step fused behind the F^T product of Q dw, as
// launch_qp_primal_fq (dqw: Q dw, stored).
void launch_dprimal_fq(hipStream_t st, const DevPlan &planFT, const CsrT<float> &FT, const float *Ftmp,
                       const float *fdiag, float *dqw, const DPrimalArgs &a);

// Gather-ceiling probe: synthetic entries shaped like the sliced SpMV's
// (T value + packed column/position word, one gather of a T each from the
// slice of block b % slices).  probe_fill writes nch chunks of kStreamNnz
// (random columns) or SortSpec<T>::chunk (column-sorted chunks) entries;
// probe_gather runs the gather phase of k_spmv_slices (random) or
// k_spmv_sorted (sorted) on them and writes per-thread sums to out (nch x
// threads per block).
template <typename T>
void launch_probe_fill(hipStream_t st, int64_t nch, int32_t width, int sorted, uint32_t *pk, T *val);
template <typename T>
void launch_probe_gather(hipStream_t st, int64_t nch, int32_t slices, int32_t width, int sorted, const uint32_t *pk,
                         const T *val, const T *x, T *out);
constexpr int kProbePosBits = 11;  // log2(kStreamNnz): position bits of the random-order probe

}  // namespace hpr
