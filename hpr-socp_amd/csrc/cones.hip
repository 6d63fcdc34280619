// cones.hip — stand-alone projection of a vector's cone blocks onto second-order
// cones, on device pointers (hprsocp_project_cones).
//
// Each cone k occupies rows [idx[k], idx[k+1]) as (t, xi) and is replaced by
// Pi_Q(t, xi) (reference CPU projection, kernels.jl:7585-7663):
//   |xi| <= -t  -> 0;   |xi| <= t -> unchanged;
//   otherwise   -> ((t + |xi|) / 2) * (1, xi / |xi|).
// A cone's xi rows are cut into tiles of kConeTile rows, one workgroup per tile,
// so a cone of 10^5 rows spreads over ~50 workgroups instead of serialising its
// norm on one.  Three launches:
//   k_cone_tiles    one workgroup: validates idx, scans the per-cone tile counts;
//   k_cone_partial  one workgroup per tile: sum of xi^2 over the tile -> part[];
//                   the cone's first tile also records t (so in == out is safe);
//   k_cone_finish   one workgroup per tile: |xi|^2 of its cone as the fixed-order
//                   sum of the cone's partials, then projects its rows.
// The scratch (tile offsets, partial sums, t values, a status word) is device
// memory beyond the input and output: allocated per call by
// hprsocp_project_cones, a fixed file-scope device array in
// hprsocp_project_cones_noalloc.  Every launch writes all of the scratch it
// reads (offsets and status in k_cone_tiles, partials and t in k_cone_partial),
// so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/hprsocp.h"
#include "devutil.h"

namespace hpr {
namespace {

constexpr int kConeTile = 2048;    // xi rows per workgroup
constexpr int kConeThreads = 256;  // 4 x wave64

// Scratch carved out of one base pointer.
struct ConeScratch {
    double *part;    // one partial |xi|^2 per tile
    double *tval;    // t of each cone
    int32_t *off;    // tile offsets, ncones + 1
    int32_t *flag;   // 0 = valid cone offsets, 1 = rejected (call was a no-op)
};

int64_t tile_bound(int64_t n, int32_t ncones) { return (n + kConeTile - 1) / kConeTile + ncones; }

size_t scratch_bytes(int64_t max_tiles, int32_t max_cones) {
    return (size_t)max_tiles * 8 + (size_t)max_cones * 8 + ((size_t)max_cones + 1) * 4 + 4;
}

ConeScratch carve(void *base, int64_t max_tiles, int32_t max_cones) {
    char *p = (char *)base;
    ConeScratch s;
    s.part = (double *)p;
    s.tval = (double *)(p + (size_t)max_tiles * 8);
    s.off = (int32_t *)(p + (size_t)max_tiles * 8 + (size_t)max_cones * 8);
    s.flag = s.off + max_cones + 1;
    return s;
}

// Cone of tile b: the k with off[k] <= b < off[k+1] (cones without tiles skipped).
__device__ __forceinline__ int32_t cone_of(const int32_t *off, int32_t ncones, int32_t b) {
    int32_t lo = 0, hi = ncones;  // off[lo] <= b < off[hi]
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (off[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One workgroup of 1024 threads.  Offsets must start at >= 0, never decrease and
// end at <= n; otherwise every offset is 0 (no tile runs) and flag = 1.
__global__ __launch_bounds__(1024) void k_cone_tiles(const int32_t *idx, int32_t ncones, int64_t n,
                                                      ConeScratch s) {
    __shared__ int32_t sh[1024];
    __shared__ int32_t carry, bad;
    const int tid = threadIdx.x;
    if (tid == 0) {
        carry = 0;
        bad = (idx[0] < 0 || (int64_t)idx[ncones] > n) ? 1 : 0;
    }
    __syncthreads();
    for (int32_t base = 0; base < ncones; base += 1024) {
        const int32_t k = base + tid;
        if (k < ncones && idx[k + 1] < idx[k]) bad = 1;
    }
    __syncthreads();
    const int32_t invalid = bad;
    for (int32_t base = 0; base < ncones; base += 1024) {
        const int32_t k = base + tid;
        int32_t c = 0;
        if (k < ncones && !invalid) {
            const int32_t len = idx[k + 1] - idx[k];
            c = len > 0 ? (len + kConeTile - 1) / kConeTile : 0;
        }
        sh[tid] = c;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {  // inclusive scan
            const int32_t v = tid >= o ? sh[tid - o] : 0;
            __syncthreads();
            sh[tid] += v;
            __syncthreads();
        }
        if (k < ncones) s.off[k] = carry + sh[tid] - c;
        __syncthreads();
        if (tid == 1023) carry += sh[1023];
        __syncthreads();
    }
    if (tid == 0) {
        s.off[ncones] = carry;
        s.flag[0] = invalid;
    }
}

__global__ __launch_bounds__(kConeThreads) void k_cone_partial(const double *in, const int32_t *idx,
                                                                int32_t ncones, ConeScratch s) {
    __shared__ double sh[4];
    const int32_t b = blockIdx.x;
    if (b >= s.off[ncones]) return;  // uniform over the workgroup
    const int32_t k = cone_of(s.off, ncones, b);
    const int32_t st = idx[k], e = idx[k + 1], j = b - s.off[k];
    const int64_t r0 = (int64_t)st + 1 + (int64_t)j * kConeTile;
    const int64_t r1 = r0 + kConeTile < e ? r0 + kConeTile : e;
    double acc = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kConeThreads) {
        const double v = in[r];
        acc = fma(v, v, acc);
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        s.part[b] = acc;
        if (j == 0) s.tval[k] = in[st];
    }
}

__global__ __launch_bounds__(kConeThreads) void k_cone_finish(const double *in, double *out, const int32_t *idx,
                                                               int32_t ncones, ConeScratch s) {
    const int32_t b = blockIdx.x;
    if (b >= s.off[ncones]) return;
    const int32_t k = cone_of(s.off, ncones, b);
    const int32_t st = idx[k], e = idx[k + 1], j = b - s.off[k];
    double ns2 = 0.0;
    for (int32_t p = s.off[k]; p < s.off[k + 1]; ++p) ns2 += s.part[p];
    const double t = s.tval[k], ns = sqrt(ns2);
    // mode 0: zero, 1: keep, 2: scale onto the boundary
    const int mode = ns <= -t ? 0 : (ns <= t ? 1 : 2);
    const double pt = 0.5 * (ns + t);
    const double alpha = mode == 2 ? pt / ns : 1.0;
    const int64_t r0 = (int64_t)st + 1 + (int64_t)j * kConeTile;
    const int64_t r1 = r0 + kConeTile < e ? r0 + kConeTile : e;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kConeThreads) {
        const double v = in[r];
        out[r] = mode == 0 ? 0.0 : (mode == 1 ? v : alpha * v);
    }
    if (j == 0 && threadIdx.x == 0) out[st] = mode == 0 ? 0.0 : (mode == 1 ? t : pt);
}

// Launches the three kernels over a scratch that holds max_tiles tiles.
hipError_t launch_cones(const double *in, double *out, int64_t n, const int32_t *idx, int32_t ncones,
                        int64_t max_tiles, const ConeScratch &s, hipStream_t st) {
    k_cone_tiles<<<1, 1024, 0, st>>>(idx, ncones, n, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const unsigned grid = (unsigned)max_tiles;
    k_cone_partial<<<grid, kConeThreads, 0, st>>>(in, idx, ncones, s);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_cone_finish<<<grid, kConeThreads, 0, st>>>(in, out, idx, ncones, s);
    return hipGetLastError();
}

// ---- the allocation-free entry point's scratch
// Limits of hprsocp_project_cones_noalloc: cones per call, and tiles per call
// (tiles = ceil(n / kConeTile) + ncones bounds the tiles a call can have).
constexpr int32_t kNoAllocCones = 65536;
constexpr int64_t kNoAllocTiles = 131072;
constexpr size_t kNoAllocBytes = (size_t)kNoAllocTiles * 8 + (size_t)kNoAllocCones * 8 + ((size_t)kNoAllocCones + 1) * 4 + 4;
// One copy per GPU, held until the process ends; calls on one GPU share it.
static __device__ __align__(256) double scratch[(kNoAllocBytes + 7) / 8];

// Device address of `scratch` on each device (the host-side name is only a
// stand-in), looked up once per device with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

thread_local std::string g_cone_error;

int cone_fail(int code, const std::string &msg) {
    g_cone_error = msg;
    return code;
}

int check_args(const double *in, double *out, int64_t n, const int32_t *idx, int32_t ncones) {
    if (n < 0 || ncones < 0) return cone_fail(HPRSOCP_ERR_INVALID, "project_cones: negative n or ncones");
    if (n > INT32_MAX) return cone_fail(HPRSOCP_ERR_INVALID, "project_cones: n exceeds the int32 row range");
    if (ncones > 0 && (!in || !out || !idx))
        return cone_fail(HPRSOCP_ERR_INVALID, "project_cones: NULL in, out or cone_idx");
    return HPRSOCP_OK;
}

}  // namespace
}  // namespace hpr

using namespace hpr;

extern "C" {

const char *hprsocp_project_cones_last_error(void) { return g_cone_error.c_str(); }

int hprsocp_project_cones(const double *in, double *out, int64_t n, const int32_t *cone_idx, int32_t ncones,
                          void *stream) {
    int rc = check_args(in, out, n, cone_idx, ncones);
    if (rc != HPRSOCP_OK || ncones == 0) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int64_t tiles = tile_bound(n, ncones);
    if (tiles > INT32_MAX) return cone_fail(HPRSOCP_ERR_INVALID, "project_cones: too many tiles");
    const size_t bytes = scratch_bytes(tiles, ncones);
    void *base = nullptr;
    hipError_t e = hipMallocAsync(&base, bytes, st);
    if (e != hipSuccess) return cone_fail(HPRSOCP_ERR_ALLOC, std::string("project_cones: ") + hipGetErrorString(e));
    const ConeScratch s = carve(base, tiles, ncones);
    e = launch_cones(in, out, n, cone_idx, ncones, tiles, s, st);
    int32_t flag = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, s.flag, sizeof(flag), hipMemcpyDeviceToHost, st);
    const hipError_t ef = hipFreeAsync(base, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = ef != hipSuccess ? ef : es;
    if (e != hipSuccess) return cone_fail(HPRSOCP_ERR_DEVICE, std::string("project_cones: ") + hipGetErrorString(e));
    if (flag)
        return cone_fail(HPRSOCP_ERR_INVALID,
                         "project_cones: cone_idx must start at >= 0, never decrease and end at <= n");
    return HPRSOCP_OK;
}

int hprsocp_project_cones_noalloc(const double *in, double *out, int64_t n, const int32_t *cone_idx,
                                  int32_t ncones, void *stream) {
    int rc = check_args(in, out, n, cone_idx, ncones);
    if (rc != HPRSOCP_OK || ncones == 0) return rc;
    if (ncones > kNoAllocCones)
        return cone_fail(HPRSOCP_ERR_INVALID, "project_cones_noalloc: ncones " + std::to_string(ncones) +
                                                  " exceeds the limit " + std::to_string(kNoAllocCones));
    const int64_t tiles = tile_bound(n, ncones);
    if (tiles > kNoAllocTiles)
        return cone_fail(HPRSOCP_ERR_INVALID, "project_cones_noalloc: ceil(n / " + std::to_string(kConeTile) +
                                                  ") + ncones = " + std::to_string(tiles) + " exceeds the limit " +
                                                  std::to_string(kNoAllocTiles));
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return cone_fail(HPRSOCP_ERR_DEVICE, "project_cones_noalloc: no usable current device");
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void *ptr = nullptr;
            e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr)
                return cone_fail(HPRSOCP_ERR_DEVICE, std::string("project_cones_noalloc: hipGetSymbolAddress: ") +
                                                         hipGetErrorString(e));
            g_scratch_ptr[dev] = ptr;
        }
        base = g_scratch_ptr[dev];
    }
    const ConeScratch s = carve(base, kNoAllocTiles, kNoAllocCones);
    e = launch_cones(in, out, n, cone_idx, ncones, tiles, s, (hipStream_t)stream);
    if (e != hipSuccess)
        return cone_fail(HPRSOCP_ERR_DEVICE, std::string("project_cones_noalloc: ") + hipGetErrorString(e));
    return HPRSOCP_OK;
}

}  // extern "C"
