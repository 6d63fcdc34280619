// solver.hip — handle, setup and host control of the HPR iteration behind the
// C ABI of include/hprsocp.h.
//
// Host control mirrors the reference's solve loop (src/algorithm.jl:2765-2905):
// residual checks, restart tests and the sigma update are scalar host logic fed
// by device reductions; everything vector-sized stays on the GPU.  Between two
// checks the stream holds only kernel launches (no host synchronisation): the
// M-norm needed by the restart test is reduced on the device and read at the
// next check.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/hprsocp.h"
#include "kernels.h"
#include "plan.h"
#include "ctl.h"
#include "ctlk.h"
#include "golden.h"
#include "scale_rules.h"
#include "npz.h"

using namespace hpr;

namespace {

const char *const kDefaultSaveName = "HPRSOCP_autosave.npz";

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Julia semantics for host scalars (max/min propagate NaN, clamp = nested ifelse)
double jmax(double a, double b) {
    if (a != a || b != b) return NAN;
    return a > b ? a : (b > a ? b : (signbit(a) ? b : a));
}
double jmin(double a, double b) {
    if (a != a || b != b) return NAN;
    return a < b ? a : (b < a ? b : (signbit(a) ? a : b));
}
double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

struct Err {
    int code;
    std::string msg;
};

#define HIPCHK(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) throw Err{HPRSOCP_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

std::string g_create_error;

// SplitMix64 + Box-Muller normals (+1e-8): start vector of the power iterations
// (eigenvalues.jl:76,159 use randn(MersenneTwister(1)) + 1e-8).
void start_vector(int64_t len, double *out) {
    uint64_t st = 1;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < len; i += 2) {
        const double u1 = ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        const double u2 = ((double)(next() >> 11)) * (1.0 / 9007199254740992.0);
        const double r = sqrt(-2.0 * log(u1));
        out[i] = r * cos(two_pi * u2) + 1e-8;
        if (i + 1 < len) out[i + 1] = r * sin(two_pi * u2) + 1e-8;
    }
}

// std::allocator that leaves trivially constructible elements uninitialized on
// resize, so large host arrays are first touched (and paged in) by the threads
// that fill them rather than by one zeroing pass.
template <class T>
struct RawAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = RawAlloc<U>;
    };
    RawAlloc() = default;
    template <class U>
    RawAlloc(const RawAlloc<U> &) {}
    template <class U>
    void construct(U *q) noexcept {
        ::new ((void *)q) U;
    }
    template <class U, class... Args>
    void construct(U *q, Args &&...a) {
        ::new ((void *)q) U(std::forward<Args>(a)...);
    }
};
template <class T>
using RawVec = std::vector<T, RawAlloc<T>>;

struct HostCsr {
    int32_t nrows = 0, ncols = 0;
    std::vector<int32_t> rp;
    RawVec<int32_t> ci;
    RawVec<double> v;
};

// fn(i) for every i in [0, count), spread over up to 16 host threads
template <class F>
void host_parallel(int64_t count, F fn) {
    const int nth = (int)std::max<int64_t>(
        1, std::min<int64_t>(std::min<int64_t>(count, 16), (int64_t)std::thread::hardware_concurrency()));
    std::atomic<int64_t> next{0};
    auto work = [&] {
        for (int64_t i; (i = next.fetch_add(1)) < count;) fn(i);
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < nth; ++w) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
}

// dst = src[0, n), copied in 1 M-element pieces on several threads
template <class T>
void par_assign(RawVec<T> &dst, const T *src, int64_t n) {
    dst.resize((size_t)n);
    const int64_t piece = 1 << 20;
    host_parallel((n + piece - 1) / piece, [&](int64_t b) {
        const int64_t lo = b * piece, hi = std::min(n, lo + piece);
        memcpy(dst.data() + lo, src + lo, (size_t)(hi - lo) * sizeof(T));
    });
}

// First column of every row of `a` on the device when each row's columns are
// consecutive (CsrT::c0), null otherwise.
int32_t *contig_c0(hprsocp_handle *h, const HostCsr &a);

// Counting-sort transpose.  Large matrices fill in parallel: each thread owns
// a range of output rows (input columns) holding about the same number of
// entries and scans the input rows in order, so every output row comes out
// sorted by column, exactly as in the sequential fill.
HostCsr transpose(const HostCsr &a) {
    // Row ranges of about equal nnz go to threads; each counts its columns, the
    // counts turn into per-(thread, column) start positions, and each thread
    // scatters its rows in order, so every column lists its rows ascending.
    // Threads are limited so the count table stays within nnz entries.
    HostCsr t;
    t.nrows = a.ncols;
    t.ncols = a.nrows;
    t.rp.assign((size_t)a.ncols + 1, 0);
    const int64_t nnz = (int64_t)a.ci.size();
    t.ci.resize((size_t)nnz);
    t.v.resize((size_t)nnz);
    const int64_t nc = a.ncols;
    const int nth = nnz < (1 << 16) ? 1
                                    : (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(16, nnz / std::max<int64_t>(nc, 1)),
                                                                                  std::thread::hardware_concurrency()));
    std::vector<int32_t> cut((size_t)nth + 1, a.nrows);
    cut[0] = 0;
    for (int w = 1; w < nth; ++w)  // first row whose entries start at or after w/nth of nnz
        cut[w] = (int32_t)(std::lower_bound(a.rp.begin(), a.rp.end() - 1, (int32_t)(nnz * w / nth)) - a.rp.begin());
    RawVec<int32_t> cnt((size_t)nth * nc);
    host_parallel(nth, [&](int64_t w) {
        int32_t *c = cnt.data() + w * nc;
        std::fill(c, c + nc, 0);
        const int32_t r1 = std::max(cut[w], cut[w + 1]);
        for (int32_t k = a.rp[cut[w]]; k < a.rp[r1]; ++k) c[a.ci[k]]++;
    });
    const int64_t piece = 1 << 16, pieces = (nc + piece - 1) / piece;
    host_parallel(pieces, [&](int64_t b) {  // column totals; counts -> offsets within the column
        for (int64_t j = b * piece; j < std::min(nc, (b + 1) * piece); ++j) {
            int32_t run = 0;
            for (int w = 0; w < nth; ++w) {
                const int32_t k = cnt[(size_t)w * nc + j];
                cnt[(size_t)w * nc + j] = run;
                run += k;
            }
            t.rp[j + 1] = run;
        }
    });
    for (int64_t j = 0; j < nc; ++j) t.rp[j + 1] += t.rp[j];
    host_parallel(pieces, [&](int64_t b) {
        for (int64_t j = b * piece; j < std::min(nc, (b + 1) * piece); ++j)
            for (int w = 0; w < nth; ++w) cnt[(size_t)w * nc + j] += t.rp[j];
    });
    host_parallel(nth, [&](int64_t w) {
        int32_t *next = cnt.data() + w * nc;
        const int32_t r1 = std::max(cut[w], cut[w + 1]);
        for (int32_t i = cut[w]; i < r1; ++i)
            for (int32_t k = a.rp[i]; k < a.rp[i + 1]; ++k) {
                const int32_t d = next[a.ci[k]]++;
                t.ci[d] = i;
                t.v[d] = a.v[k];
            }
    });
    return t;
}

// rows sorted by column, duplicates summed
void canonicalize(HostCsr &a) {
    std::vector<int32_t> rp(1, 0);
    RawVec<int32_t> ci;
    RawVec<double> v;
    std::vector<std::pair<int32_t, double>> row;
    for (int32_t i = 0; i < a.nrows; ++i) {
        row.clear();
        for (int32_t k = a.rp[i]; k < a.rp[i + 1]; ++k) row.push_back({a.ci[k], a.v[k]});
        std::stable_sort(row.begin(), row.end(), [](auto &x, auto &y) { return x.first < y.first; });
        for (size_t k = 0; k < row.size(); ++k) {
            if (!ci.empty() && (int32_t)ci.size() > rp.back() && ci.back() == row[k].first) v.back() += row[k].second;
            else {
                ci.push_back(row[k].first);
                v.push_back(row[k].second);
            }
        }
        rp.push_back((int32_t)ci.size());
    }
    a.rp.swap(rp);
    a.ci.swap(ci);
    a.v.swap(v);
}

}  // namespace

// ------------------------------------------------------------------ handle
// Vectors touched by the update kernels.  STATE vectors carry the iteration
// (and its check outputs) and travel both ways between FP64 and the FP32
// mirrors; DATA vectors (problem data and sigma-dependent factors) are only
// demoted.
#define HPR_STATE_N(X) X(x) X(x_hat) X(x_bar) X(z_bar) X(dx) X(last_x) X(w) X(w_bar) X(last_w) X(dw) \
    X(aty) X(aty_bar) X(last_aty)
#define HPR_STATE_M(X) X(y) X(y_bar) X(dy) X(last_y) X(s)
#define HPR_DATA_N(X) X(c) X(l) X(u) X(diag_Q) X(fact1) X(fact2) X(fact)
#define HPR_DATA_M(X) X(AL) X(AU) X(rhs)
#define HPR_SCRATCH_N(X) X(atdy) X(qw) X(tempv) X(raw_n)
#define HPR_SCRATCH_M(X) X(raw_m)

// Matrices and vectors one precision of the update runs on.
template <typename T>
struct IterView {
    CsrT<T> A{}, AT{}, Q{}, F{}, FT{};
    T *Fdiag = nullptr, *Ftmp = nullptr;
#define HPR_FIELD(f) T *f = nullptr;
    HPR_STATE_N(HPR_FIELD) HPR_STATE_M(HPR_FIELD) HPR_DATA_N(HPR_FIELD) HPR_DATA_M(HPR_FIELD)
    HPR_SCRATCH_N(HPR_FIELD) HPR_SCRATCH_M(HPR_FIELD)
#undef HPR_FIELD
};
// One problem instance inside the block-diagonal layout (half-open ranges).
// Rows: linear rows of all instances first, then their SOC rows; columns:
// boxed columns of all instances, then their SOC columns (the canonical order
// of a single problem, applied to the block-diagonal union).
struct InstRange {
    int32_t lin0, lin1, con0, con1;  // rows
    int32_t box0, box1, var0, var1;  // columns
    int32_t kc0, kc1, kv0, kv1;      // constraint / variable cones (indices into con_idx / var_idx)
    int32_t m() const { return (lin1 - lin0) + (con1 - con0); }
    int32_t n() const { return (box1 - box0) + (var1 - var0); }
};

// Host-side scalars of one instance.
struct InstHost {
    double norm_b = 0, norm_c = 0, norm_b_org = 0, norm_c_org = 0, obj_const = 0;
    double lambda_Q_diag = 0;  // max Q entry (diagonal Q), from check_Q_diagonal
    int32_t Q_is_diag = 0;
};

struct SolveLoop;
void free_loop(SolveLoop *s);

struct hprsocp_handle {
    std::string err;
    hprsocp_params params{};
    hipStream_t st = nullptr;
    int device = 0;
    int32_t m = 0, n = 0, m1 = 0, n1 = 0, ncon = 0, nvar = 0;
    int64_t nnzA = 0, nnzQ = 0;
    std::vector<void *> allocs;
    // instances (1 for a single problem)
    int32_t nb = 1;
    std::vector<InstRange> rng;
    std::vector<InstHost> ih;
    InstScal *isc = nullptr;          // device, nb
    InstScal *isc_h = nullptr;        // pinned host mirror
    SegPlan seg_n, seg_m;             // per-instance chunks of columns / rows
    std::vector<int32_t> seg_n_first, seg_m_first;  // host copies of their first-chunk tables
    int32_t max_chunks = 1;
    int32_t *cinst_con = nullptr, *cinst_var = nullptr;  // instance of each cone
    int32_t *ccnt_con = nullptr, *ccnt_var = nullptr;    // cone count of that instance
    int32_t *flag = nullptr, *flag_h = nullptr;          // per-instance flags (restart copies)
    std::vector<int32_t> h_row_inst, h_col_inst;
    // matrices
    CsrView A{}, AT{}, Q{}, F{}, FT{};
    double *A_v = nullptr, *AT_v = nullptr, *Q_v = nullptr, *Q_v2 = nullptr;
    int32_t *Q_tpos = nullptr;
    int q_kind = HPRSOCP_Q_CSR, has_quad = 0, Q_is_diag = 0, q_struct_diag = 0;
    // factored Q and no variable cones: the primal step runs in the F^T product's
    // tile kernel (launch_qp_primal_fq); HPRSOCP_FUSE_QW=0 keeps them apart
    int fuse_qw = 0;
    int32_t Frows = 0;
    double *Fdiag = nullptr, *Ftmp = nullptr;
    DevPlan plan_A, plan_AT, plan_Q, plan_F, plan_FT;
    // batches: host tiles of A, A^T, Q and the compacted (active-instance) plans
    // the update launches once instances retire
    HostPlan hp_A, hp_AT, hp_Q;
    DevPlan act_A, act_AT, act_Q;
    DevPlan act32_A, act32_AT, act32_Q;  // the same for the instances running in FP32
    bool use_act = false;
    int32_t *con_idx = nullptr, *var_idx = nullptr, *row_cone = nullptr;
    std::vector<int32_t> h_con_idx, h_var_idx;
    std::vector<int32_t> h_A_rp;  // for nnz bookkeeping
    // vectors
    double *c = nullptr, *AL = nullptr, *AU = nullptr, *l = nullptr, *u = nullptr, *rhs = nullptr;
    double *row_norm = nullptr, *col_norm = nullptr, *diag_Q = nullptr;
    double *x = nullptr, *x_hat = nullptr, *x_bar = nullptr, *z_bar = nullptr, *dx = nullptr, *last_x = nullptr;
    double *y = nullptr, *y_bar = nullptr, *dy = nullptr, *last_y = nullptr, *s = nullptr;
    double *w = nullptr, *w_bar = nullptr, *last_w = nullptr, *dw = nullptr;
    double *aty = nullptr, *aty_bar = nullptr, *last_aty = nullptr, *atdy = nullptr, *qatdy = nullptr;
    double *qw = nullptr, *dqw = nullptr, *qx = nullptr, *tempv = nullptr, *ax = nullptr, *rp = nullptr, *rd = nullptr;
    double *fact1 = nullptr, *fact2 = nullptr, *fact = nullptr, *fact_m = nullptr;
    double *raw_n = nullptr, *raw_m = nullptr, *part_n = nullptr, *part_m = nullptr;
    double *cone_t = nullptr, *cone_ns = nullptr;
    ConePieces cone_pieces;  // constraint cones cut for the residual's cone statistics
    double *red_part = nullptr;  // reduction partials, 3 regions of red_region doubles
    int64_t red_region = 0;
    double *dscal = nullptr;     // device scalar slots, nb x SL_COUNT
    double *hscal = nullptr;     // pinned host mirror
    double *ivtmp = nullptr, *ivtmp_h = nullptr;  // per-instance scratch values (2 x nb)
    PowerState *pstate = nullptr;
    double *tmp_m = nullptr, *tmp_n = nullptr, *tmp_n2 = nullptr;
    // auto_save: checkpoint path, the initial points given in the parameters,
    // and the best iterate of the solve (scaled, device) with its residuals
    std::string save_path;
    std::vector<double> init_x, init_y;
    double *sv_x = nullptr, *sv_y = nullptr, *sv_z = nullptr, *sv_w = nullptr;
    struct Saved {
        double sigma = 0, err_Rp = INFINITY, err_Rd = INFINITY, primal_obj = INFINITY, dual_obj = INFINITY,
               rel_gap = INFINITY;
        int64_t iter = 0;
    } saved;
    bool lambda_set = false;
    double power_time = 0, setup_time = 0;
    int to_check = 1;
    int64_t kstep = 0;  // updates since InstScal::inner0 was set
    // FP32 mirrors of the iteration (precision FP32_REFINE)
    bool fp32 = false;         // mirrors allocated
    int32_t n32 = 0, n64 = 0;  // active instances whose updates run in FP32 / FP64
    bool fp32_dirty = true;    // FP64 state changed since the last demotion
    int64_t nnzF = 0;
    IterView<float> v32{};
    // precision FP32_DELTA (single problem): the FP32 mirrors of x, y, w, aty
    // hold differences from last_x, last_y, last_w, last_aty; ra and rya hold
    // the anchors' images (see DPrimalArgs)
    bool delta = false;
    bool delta_live = false;    // the deltas hold the current state (the FP64 x, y, w, aty are stale)
    bool images_valid = false;  // ra, rya match the anchors, sigma and lambda_A
    double *d_ra = nullptr, *d_rya = nullptr, *d_img_n = nullptr, *d_img_m = nullptr;
    // the solve loop in progress (hprsocp_iterate)
    SolveLoop *loop = nullptr;
    // device controller of batch solves (window-edge steps queued on the stream)
    bool dctl = false;
    ctl::CtlInst *ci = nullptr;        // device, nb
    ctl::InstConst *ic = nullptr;      // device, nb
    int32_t *ctl_status = nullptr;     // device, nb
    int32_t *noflag = nullptr;         // device, nb: !flag
    GoldenWalk *gwalk = nullptr;       // device, nb: golden-section walks of the diagonal-Q sigma update
    int32_t *gnp = nullptr;            // device, nb: points of each walk's current pass
    int32_t *ctl_status_h = nullptr;   // pinned, 2 x nb (one per event in flight)
    hipEvent_t ctl_ev[2] = {nullptr, nullptr};
    Tile *stage_h[3] = {nullptr, nullptr, nullptr};  // pinned staging of compacted tile lists
    hipEvent_t stage_ev = nullptr;
    bool stage_pending = false;
    std::vector<hipGraphExec_t> graphs_retired;      // replaced while possibly still running
    // launch-plan choices of the autotuner, one record per operator
    struct TuneRecord {
        std::string name, form = "plain";
        PlanConfig cfg;
        bool tuned = false, cached = false;
        int32_t candidates = 0, tiles = 0;
        float ms_default = 0, ms = 0;    // plain form: default / chosen config
        float ms_form = 0;               // the chosen form (sliced or plain)
        int32_t block = 0;               // sorted form: threads of the sorted-slice kernel (FP64)
        std::string form32;              // FP32 copy: "own" sorted layout, "fp64" layout, or "" (none)
        int32_t block32 = 0;             // FP32 copy: threads of its sorted-slice kernel
        float ms32 = 0;                  // FP32 copy: the chosen layout and block
    };
    std::vector<TuneRecord> tune;
    // timing (profiled iterations)
    bool profile = false;
    std::vector<hipEvent_t> ev;
    // captured runs of partial updates: (first kstep, length, fp32) -> executable graph
    std::map<std::tuple<int64_t, int32_t, int>, hipGraphExec_t> graphs;
    // batches: windows requested since the last compaction, not yet captured
    std::map<std::tuple<int64_t, int32_t, int>, int> graph_seen;
    void clear_graphs() {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
        graph_seen.clear();
    }
    // drops the captured windows while the stream may still run one of them;
    // they are destroyed after the next full synchronisation (free_retired_graphs)
    void retire_graphs() {
        for (auto &kv : graphs) graphs_retired.push_back(kv.second);
        graphs.clear();
        graph_seen.clear();
    }
    void free_retired_graphs() {
        for (auto g : graphs_retired) (void)hipGraphExecDestroy(g);
        graphs_retired.clear();
    }

    InstScal &I(int i) { return isc_h[i]; }
    template <class T>
    T *dalloc(int64_t count) {
        void *p = nullptr;
        const size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) throw Err{HPRSOCP_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
        allocs.push_back(p);
        HIPCHK(hipMemsetAsync(p, 0, bytes, st));
        return (T *)p;
    }
    // frees an allocation made by dalloc/upload
    void dfree(const void *p) {
        if (!p) return;
        for (size_t k = 0; k < allocs.size(); ++k)
            if (allocs[k] == p) {
                (void)hipStreamSynchronize(st);
                (void)hipFree(allocs[k]);
                allocs.erase(allocs.begin() + (int64_t)k);
                return;
            }
    }
    // synchronous: the host source may be a temporary
    template <class T>
    T *upload(const T *src, int64_t count) {
        T *d = dalloc<T>(count);
        if (count > 0 && src) HIPCHK(hipMemcpyAsync(d, src, (size_t)count * sizeof(T), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        return d;
    }
    void sync() { HIPCHK(hipStreamSynchronize(st)); }
    ~hprsocp_handle() {
        if (st) (void)hipStreamSynchronize(st);
        free_loop(loop);
        clear_graphs();
        free_retired_graphs();
        if (ctl_status_h) (void)hipHostFree(ctl_status_h);
        for (auto &e : ctl_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &b : stage_h)
            if (b) (void)hipHostFree(b);
        if (stage_ev) (void)hipEventDestroy(stage_ev);
        for (void *p : allocs) (void)hipFree(p);
        if (hscal) (void)hipHostFree(hscal);
        if (isc_h) (void)hipHostFree(isc_h);
        if (flag_h) (void)hipHostFree(flag_h);
        if (ivtmp_h) (void)hipHostFree(ivtmp_h);
        for (auto e : ev) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};

// device scalar slots
enum : int {
    SL_RN = 0,      // residual n-side (6)
    SL_RM = 8,      // residual m-side (5)
    SL_BV = 16,     // bounds violation (1)
    SL_MN0 = 24,    // M-norm dots, slot "last gap" (8)
    SL_MN1 = 32,    // M-norm dots, slot "current gap" (8)
    SL_SIG = 40,    // sigma dots (8)
    SL_PW = 48,     // power iteration scratch (4)
    SL_GEN = 56,    // generic (8)
    SL_GOLD = 64,   // golden section evaluations (64)
    SL_COUNT = 128
};

namespace {

void upload_plan(hprsocp_handle *h, const HostPlan &hp, DevPlan &dp) {
    dp.ntiles = (int32_t)hp.tiles.size();
    dp.first_huge = hp.first_huge;
    dp.nslots = hp.nslots;
    dp.nhuge = (int32_t)hp.hstart.size();
    dp.tiles = h->upload<Tile>(hp.tiles.data(), (int64_t)hp.tiles.size());
    const int64_t nh = dp.nhuge;
    std::vector<int32_t> buf;
    buf.reserve((size_t)nh * 4);
    buf.insert(buf.end(), hp.hstart.begin(), hp.hstart.end());
    buf.insert(buf.end(), hp.hend.begin(), hp.hend.end());
    buf.insert(buf.end(), hp.hslot0.begin(), hp.hslot0.end());
    buf.insert(buf.end(), hp.hnslot.begin(), hp.hnslot.end());
    dp.huge_buf = h->upload<int32_t>(buf.data(), (int64_t)buf.size());
    dp.huge.start = dp.huge_buf;
    dp.huge.end = dp.huge_buf + nh;
    dp.huge.slot0 = dp.huge_buf + 2 * nh;
    dp.huge.nslot = dp.huge_buf + 3 * nh;
}

DevPlan make_plan(hprsocp_handle *h, const std::vector<int32_t> &rp, int32_t nrows, int32_t lin_end,
                  const std::vector<int32_t> &cidx, int32_t ncones, const std::vector<int32_t> *row_inst = nullptr,
                  HostPlan *keep = nullptr, const PlanConfig &cfg = PlanConfig{}) {
    const int32_t *ri = (row_inst && h->nb > 1) ? row_inst->data() : nullptr;
    HostPlan hp = build_plan(rp.empty() ? nullptr : rp.data(), nrows, lin_end, cidx.empty() ? nullptr : cidx.data(),
                             ncones, cfg, ri);
    DevPlan dp;
    upload_plan(h, hp, dp);
    if (keep) *keep = hp;
    return dp;
}

// A second tile buffer of the plan, filled by compact_plan.
DevPlan alloc_active(hprsocp_handle *h, const DevPlan &full) {
    DevPlan a = full;
    a.tiles = h->dalloc<Tile>(std::max(full.ntiles, 1));
    return a;
}

// Keeps the tiles of the instances with keep[inst] (in order; HUGE segments stay last).
void compact_plan(hprsocp_handle *h, const HostPlan &hp, DevPlan &act, const std::vector<char> &keep) {
    std::vector<Tile> t;
    t.reserve(hp.tiles.size());
    int32_t first_huge = 0;
    for (size_t k = 0; k < hp.tiles.size(); ++k) {
        const Tile &x = hp.tiles[k];
        if (!keep[x.inst]) continue;
        t.push_back(x);
        if ((int32_t)k < hp.first_huge) ++first_huge;
    }
    act.ntiles = (int32_t)t.size();
    act.first_huge = first_huge;
    if (!t.empty())
        HIPCHK(hipMemcpyAsync(act.tiles, t.data(), sizeof(Tile) * t.size(), hipMemcpyHostToDevice, h->st));
    h->sync();  // t is a host temporary
}

// compact_plan without a synchronisation: the kept tiles go through pinned
// staging buffer `which` (one per plan) and an in-stream copy, so the running
// window still reads the old list and the next window the new one.  A staging
// buffer is rewritten only after the previous staged copies have landed.
void compact_plan_staged(hprsocp_handle *h, const HostPlan &hp, DevPlan &act, const std::vector<char> &keep,
                         int which) {
    if (!h->stage_h[which])
        HIPCHK(hipHostMalloc((void **)&h->stage_h[which], sizeof(Tile) * std::max<size_t>(hp.tiles.size(), 1)));
    if (h->stage_pending) {
        HIPCHK(hipEventSynchronize(h->stage_ev));
        h->stage_pending = false;
    }
    Tile *t = h->stage_h[which];
    int32_t n = 0, first_huge = 0;
    for (size_t k = 0; k < hp.tiles.size(); ++k) {
        const Tile &x = hp.tiles[k];
        if (!keep[x.inst]) continue;
        t[n++] = x;
        if ((int32_t)k < hp.first_huge) ++first_huge;
    }
    act.ntiles = n;
    act.first_huge = first_huge;
    if (n > 0) HIPCHK(hipMemcpyAsync(act.tiles, t, sizeof(Tile) * n, hipMemcpyHostToDevice, h->st));
}

double *slot(hprsocp_handle *h, int k) { return h->dscal + k; }
// slot k of instance i on the device / in the host mirror after read_slots
InstVal islot(hprsocp_handle *h, int k) { return InstVal{h->dscal + k, SL_COUNT}; }
double hs(hprsocp_handle *h, int i, int k) { return h->hscal[(int64_t)i * SL_COUNT + k]; }

// copy every instance's device slots to the host mirror and wait
void read_slots(hprsocp_handle *h) {
    HIPCHK(hipMemcpyAsync(h->hscal, h->dscal, sizeof(double) * SL_COUNT * h->nb, hipMemcpyDeviceToHost, h->st));
    h->sync();
}
// publish the per-instance scalars (synchronous: the pinned mirror may change right after)
void upload_isc(hprsocp_handle *h) {
    HIPCHK(hipMemcpyAsync(h->isc, h->isc_h, sizeof(InstScal) * h->nb, hipMemcpyHostToDevice, h->st));
    h->sync();
    h->n32 = h->n64 = 0;
    for (int i = 0; i < h->nb; ++i)
        if (h->isc_h[i].active) ++(h->isc_h[i].fp32 ? h->n32 : h->n64);
}
// per-instance scratch values: ivtmp[k*nb + i]
InstVal ivals(hprsocp_handle *h, int k, const std::vector<double> &v) {
    memcpy(h->ivtmp_h + (int64_t)k * h->nb, v.data(), sizeof(double) * h->nb);
    HIPCHK(hipMemcpyAsync(h->ivtmp + (int64_t)k * h->nb, h->ivtmp_h + (int64_t)k * h->nb, sizeof(double) * h->nb,
                          hipMemcpyHostToDevice, h->st));
    h->sync();
    return InstVal{h->ivtmp + (int64_t)k * h->nb, 1};
}
InstVal isc_field(hprsocp_handle *h, const double *field0) {
    return InstVal{field0, (int32_t)(sizeof(InstScal) / sizeof(double))};
}

double *region(hprsocp_handle *h, int r) { return h->red_part + (int64_t)r * h->red_region; }

void reduce(hprsocp_handle *h, const SegPlan &sp, std::initializer_list<RedTerm> terms, int out_slot, int reg = 0) {
    std::vector<RedTerm> t(terms);
    launch_reduce(h->st, sp, t.data(), (int)t.size(), region(h, reg), slot(h, out_slot), SL_COUNT);
}

// Q v into out (general Q; Q must have quadratic terms)
void qmap(hprsocp_handle *h, const double *v, double *out) {
    if (h->q_kind == HPRSOCP_Q_FACTORED) {
        launch_spmv(h->st, h->plan_F, h->F, v, h->Ftmp);
        launch_spmv(h->st, h->plan_FT, h->FT, h->Ftmp, out, h->Fdiag, v);
    } else if (h->nnzQ == 0) {
        launch_fill(h->st, h->n, 0.0, out);
    } else if (h->q_struct_diag) {
        launch_mul(h->st, h->n, h->diag_Q, v, out);
    } else {
        launch_spmv(h->st, h->plan_Q, h->Q, v, out);
    }
}

// ------------------------------------------------------------- scaling
void equilibrate(hprsocp_handle *h, int use_sum, int phase, double *trow, double *tcol) {
    hipStream_t st = h->st;
    const CsrView *qv = (h->nnzQ > 0) ? &h->Q : nullptr;
    launch_row_stat(st, h->AT, qv, use_sum, tcol);
    if (h->m > 0) {
        launch_row_stat(st, h->A, nullptr, use_sum, trow);
        launch_block_scale(st, h->ncon, h->con_idx, h->ccnt_con, trow, h->params.soc_block_scaling_strategy, phase,
                           LOC_CONSTRAINT);
        launch_clip(st, h->m, trow);
    }
    launch_block_scale(st, h->nvar, h->var_idx, h->ccnt_var, tcol, h->params.soc_block_scaling_strategy, phase,
                       LOC_VARIABLE);
    launch_clip(st, h->n, tcol);
    launch_scale_rows(st, h->m, trow, h->AL, h->AU, h->rhs, h->row_norm);
    launch_scale_cols(st, h->n, tcol, h->c, h->l, h->u, h->col_norm);
    if (h->nnzQ > 0) launch_scale_csr(st, h->n, h->Q.rowptr, h->Q.col, h->Q_v, tcol, tcol);
    if (h->m > 0) {
        launch_scale_csr(st, h->m, h->A.rowptr, h->A.col, h->A_v, trow, tcol);
        launch_scale_csr_t(st, h->n, h->AT.rowptr, h->AT.col, h->AT_v, trow, tcol);
    }
}

// per instance: l2 and inf norm of max(|AL_nInf|,|AU_nInf|)
void bound_norms(hprsocp_handle *h, std::vector<double> &l2, std::vector<double> &linf) {
    l2.assign(h->nb, 0.0);
    linf.assign(h->nb, 0.0);
    if (h->m == 0) return;
    launch_bound_stats(h->st, h->seg_m, h->AL, h->AU, region(h, 0), slot(h, SL_GEN), SL_COUNT);
    read_slots(h);
    for (int i = 0; i < h->nb; ++i) {
        if (h->rng[i].m() == 0) continue;
        l2[i] = sqrt(hs(h, i, SL_GEN));
        linf[i] = hs(h, i, SL_GEN + 1);
    }
}
// per instance [sumsq, count, absmax] of the finite entries of an m- or n-vector
void finite_stats(hprsocp_handle *h, const SegPlan &sp, const double *v, std::vector<std::array<double, 3>> &out) {
    launch_finite_stats(h->st, sp, v, region(h, 0), slot(h, SL_GEN), SL_COUNT);
    read_slots(h);
    out.resize(h->nb);
    for (int i = 0; i < h->nb; ++i) out[i] = {hs(h, i, SL_GEN), hs(h, i, SL_GEN + 1), hs(h, i, SL_GEN + 2)};
}
// per instance absmax of soc_rhs (zero on linear rows)
std::vector<double> soc_rhs_inf(hprsocp_handle *h) {
    std::vector<double> r(h->nb, 0.0);
    if (h->m == 0) return r;
    reduce(h, h->seg_m, {RedTerm{h->rhs, nullptr, nullptr, 1}}, SL_GEN);
    read_slots(h);
    for (int i = 0; i < h->nb; ++i) r[i] = (h->rng[i].con1 > h->rng[i].con0) ? hs(h, i, SL_GEN) : 0.0;
    return r;
}
std::vector<double> vec_absmax_n(hprsocp_handle *h, const double *v) {
    std::vector<double> r(h->nb, 0.0);
    if (h->n == 0) return r;
    reduce(h, h->seg_n, {RedTerm{v, nullptr, nullptr, 1}}, SL_GEN);
    read_slots(h);
    for (int i = 0; i < h->nb; ++i) r[i] = hs(h, i, SL_GEN);
    return r;
}
std::vector<double> vec_nrm2_n(hprsocp_handle *h, const double *v) {
    std::vector<double> r(h->nb, 0.0);
    if (h->n == 0) return r;
    reduce(h, h->seg_n, {RedTerm{v, v, nullptr, 0}}, SL_GEN);
    read_slots(h);
    for (int i = 0; i < h->nb; ++i) r[i] = sqrt(hs(h, i, SL_GEN));
    return r;
}

// scaling!, reference src/utils/scaling.jl:785-1289 (CPU-path semantics: no
// clipping before the SOC block aggregation; A^T scaled bit-identically to A).
// Row/column equilibration is local to each instance of a block-diagonal
// batch; the b/c scales and all norms are per instance.
void do_scaling(hprsocp_handle *h) {
    hipStream_t st = h->st;
    const int32_t m = h->m, n = h->n, nb = h->nb;
    launch_fill(st, m, 1.0, h->row_norm);
    launch_fill(st, n, 1.0, h->col_norm);
    for (int i = 0; i < nb; ++i) h->I(i).b_scale = h->I(i).c_scale = 1.0;
    std::vector<double> bl2, blinf;
    bound_norms(h, bl2, blinf);
    std::vector<double> rinf = soc_rhs_inf(h);
    std::vector<double> cinf = vec_absmax_n(h, h->c);
    for (int i = 0; i < nb; ++i) {
        const bool hm = h->rng[i].m() > 0;
        h->ih[i].norm_b_org = hm ? jmax(blinf[i], rinf[i]) : 0.0;
        h->ih[i].norm_c_org = cinf[i];
    }
    if (h->q_kind == HPRSOCP_Q_FACTORED) {  // operator Q: skip ALL scaling (scaling.jl:797-846)
        std::vector<double> cn = vec_nrm2_n(h, h->c);
        for (int i = 0; i < nb; ++i) {
            h->ih[i].norm_b = h->rng[i].m() > 0 ? jmax(bl2[i], rinf[i]) : 0.0;
            h->ih[i].norm_c = cn[i];
        }
        upload_isc(h);
        return;
    }
    double *trow = h->tmp_m, *tcol = h->tmp_n;
    launch_fill(st, m, 1.0, trow);
    launch_fill(st, n, 1.0, tcol);
    const hprsocp_params &p = h->params;
    if (p.use_Ruiz_scaling)
        for (int it = 0; it < std::max(p.ruiz_iterations, 0); ++it) equilibrate(h, 0, PHASE_RUIZ, trow, tcol);
    if (uses_pre_pock(p.soc_block_scaling_strategy))
        for (int it = 0; it < 2; ++it) equilibrate(h, 0, PHASE_PRE_POCK, trow, tcol);
    if (p.use_Pock_Chambolle_scaling) equilibrate(h, 1, PHASE_POCK, trow, tcol);
    if (p.use_bc_scaling) {
        // compute_bc_scales, scaling.jl:721-760
        const int t = p.bc_scaling_norm_type;
        std::vector<std::array<double, 3>> fAL, fAU, fr, fc;
        if (m > 0) {
            finite_stats(h, h->seg_m, h->AL, fAL);
            finite_stats(h, h->seg_m, h->AU, fAU);
            finite_stats(h, h->seg_m, h->rhs, fr);
        }
        finite_stats(h, h->seg_n, h->c, fc);
        auto finite_norm = [&](const std::array<double, 3> &f) {
            if (t == HPRSOCP_BC_LINF) return f[2];
            if (f[1] == 0) return 0.0;
            return t == HPRSOCP_BC_RMS ? sqrt(f[0] / f[1]) : sqrt(f[0]);
        };
        auto vec_norm = [&](const std::array<double, 3> &f, int64_t len) {
            if (len == 0) return 0.0;
            if (t == HPRSOCP_BC_LINF) return f[2];
            return t == HPRSOCP_BC_RMS ? sqrt(f[0] / (double)len) : sqrt(f[0]);
        };
        std::vector<double> qmul(nb), bdiv(nb), cdiv(nb);
        for (int i = 0; i < nb; ++i) {
            const InstRange &R = h->rng[i];
            double hb = 1.0;
            if (R.m() > 0) {
                hb = jmax(hb, finite_norm(fAL[i]));
                hb = jmax(hb, finite_norm(fAU[i]));
                const int32_t ncr = R.con1 - R.con0;
                hb = jmax(hb, (R.kc1 == R.kc0 || ncr == 0) ? 0.0 : vec_norm(fr[i], ncr));
            }
            hb = jclamp(hb, 1e-4, 1e4);
            const double hc = jclamp(jmax(1.0, vec_norm(fc[i], R.n())), 1e-4, 1e4);
            const double bs = hb, cs = jclamp(hc, hb / 1e2, hb / 1e-2);
            h->I(i).b_scale = bs;
            h->I(i).c_scale = cs;
            qmul[i] = bs / cs;
            bdiv[i] = bs;
            cdiv[i] = cs;
        }
        const InstVal one = ivals(h, 0, std::vector<double>(nb, 1.0));
        if (h->nnzQ > 0) {
            const InstVal qm = ivals(h, 1, qmul);
            launch_scale_rows_inst(st, h->seg_n, h->Q.rowptr, h->Q_v, qm);
        }
        const InstVal bd = ivals(h, 1, bdiv);
        if (m > 0) {
            launch_scale_scalar(st, h->seg_m, h->AL, one, bd);
            launch_scale_scalar(st, h->seg_m, h->AU, one, bd);
            launch_scale_scalar(st, h->seg_m, h->rhs, one, bd);
        }
        launch_scale_scalar(st, h->seg_n, h->l, one, bd);
        launch_scale_scalar(st, h->seg_n, h->u, one, bd);
        const InstVal cd = ivals(h, 1, cdiv);  // (synchronises: the previous uses of slot 1 are done)
        launch_scale_scalar(st, h->seg_n, h->c, one, cd);
        h->sync();
    }
    // Q = (Q + Q')/2 (scaling.jl:1265-1269)
    if (h->nnzQ > 0) {
        launch_sym_q(st, h->nnzQ, h->Q_v, h->Q_tpos, h->Q_v2);
        launch_copy(st, h->nnzQ, h->Q_v2, h->Q_v);
    }
    bound_norms(h, bl2, blinf);
    rinf = soc_rhs_inf(h);
    std::vector<double> cn = vec_nrm2_n(h, h->c);
    for (int i = 0; i < nb; ++i) {
        h->ih[i].norm_b = jmax(bl2[i], rinf[i]);
        h->ih[i].norm_c = cn[i];
    }
    upload_isc(h);
}

void update_q_factors(hprsocp_handle *h) {
    launch_q_factors(h->st, h->seg_n, h->diag_Q, h->isc, h->fact2, h->fact, h->fact1, h->fact_m);
}

// ------------------------------------------------------------- power iteration
// power_iteration_A/Q (eigenvalues.jl:68-196) for every instance at once; each
// instance starts from its own start vector and stops on its own test.
std::vector<double> power_iteration(hprsocp_handle *h, bool onA) {
    const int nb = h->nb;
    const int64_t len = onA ? h->m : h->n;
    const SegPlan &sp = onA ? h->seg_m : h->seg_n;
    std::vector<double> z0((size_t)len, 0.0);
    for (int i = 0; i < nb; ++i) {
        const InstRange &R = h->rng[i];
        const int32_t a0 = onA ? R.lin0 : R.box0, a1 = onA ? R.lin1 : R.box1;
        const int32_t b0 = onA ? R.con0 : R.var0, b1 = onA ? R.con1 : R.var1;
        const int64_t li = (int64_t)(a1 - a0) + (b1 - b0);
        std::vector<double> zi((size_t)li);
        start_vector(li, zi.data());
        std::copy(zi.begin(), zi.begin() + (a1 - a0), z0.begin() + a0);
        std::copy(zi.begin() + (a1 - a0), zi.end(), z0.begin() + b0);
    }
    double *z = onA ? h->tmp_m : h->tmp_n;
    double *q = onA ? h->ax : h->tmp_n2;
    HIPCHK(hipMemcpyAsync(z, z0.data(), sizeof(double) * len, hipMemcpyHostToDevice, h->st));
    std::vector<PowerState> ps((size_t)nb, PowerState{1.0, 1.0, 0, 0});
    for (int i = 0; i < nb; ++i) {
        const InstRange &R = h->rng[i];
        if ((onA ? R.m() : R.n()) == 0) ps[i].done = 1, ps[i].lambda = 0.0;
    }
    HIPCHK(hipMemcpyAsync(h->pstate, ps.data(), sizeof(PowerState) * nb, hipMemcpyHostToDevice, h->st));
    h->sync();
    const int max_iter = 5000, batch = 10;
    bool cpt = false;  // compacted plans in use
    int done_seen = 0;
    for (int it = 0; it < max_iter; it += batch) {
        for (int k = 0; k < batch; ++k) {
            reduce(h, sp, {RedTerm{z, z, nullptr, 0}}, SL_PW);
            launch_scale_by_rnorm(h->st, sp, z, islot(h, SL_PW), q, h->pstate);
            if (onA) {
                launch_spmv(h->st, cpt ? h->act_AT : h->plan_AT, h->AT, q, h->tempv);
                launch_spmv(h->st, cpt ? h->act_A : h->plan_A, h->A, h->tempv, z);
            } else if (cpt && h->q_kind == HPRSOCP_Q_CSR && h->nnzQ > 0 && !h->q_struct_diag) {
                launch_spmv(h->st, h->act_Q, h->Q, q, z);
            } else {
                qmap(h, q, z);
            }
            reduce(h, sp, {RedTerm{q, z, nullptr, 0}}, SL_PW + 1);
            launch_power_residual(h->st, sp, z, islot(h, SL_PW + 1), q, h->pstate);
            reduce(h, sp, {RedTerm{q, q, nullptr, 0}, RedTerm{z, z, nullptr, 0}}, SL_PW + 2);
            launch_power_check(h->st, nb, islot(h, SL_PW + 1), islot(h, SL_PW + 2), islot(h, SL_PW + 3), h->pstate,
                               1e-4, max_iter);
        }
        HIPCHK(hipMemcpyAsync(ps.data(), h->pstate, sizeof(PowerState) * nb, hipMemcpyDeviceToHost, h->st));
        h->sync();
        bool all = true;
        int ndone = 0;
        for (int i = 0; i < nb; ++i) {
            all = all && ps[i].done;
            ndone += ps[i].done != 0;
        }
        if (all) break;
        // batches: once instances converge, run the SpMVs over the others' tiles only
        if (nb > 1 && ndone > done_seen) {
            std::vector<char> keep(nb);
            for (int i = 0; i < nb; ++i) keep[i] = !ps[i].done;
            compact_plan(h, h->hp_A, h->act_A, keep);
            compact_plan(h, h->hp_AT, h->act_AT, keep);
            if (h->nnzQ > 0) compact_plan(h, h->hp_Q, h->act_Q, keep);
            cpt = true;
            done_seen = ndone;
        }
    }
    std::vector<double> lam((size_t)nb);
    for (int i = 0; i < nb; ++i) lam[i] = ps[i].lambda;
    return lam;
}

// ------------------------------------------------------------- XCD slicing
// One column slice per XCD (8 on MI355X) lets each XCD's 4 MB L2 keep its
// slice of the gathered vector resident.

// Regroups the entries of M slice-major (slice p = columns [p*C/P, (p+1)*C/P)),
// fills the device copy of the (already scaled) values through a permutation,
// and builds the slice row tiles.  `sorted` selects the column-sorted layout of
// k_spmv_sorted: tiles of <= kSortRows rows whose entries in each slice fill
// about one chunk of kSortChunk, each chunk stored in column order as packed
// (column offset, row-major position) words.  `fused_tiles` > 0 (sorted only)
// builds the layout of the slice-fused kernel instead: at most fused_tiles
// contiguous row tiles of about equal nonzero counts (one workgroup each walks
// all slices of its tile), cut earlier where a tile would exceed the row or
// per-slice chunk limits.  Row sums do not depend on the tiling: every row's
// products are summed slice by slice in row-major order either way.
// `round_tiles` > 0 (sorted, unfused): cut the tiles nnz-balanced into the
// next multiple of round_tiles at or above the natural count, so that the
// P workgroups per tile fill whole rounds of the device's workgroup slots
// instead of ending on a partial round.
template <typename T>
bool build_slices(hprsocp_handle *h, CsrT<T> &M, const HostCsr &hm, const T *dval, int32_t P, bool sorted,
                  int32_t fused_tiles = 0, int32_t round_tiles = 0) {
    using Spec = SortSpec<T>;
    if (P <= 1 || hm.ci.empty()) return false;
    const int32_t R = hm.nrows;
    const int64_t nnz = (int64_t)hm.ci.size();
    for (int32_t i = 0; i < R; ++i)
        if (hm.rp[i + 1] - hm.rp[i] > kSliceTileNnz) return false;  // per-slice offsets must fit uint16
    std::vector<int64_t> bound(P + 1);
    for (int p = 0; p <= P; ++p) bound[p] = (int64_t)hm.ncols * p / P;
    if (sorted)
        for (int p = 0; p < P; ++p)
            if (bound[p + 1] - bound[p] > (1ll << Spec::col_bits)) return false;  // column offsets must fit
    auto slice_of = [&](int32_t c) {
        int p = (int)((int64_t)c * P / std::max<int64_t>(hm.ncols, 1));
        while (p + 1 < P && c >= bound[p + 1]) ++p;
        while (p > 0 && c < bound[p]) --p;
        return p;
    };
    std::vector<int32_t> cnt((size_t)P * R, 0);
    for (int32_t i = 0; i < R; ++i)
        for (int32_t k = hm.rp[i]; k < hm.rp[i + 1]; ++k) cnt[(size_t)slice_of(hm.ci[k]) * R + i]++;
    std::vector<SliceTile> tiles;
    if (!sorted) {
        // <= kSliceRows rows, about P * kStreamNnz nonzeros, <= kSliceTileNnz
        const int64_t cap = std::min<int64_t>((int64_t)P * kStreamNnz, kSliceTileNnz);
        int32_t r = 0;
        while (r < R) {
            int32_t e = r + 1;
            int64_t nz = hm.rp[e] - hm.rp[r];
            while (e < R && e - r < kSliceRows && nz + (hm.rp[e + 1] - hm.rp[e]) <= cap) {
                nz += hm.rp[e + 1] - hm.rp[e];
                ++e;
            }
            tiles.push_back(SliceTile{r, e});
            r = e;
        }
    } else {
        // <= Spec::rows rows, <= Spec::chunk entries in every slice (a single
        // longer row forms a tile of its own; the kernel then runs several chunks);
        // fused layout: also end a tile where the running nonzero count reaches
        // the next multiple of nnz / fused_tiles
        std::vector<int64_t> tc((size_t)P);
        auto cut = [&](int32_t target) {
            tiles.clear();
            const double per_tile = target > 0 ? (double)nnz / target : 0.0;
            int32_t r = 0;
            while (r < R) {
                for (int p = 0; p < P; ++p) tc[p] = cnt[(size_t)p * R + r];
                int64_t goal = INT64_MAX;
                if (target > 0) {
                    const int64_t k = (int64_t)std::floor((double)hm.rp[r] / per_tile) + 1;
                    goal = std::max<int64_t>((int64_t)std::llround(k * per_tile), (int64_t)hm.rp[r] + 1);
                }
                int32_t e = r + 1;
                while (e < R && e - r < Spec::rows && hm.rp[e] < goal) {
                    bool fits = true;
                    for (int p = 0; p < P && fits; ++p) fits = tc[p] + cnt[(size_t)p * R + e] <= Spec::chunk;
                    if (!fits) break;
                    for (int p = 0; p < P; ++p) tc[p] += cnt[(size_t)p * R + e];
                    ++e;
                }
                tiles.push_back(SliceTile{r, e});
                r = e;
            }
        };
        cut(fused_tiles);
        // the fused kernel's tiles must not outnumber its workgroup slots (a
        // second round of a few workgroups doubles the time): lower the target
        // by the excess until they fit, or give the layout up
        for (int32_t target = fused_tiles; fused_tiles > 0 && (int32_t)tiles.size() > fused_tiles;) {
            target -= (int32_t)tiles.size() - fused_tiles;
            if (target < 1) return false;
            cut(target);
        }
        if (fused_tiles == 0 && round_tiles > 0) {
            const int32_t natural = (int32_t)tiles.size();
            const int32_t want = (natural + round_tiles - 1) / round_tiles * round_tiles;
            if (want > natural) {
                cut(want);
                for (int32_t target = want; (int32_t)tiles.size() > want;) {
                    target -= (int32_t)tiles.size() - want;
                    if (target < 1) {
                        cut(0);  // keep the natural tiling
                        break;
                    }
                    cut(target);
                }
            }
        }
    }
    const int32_t NT = (int32_t)tiles.size();
    // slice-major entry order: slice p, then tile, then row
    std::vector<int32_t> base((size_t)P * (NT + 1));
    std::vector<uint16_t> off((size_t)P * R);
    std::vector<int64_t> rowstart((size_t)P * R);
    int64_t pos = 0;
    for (int p = 0; p < P; ++p) {
        for (int32_t t = 0; t < NT; ++t) {
            base[(size_t)p * (NT + 1) + t] = (int32_t)pos;
            int64_t o = 0;
            for (int32_t i = tiles[t].rb; i < tiles[t].re; ++i) {
                off[(size_t)p * R + i] = (uint16_t)o;
                rowstart[(size_t)p * R + i] = pos + o;
                o += cnt[(size_t)p * R + i];
            }
            pos += o;
        }
        base[(size_t)p * (NT + 1) + NT] = (int32_t)pos;
    }
    std::vector<int32_t> scol((size_t)nnz), perm((size_t)nnz);
    for (int32_t i = 0; i < R; ++i)
        for (int32_t k = hm.rp[i]; k < hm.rp[i + 1]; ++k) {
            const int p = slice_of(hm.ci[k]);
            const int64_t d = rowstart[(size_t)p * R + i]++;
            scol[d] = hm.ci[k];
            perm[d] = k;
        }
    if (sorted) {
        // each chunk: packed words sorted ascending (= by column, then by
        // row-major position), the value permutation reordered alongside
        const int64_t items = (int64_t)P * NT;
        const int nth = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(16, items),
                                                                    std::thread::hardware_concurrency()));
        auto work = [&](int w) {
            std::vector<uint32_t> key;
            std::vector<int32_t> pv;
            for (int64_t it = w; it < items; it += nth) {
                const int p = (int)(it / NT);
                const int32_t t = (int32_t)(it % NT);
                const int64_t b0 = base[(size_t)p * (NT + 1) + t], b1 = base[(size_t)p * (NT + 1) + t + 1];
                for (int64_t c0 = b0; c0 < b1; c0 += Spec::chunk) {
                    const int64_t c1 = std::min<int64_t>(c0 + Spec::chunk, b1);
                    const int32_t len = (int32_t)(c1 - c0);
                    key.resize(len);
                    pv.assign(perm.begin() + c0, perm.begin() + c1);
                    for (int32_t j = 0; j < len; ++j)
                        key[j] = ((uint32_t)(scol[c0 + j] - bound[p]) << Spec::pos_bits) | (uint32_t)j;
                    std::sort(key.begin(), key.end());
                    for (int32_t j = 0; j < len; ++j) {
                        scol[c0 + j] = (int32_t)key[j];
                        perm[c0 + j] = pv[key[j] & ((1u << Spec::pos_bits) - 1)];
                    }
                }
            }
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < nth; ++w) pool.emplace_back(work, w);
        work(0);
        for (auto &th : pool) th.join();
    }
    M.nslices = P;
    M.s_sorted = sorted ? 1 : 0;
    M.s_fused = sorted && fused_tiles > 0 ? 1 : 0;
    M.s_pos_bits = sorted ? Spec::pos_bits : 0;
    M.s_ncols = hm.ncols;
    M.s_base = h->upload<int32_t>(base.data(), (int64_t)base.size());
    M.s_off = h->upload<uint16_t>(off.data(), (int64_t)off.size());
    M.s_col = h->upload<int32_t>(scol.data(), nnz);
    int32_t *dperm = h->upload<int32_t>(perm.data(), nnz);
    T *sval = h->dalloc<T>(nnz);
    launch_gather_perm(h->st, nnz, dval, dperm, sval);
    h->dfree(dperm);
    M.s_val = sval;
    M.s_tiles = h->upload<SliceTile>(tiles.data(), NT);
    M.s_ntiles = NT;
    M.s_part = h->dalloc<T>((int64_t)P * R);
    h->sync();
    return true;
}

// Releases the device arrays of a sliced form built by build_slices.
template <typename T>
void free_slices(hprsocp_handle *h, const CsrT<T> &S) {
    h->dfree(S.s_base);
    h->dfree(S.s_off);
    h->dfree(S.s_col);
    h->dfree(S.s_val);
    h->dfree(S.s_tiles);
    h->dfree(S.s_part);
}
void free_plan(hprsocp_handle *h, const DevPlan &p) {
    h->dfree(p.tiles);
    h->dfree(p.huge_buf);
}

// Autotuner of the tile plan (the reference fixes its launch shapes:
// gpu_launch_config, kernels.jl:349-357, and the cone size classes,
// algorithm.jl:1914-1919).  For an operator of at least kTuneMinNnz nonzeros
// (HPRSOCP_TUNE=1: any size, HPRSOCP_TUNE=0: never) it times one SpMV over the
// plan of each PlanConfig candidate (median of 5 HIP-event timed launches after
// a warm-up), coordinate-wise around the defaults: nonzeros per tile, lane
// width divisor, HUGE-segment targets (when the operator has split cones) and
// the CSR-stream path; then the combination of the per-coordinate winners.  A
// candidate replaces the current best only when it is at least 2% faster, and
// candidates are visited in a fixed order, so near-ties keep the earlier one.
// Results are cached per process by the operator's shape: the log2 histogram
// of its row lengths (linear rows and cone rows apart), the log2 histogram of
// its cone sizes and the instance count.  The choice changes only launch shapes
// (and with them the summation order of the row sums), never the arithmetic.
constexpr int64_t kTuneMinNnz = 1 << 21;

std::mutex g_tune_mu;
std::map<std::string, PlanConfig> g_tune_cache;

std::string tune_key(const std::vector<int32_t> &rp, int32_t nrows, int32_t lin_end, const std::vector<int32_t> &cidx,
                     int32_t ncones, int32_t nb) {
    std::map<std::tuple<int, int>, int64_t> hist;  // (class, log2 bucket) -> count
    auto l2 = [](int64_t v) {
        int b = 0;
        while ((1ll << b) <= v && b < 62) ++b;
        return b;
    };
    for (int32_t i = 0; i < nrows; ++i) hist[{i < lin_end ? 0 : 1, l2(rp[i + 1] - rp[i])}]++;
    for (int32_t k = 0; k < ncones; ++k) hist[{2, l2(cidx[k + 1] - cidx[k])}]++;
    std::string key = std::to_string(nrows) + "/" + std::to_string(lin_end) + "/" + std::to_string(ncones) + "/" +
                      std::to_string(nb > 1 ? nb : 1) + ":";
    for (const auto &kv : hist)
        key += std::to_string(std::get<0>(kv.first)) + "." + std::to_string(std::get<1>(kv.first)) + "=" +
               std::to_string(kv.second) + ";";
    return key;
}

PlanConfig tune_plan_config(hprsocp_handle *h, const CsrView &M, const std::vector<int32_t> &rp, int32_t nrows,
                            int32_t ncols, int32_t lin_end, const std::vector<int32_t> &cidx, int32_t ncones,
                            const std::vector<int32_t> *row_inst, const char *name) {
    hprsocp_handle::TuneRecord rec;
    rec.name = name;
    PlanConfig best;
    const int64_t nnz = rp.empty() ? 0 : rp[nrows];
    const char *env = getenv("HPRSOCP_TUNE");
    const bool force = env && atoi(env) == 1, off = env && atoi(env) == 0;
    const int32_t *ri = (row_inst && h->nb > 1) ? row_inst->data() : nullptr;
    auto plan_of = [&](const PlanConfig &c) {
        return build_plan(rp.data(), nrows, lin_end, cidx.empty() ? nullptr : cidx.data(), ncones, c, ri);
    };
    if (off || nnz == 0 || (!force && nnz < kTuneMinNnz)) {
        rec.cfg = best;
        rec.tiles = (int32_t)plan_of(best).tiles.size();
        h->tune.push_back(rec);
        return best;
    }
    const std::string key = tune_key(rp, nrows, lin_end, cidx, ncones, h->nb);
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            rec.cfg = it->second;
            rec.tuned = rec.cached = true;
            rec.tiles = (int32_t)plan_of(rec.cfg).tiles.size();
            h->tune.push_back(rec);
            return rec.cfg;
        }
    }
    double *xin = h->dalloc<double>(ncols);
    double *yout = h->dalloc<double>(nrows);
    launch_fill(h->st, ncols, 1.0, xin);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto time_cfg = [&](const PlanConfig &c) {
        HostPlan hp = plan_of(c);
        DevPlan dp;
        upload_plan(h, hp, dp);
        std::vector<float> t;
        for (int rep = 0; rep < 6; ++rep) {
            HIPCHK(hipEventRecord(e0, h->st));
            launch_spmv(h->st, dp, M, xin, yout);
            HIPCHK(hipEventRecord(e1, h->st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (rep) t.push_back(ms);
        }
        free_plan(h, dp);
        std::sort(t.begin(), t.end());
        ++rec.candidates;
        return t[t.size() / 2];
    };
    const PlanConfig def;
    const float t_def = time_cfg(def);
    float t_best = t_def;
    bool has_huge = false;
    for (int32_t k = 0; k < ncones && !has_huge; ++k) has_huge = cidx[k + 1] - cidx[k] > kTileRows;
    // coordinate candidates; each coordinate remembers its own winner
    struct Coord {
        std::vector<PlanConfig> alts;
        PlanConfig win;
        float t;
        bool improved = false;
    };
    std::vector<Coord> coords;
    auto with = [&](auto setter) {
        PlanConfig c = def;
        setter(c);
        return c;
    };
    coords.push_back(Coord{{with([](PlanConfig &c) { c.nnz_target = 1024; }),
                            with([](PlanConfig &c) { c.nnz_target = 4096; })}, def, t_def});
    coords.push_back(Coord{{with([](PlanConfig &c) { c.g_divisor = 1; }),
                            with([](PlanConfig &c) { c.g_divisor = 4; })}, def, t_def});
    if (has_huge)
        coords.push_back(Coord{{with([](PlanConfig &c) { c.huge_target = 8192; c.huge_whole = 4096; }),
                                with([](PlanConfig &c) { c.huge_target = 32768; c.huge_whole = 16384; })},
                               def, t_def});
    coords.push_back(Coord{{with([](PlanConfig &c) { c.stream = false; })}, def, t_def});
    for (Coord &co : coords)
        for (const PlanConfig &c : co.alts) {
            const float t = time_cfg(c);
            if (t < 0.98f * co.t) {
                co.t = t;
                co.win = c;
                co.improved = true;
            }
            if (t < 0.98f * t_best) {
                t_best = t;
                best = c;
            }
        }
    int nimp = 0;
    PlanConfig comb = def;
    for (size_t k = 0; k < coords.size(); ++k) {
        if (!coords[k].improved) continue;
        ++nimp;
        const PlanConfig &w = coords[k].win;
        if (w.nnz_target != def.nnz_target) comb.nnz_target = w.nnz_target;
        if (w.g_divisor != def.g_divisor) comb.g_divisor = w.g_divisor;
        if (w.huge_target != def.huge_target) {
            comb.huge_target = w.huge_target;
            comb.huge_whole = w.huge_whole;
        }
        if (w.stream != def.stream) comb.stream = w.stream;
    }
    if (nimp >= 2) {
        const float t = time_cfg(comb);
        if (t < 0.98f * t_best) {
            t_best = t;
            best = comb;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->dfree(xin);
    h->dfree(yout);
    if (h->params.verbose)
        printf("plan autotune %s: nnz/tile %d, g divisor %d, huge %d/%d, stream %d: %.4f ms (default %.4f ms, %d "
               "candidates)\n",
               name, best.nnz_target, best.g_divisor, best.huge_target, best.huge_whole, (int)best.stream, t_best,
               t_def, rec.candidates);
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tune_cache[key] = best;
    }
    rec.cfg = best;
    rec.tuned = true;
    rec.ms_default = t_def;
    rec.ms = rec.ms_form = t_best;
    rec.tiles = (int32_t)plan_of(best).tiles.size();
    h->tune.push_back(rec);
    return best;
}

hprsocp_handle::TuneRecord *tune_record(hprsocp_handle *h, const char *name) {
    for (auto &r : h->tune)
        if (r.name == name) return &r;
    return nullptr;
}

// Threads of the sorted-slice kernel for one sorted operator: 512 (each thread
// sums more rows of a chunk) or 1024 (twice the waves to hide the gathers);
// both sum every row in the same order.  `time` times one SpMV of the operator
// as it stands; 1024 is kept when it is more than 2% faster.
// HPRSOCP_SORT_BLOCK=<512|1024> forces the choice.
template <typename T, class Time>
float pick_sort_block(CsrT<T> &M, float t512, Time time) {
    if (!M.s_sorted) return t512;
    if (const char *env = getenv("HPRSOCP_SORT_BLOCK")) {
        M.s_block = atoi(env) == 1024 ? 1024 : kSortBlock;
        return t512;
    }
    M.s_block = 1024;
    const float t1024 = time(M);
    if (t1024 < 0.98f * t512) return t1024;
    M.s_block = kSortBlock;
    return t512;
}

// Workgroup slots of the sorted-slice kernels on the handle's device: two per
// CU (each workgroup holds a 64 KB LDS product buffer).  The slice-fused layout
// is cut into at most one (or two) tiles per slot, so that all its workgroups
// run in one round (or two) with about equal work.
int32_t sorted_slots(hprsocp_handle *h) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess || ncu <= 0)
        ncu = 256;
    return 2 * ncu;
}

// Autotuner of the XCD slicing: for an operator whose input vector outgrows
// ~2 MB it builds the sliced forms with 4 and 8 column slices (8: one slice per
// XCD; 4: two XCDs share a slice, halving the partial-sum traffic while the
// slice still fits one L2), each in row-major and in column-sorted layout,
// times one SpMV of each form and of the plain form with HIP events (median of
// 5 after a warm-up) and keeps the fastest with its epilogue plan
// (nnz-balanced tiles for the plain form, row tiles for sliced forms).
// HPRSOCP_XCD_SLICES=<P> forces P slices and HPRSOCP_XCD_SORTED=<0|1> the
// layout (tests).
void tune_slices(hprsocp_handle *h, CsrView &M, DevPlan &plan, const HostCsr &hm, const double *dval,
                 int32_t lin_end, const std::vector<int32_t> &cidx, int32_t ncones, const char *name,
                 const std::vector<int32_t> *row_inst) {
    const char *env = getenv("HPRSOCP_XCD_SLICES");
    const char *env_sorted = getenv("HPRSOCP_XCD_SORTED");
    const char *env_fused = getenv("HPRSOCP_SLICE_FUSED");
    if (hm.ci.empty()) return;
    const char *env_rounds = getenv("HPRSOCP_SLICE_ROUNDS");
    struct Cand {
        int32_t P;
        bool sorted;
        int32_t fused;  // slice-fused layout with fused * sorted_slots tiles (0: unfused)
        bool rounds;    // unfused sorted layout cut into whole rounds of workgroup slots
    };
    std::vector<Cand> cands;
    if (env) {
        const int32_t P = std::max(1, atoi(env));
        const bool srt = env_sorted && atoi(env_sorted) != 0;
        const int32_t fz = srt && env_fused ? std::max(0, atoi(env_fused)) : 0;
        if (P > 1) cands.push_back({P, srt, fz, srt && fz == 0 && env_rounds && atoi(env_rounds) != 0});
    } else if (hm.ncols * (int64_t)sizeof(double) > (2ll << 20)) {
        cands = {{4, false, 0, false}, {8, false, 0, false}, {4, true, 0, false}, {8, true, 0, false},
                 {8, true, 1, false},  {8, true, 2, false},  {8, true, 0, true}};
    }
    if (cands.empty()) return;
    double *xin = nullptr, *yout = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto time_one = [&](const CsrView &V, const DevPlan &pl) {
        std::vector<float> t;
        for (int rep = 0; rep < 6; ++rep) {
            HIPCHK(hipEventRecord(e0, h->st));
            launch_spmv(h->st, pl, V, xin, yout);
            HIPCHK(hipEventRecord(e1, h->st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (rep) t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        return t[t.size() / 2];
    };
    if (!env) {
        xin = h->dalloc<double>(hm.ncols);
        yout = h->dalloc<double>(hm.nrows);
        launch_fill(h->st, hm.ncols, 1.0, xin);
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
    }
    // every candidate is timed; the fastest wins unless a form earlier in the
    // preference order (see the ranks below) is within 2% of it,
    // which keeps the choice stable between runs when forms time alike
    struct Form {
        CsrView M;
        DevPlan plan;
        float t;
        int rank;
        bool sliced;
        bool rounds = false;
    };
    std::vector<Form> forms;
    forms.push_back(Form{M, plan, env ? INFINITY : time_one(M, plan), 4, false});  // plain
    int32_t natural_tiles[2] = {-1, -1};  // tiles of the natural column-sorted tiling, P = 4 / 8
    for (const auto &cand : cands) {
        const int32_t P = cand.P;
        CsrView S = M;
        if (!build_slices(h, S, hm, dval, P, cand.sorted, cand.fused * sorted_slots(h),
                          cand.rounds ? std::max(1, sorted_slots(h) / P) : 0))
            continue;
        if (cand.sorted && !cand.fused && (P == 4 || P == 8)) {
            int32_t &nat = natural_tiles[P == 8];
            if (!cand.rounds) {
                nat = S.s_ntiles;
            } else if (S.s_ntiles == nat) {  // the natural tiling fills whole rounds already
                free_slices(h, S);
                continue;
            }
        }
        DevPlan splan = make_plan(h, {}, hm.nrows, lin_end, cidx, ncones, row_inst);
        float t = env ? 0.0f : time_one(S, splan);
        if (env) {
            if (const char *eb = getenv("HPRSOCP_SORT_BLOCK")) S.s_block = atoi(eb) == 1024 ? 1024 : kSortBlock;
        } else {
            t = pick_sort_block(S, t, [&](const CsrView &V) { return time_one(V, splan); });
        }
        if (h->params.verbose)
            printf("XCD slicing autotune %s: %d slices%s%s%s (%d tiles, %d threads) %.4f ms (plain %.4f ms)\n", name,
                   P, cand.sorted ? " column-sorted" : "", S.s_fused ? " slice-fused" : "",
                   cand.rounds ? " whole rounds" : "", S.s_ntiles, S.s_block, t, forms[0].t);
        // preference order on near-ties: 8 sorted in whole rounds (the natural 8-sorted tiling
        // with its last round filled: as fast or faster inside the iteration, where the two
        // often time alike alone), 8 sorted, 8, 4 sorted, 4, plain, slice-fused
        const int rank = cand.rounds ? -1 : cand.fused ? 4 + cand.fused : (P >= 8 ? 0 : 2) + (cand.sorted ? 0 : 1);
        forms.push_back(Form{S, splan, t, rank, true, cand.rounds});
    }
    float tmin = INFINITY;
    for (const Form &f : forms) tmin = std::min(tmin, f.t);
    size_t pick = 0;
    for (size_t k = 0; k < forms.size(); ++k)
        if (forms[k].t <= 1.02f * tmin && (forms[pick].t > 1.02f * tmin || forms[k].rank < forms[pick].rank)) pick = k;
    for (size_t k = 1; k < forms.size(); ++k)
        if (k != pick) {
            free_slices(h, forms[k].M);
            free_plan(h, forms[k].plan);
        }
    const bool sliced = forms[pick].sliced;
    const CsrView bestM = forms[pick].M;
    const DevPlan bestP = forms[pick].plan;
    if (hprsocp_handle::TuneRecord *rec = tune_record(h, name)) {
        if (sliced)
            rec->form = std::string(bestM.s_sorted ? "sorted" : "slices") + std::to_string(bestM.nslices) +
                        (bestM.s_fused ? "-fused" : "") + (forms[pick].rounds ? "-rounds" : "");
        rec->block = sliced && bestM.s_sorted ? bestM.s_block : 0;
        rec->ms_form = forms[pick].t;
        if (rec->ms == 0) rec->ms = forms[0].t;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    h->dfree(xin);
    h->dfree(yout);
    if (sliced) {
        free_plan(h, plan);  // the plain form's epilogue plan is replaced (its CSR stays for other uses)
        M = bestM;
        plan = bestP;
    }
}

int32_t *contig_c0(hprsocp_handle *h, const HostCsr &a) {
    if (a.ci.empty()) return nullptr;
    std::vector<int32_t> c0((size_t)a.nrows, 0);
    for (int32_t r = 0; r < a.nrows; ++r) {
        const int32_t s = a.rp[r], e = a.rp[r + 1];
        if (s == e) continue;
        c0[r] = a.ci[s];
        for (int32_t k = s + 1; k < e; ++k)
            if (a.ci[k] != a.ci[s] + (k - s)) return nullptr;
    }
    return h->upload<int32_t>(c0.data(), a.nrows);
}

// ------------------------------------------------------------- validation
int validate_csr(std::string &err, const char *name, int32_t nr, int32_t nc, int64_t nnz, const int32_t *rp,
                 const int32_t *ci, const double *v) {
    char buf[256];
    if (nnz < 0) {
        err = std::string(name) + ": negative nnz";
        return HPRSOCP_ERR_INVALID;
    }
    if (nnz == 0 && !rp) return 0;
    if (!rp || (nnz > 0 && (!ci || !v))) {
        err = std::string(name) + ": missing CSR array";
        return HPRSOCP_ERR_INVALID;
    }
    if (rp[0] != 0 || rp[nr] != nnz) {
        err = std::string(name) + ": rowptr must start at 0 and end at nnz";
        return HPRSOCP_ERR_INVALID;
    }
    for (int32_t i = 0; i < nr; ++i)
        if (rp[i + 1] < rp[i]) {
            snprintf(buf, sizeof buf, "%s: rowptr not monotone at row %d", name, i);
            err = buf;
            return HPRSOCP_ERR_INVALID;
        }
    for (int64_t k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= nc) {
            snprintf(buf, sizeof buf, "%s: column index out of range at %lld", name, (long long)k);
            err = buf;
            return HPRSOCP_ERR_INVALID;
        }
    return 0;
}

// ------------------------------------------------------------- instances
SegPlan make_seg(hprsocp_handle *h, bool cols, std::vector<int32_t> &first) {
    std::vector<SegChunk> ch;
    first.clear();
    for (int i = 0; i < h->nb; ++i) {
        first.push_back((int32_t)ch.size());
        const InstRange &R = h->rng[i];
        const int32_t a[2][2] = {{cols ? R.box0 : R.lin0, cols ? R.box1 : R.lin1},
                                 {cols ? R.var0 : R.con0, cols ? R.var1 : R.con1}};
        for (int part = 0; part < 2; ++part)
            for (int32_t s0 = a[part][0]; s0 < a[part][1]; s0 += kChunk)
                ch.push_back(SegChunk{s0, std::min(s0 + kChunk, a[part][1]), i, 0});
    }
    first.push_back((int32_t)ch.size());
    SegPlan sp;
    sp.nchunks = (int32_t)ch.size();
    sp.ninst = h->nb;
    sp.chunks = h->upload<SegChunk>(ch.data(), (int64_t)ch.size());
    sp.first = h->upload<int32_t>(first.data(), (int64_t)first.size());
    return sp;
}

// Per-instance device state: chunk plans, cone -> instance maps, scalars, slots.
void setup_instances(hprsocp_handle *h) {
    const int nb = h->nb;
    h->seg_n = make_seg(h, true, h->seg_n_first);
    h->seg_m = make_seg(h, false, h->seg_m_first);
    h->max_chunks = std::max(1, std::max(h->seg_n.nchunks, h->seg_m.nchunks));
    std::vector<int32_t> ci((size_t)std::max(h->ncon, 1), 0), cc((size_t)std::max(h->ncon, 1), h->ncon);
    std::vector<int32_t> vi((size_t)std::max(h->nvar, 1), 0), vc((size_t)std::max(h->nvar, 1), h->nvar);
    for (int i = 0; i < nb; ++i) {
        const InstRange &R = h->rng[i];
        for (int32_t k = R.kc0; k < R.kc1; ++k) ci[k] = i, cc[k] = R.kc1 - R.kc0;
        for (int32_t k = R.kv0; k < R.kv1; ++k) vi[k] = i, vc[k] = R.kv1 - R.kv0;
    }
    h->cinst_con = h->upload<int32_t>(ci.data(), (int64_t)ci.size());
    h->ccnt_con = h->upload<int32_t>(cc.data(), (int64_t)cc.size());
    h->cinst_var = h->upload<int32_t>(vi.data(), (int64_t)vi.size());
    h->ccnt_var = h->upload<int32_t>(vc.data(), (int64_t)vc.size());
    h->isc = h->dalloc<InstScal>(nb);
    HIPCHK(hipHostMalloc((void **)&h->isc_h, sizeof(InstScal) * nb, hipHostMallocDefault));
    for (int i = 0; i < nb; ++i) h->isc_h[i] = InstScal{1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 1, 0};
    h->flag = h->dalloc<int32_t>(nb);
    HIPCHK(hipHostMalloc((void **)&h->flag_h, sizeof(int32_t) * nb, hipHostMallocDefault));
    h->ivtmp = h->dalloc<double>(2 * (int64_t)nb);
    HIPCHK(hipHostMalloc((void **)&h->ivtmp_h, sizeof(double) * 2 * nb, hipHostMallocDefault));
    h->dscal = h->dalloc<double>((int64_t)SL_COUNT * nb);
    HIPCHK(hipHostMalloc((void **)&h->hscal, sizeof(double) * SL_COUNT * nb, hipHostMallocDefault));
    h->pstate = h->dalloc<PowerState>(nb);
    h->red_region = (int64_t)std::max(h->max_chunks, nb) * kMaxGolden;
    h->red_part = h->dalloc<double>(3 * h->red_region);
    upload_isc(h);
}

// ------------------------------------------------------------- precision views
IterView<double> view64(hprsocp_handle *h) {
    IterView<double> v;
    v.A = h->A;
    v.AT = h->AT;
    v.Q = h->Q;
    v.F = h->F;
    v.FT = h->FT;
    v.Fdiag = h->Fdiag;
    v.Ftmp = h->Ftmp;
#define HPR_COPY(f) v.f = h->f;
    HPR_STATE_N(HPR_COPY) HPR_STATE_M(HPR_COPY) HPR_DATA_N(HPR_COPY) HPR_DATA_M(HPR_COPY)
    HPR_SCRATCH_N(HPR_COPY) HPR_SCRATCH_M(HPR_COPY)
#undef HPR_COPY
    return v;
}

// FP32 copy of a (scaled) operator: values only; index arrays and the slice
// layout are shared with the FP64 operator.
// FP32 copy of an operator: the FP64 index arrays and slice layout with FP32
// values.  A sliced operator also gets a column-sorted layout of its own with
// the same slices and FP32 chunks (SortSpec<float>: twice the entries per
// chunk, so the gathers of one instruction fall on fewer cache lines); the
// copy keeps whichever of the two runs one timed SpMV faster.  HPRSOCP_FP32_OWN_LAYOUT=<0|1> forces the
// choice (tests); with the slicing forced through HPRSOCP_XCD_SLICES the own
// layout is kept.  Both sum every row in row-major order, so the results are
// the same.  The float kernel reads either layout: it takes the chunk size
// from s_pos_bits, and the FP64 layout's chunks and tiles fit its LDS buffer
// and row registers.
static_assert(SortSpec<double>::chunk <= SortSpec<float>::chunk && SortSpec<double>::rows <= SortSpec<float>::rows,
              "the FP32 sorted-slice kernel must hold the FP64 layout's chunks");
CsrT<float> csr_to_f32(hprsocp_handle *h, const CsrView &M, int64_t nnz, const HostCsr *hm, const DevPlan *plan,
                       const char *name) {
    CsrT<float> f;
    f.rowptr = M.rowptr;
    f.col = M.col;
    f.nrows = M.nrows;
    f.c0 = M.c0;
    float *v = h->dalloc<float>(nnz);
    launch_demote(h->st, nnz, M.val, v);
    f.val = v;
    if (M.nslices <= 1) return f;
    CsrT<float> g = f;
    f.nslices = M.nslices;
    f.s_sorted = M.s_sorted;
    f.s_pos_bits = M.s_pos_bits;
    f.s_block = kSortBlock;
    f.s_ncols = M.s_ncols;
    f.s_base = M.s_base;
    f.s_off = M.s_off;
    f.s_col = M.s_col;
    f.s_tiles = M.s_tiles;
    f.s_ntiles = M.s_ntiles;
    float *sv = h->dalloc<float>(nnz);
    launch_demote(h->st, nnz, M.s_val, sv);
    f.s_val = sv;
    f.s_part = h->dalloc<float>((int64_t)M.nslices * M.nrows);
    f.s_fused = M.s_fused;
    hprsocp_handle::TuneRecord *rec = tune_record(h, name);
    if (!hm || !plan || !build_slices<float>(h, g, *hm, f.val, M.nslices, true)) {
        if (rec) rec->form32 = M.s_fused ? "fp64-fused" : "fp64";
        return f;
    }
    // the own layout also in the slice-fused tiling (as many tiles as the FP64
    // operator's fused layout, else one per workgroup slot)
    auto fused_own = [&](int32_t ntiles, int32_t round_tiles = 0) {
        CsrT<float> t = f;
        t.nslices = 1;
        t.s_ntiles = 0;
        t.s_sorted = 0;
        t.s_fused = 0;
        t.s_base = nullptr;
        t.s_off = nullptr;
        t.s_col = nullptr;
        t.s_val = nullptr;
        t.s_tiles = nullptr;
        t.s_part = nullptr;
        if (!build_slices<float>(h, t, *hm, f.val, M.nslices, true, ntiles, round_tiles)) t.s_ntiles = 0;
        return t;
    };
    CsrT<float> g2 = fused_own(M.s_fused ? M.s_ntiles : sorted_slots(h));
    // and the unfused own layout cut into whole rounds of workgroup slots
    CsrT<float> g3 = fused_own(0, std::max(1, sorted_slots(h) / M.nslices));
    if (g3.s_ntiles > 0 && g3.s_ntiles == g.s_ntiles) {  // the natural tiling already fills whole rounds
        free_slices(h, g3);
        g3.s_ntiles = 0;
    }
    const bool have_g2 = g2.s_ntiles > 0, have_g3 = g3.s_ntiles > 0;
    const char *force = getenv("HPRSOCP_FP32_OWN_LAYOUT");
    const char *force_fused = getenv("HPRSOCP_SLICE_FUSED");
    const bool timed = !force && !getenv("HPRSOCP_XCD_SLICES");
    // 0: FP64 layout, 1: own layout, 2: own slice-fused layout, 3: own layout in whole rounds
    int pick = !force || atoi(force) != 0 ? (force_fused && atoi(force_fused) != 0 && have_g2 ? 2 : 1) : 0;
    float *xin = nullptr, *yout = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto time_one = [&](const CsrT<float> &V) {
        std::vector<float> t;
        for (int rep = 0; rep < 6; ++rep) {
            HIPCHK(hipEventRecord(e0, h->st));
            launch_spmv<float>(h->st, *plan, V, xin, yout);
            HIPCHK(hipEventRecord(e1, h->st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (rep) t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        return t[t.size() / 2];
    };
    float tpick = 0;
    if (timed) {
        xin = h->dalloc<float>(hm->ncols);
        yout = h->dalloc<float>(M.nrows);
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        // every layout with its faster block size
        float tf = time_one(f), tg = time_one(g);
        float tg2 = have_g2 ? time_one(g2) : INFINITY, tg3 = have_g3 ? time_one(g3) : INFINITY;
        tf = pick_sort_block(f, tf, time_one);
        tg = pick_sort_block(g, tg, time_one);
        if (have_g2) tg2 = pick_sort_block(g2, tg2, time_one);
        if (have_g3) tg3 = pick_sort_block(g3, tg3, time_one);
        if (h->params.verbose)
            printf("FP32 column-sorted layout: %.4f ms with %d-entry chunks (%d tiles), %.4f ms in whole rounds (%d "
                   "tiles), %.4f ms slice-fused (%d tiles), %.4f ms with the FP64 layout%s\n",
                   tg, SortSpec<float>::chunk, g.s_ntiles, tg3, g3.s_ntiles, tg2, g2.s_ntiles, tf,
                   f.s_fused ? " (slice-fused)" : "");
        // near-ties (2%) keep the own unfused layout, then the FP64 one, then the others
        const float tmin = std::min(std::min(tf, tg), std::min(tg2, tg3));
        pick = tg <= 1.02f * tmin ? 1 : (tf <= 1.02f * tmin ? 0 : (tg2 <= 1.02f * tmin ? 2 : 3));
        tpick = pick == 0 ? tf : (pick == 1 ? tg : (pick == 2 ? tg2 : tg3));
    }
    CsrT<float> &keep = pick == 0 ? f : (pick == 1 ? g : (pick == 2 ? g2 : g3));
    if (!timed)
        if (const char *eb = getenv("HPRSOCP_SORT_BLOCK")) keep.s_block = atoi(eb) == 1024 ? 1024 : kSortBlock;
    if (rec) {
        rec->form32 = std::string(pick == 0 ? "fp64" : "own") + (keep.s_fused ? "-fused" : "") +
                      (pick == 3 ? "-rounds" : "");
        rec->block32 = keep.s_block;
        rec->ms32 = tpick;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    h->dfree(xin);
    h->dfree(yout);
    if (pick != 1) free_slices(h, g);
    if (pick != 2 && have_g2) free_slices(h, g2);
    if (pick != 3 && have_g3) free_slices(h, g3);
    if (pick == 0) return f;
    h->dfree(f.s_val);
    h->dfree(f.s_part);
    return keep;
}

// Host structure of the operators (for the FP32 sorted layouts).
struct HostOps {
    const HostCsr *A, *AT, *Q, *F, *FT;
};

// Allocates the FP32 mirrors; called once the FP64 operators are final.
void setup_fp32(hprsocp_handle *h, const HostOps &ho) {
    IterView<float> &v = h->v32;
    const int64_t n = h->n, m = h->m;
    v.A = csr_to_f32(h, h->A, h->nnzA, ho.A, &h->plan_A, "A");
    v.AT = csr_to_f32(h, h->AT, h->nnzA, ho.AT, &h->plan_AT, "A^T");
    if (h->nnzQ > 0) v.Q = csr_to_f32(h, h->Q, h->nnzQ, ho.Q, &h->plan_Q, "Q");
    if (h->q_kind == HPRSOCP_Q_FACTORED) {
        v.F = csr_to_f32(h, h->F, h->nnzF, ho.F, &h->plan_F, "F");
        v.FT = csr_to_f32(h, h->FT, h->nnzF, ho.FT, &h->plan_FT, "F^T");
        v.Ftmp = h->dalloc<float>(h->Frows);
        if (h->Fdiag) v.Fdiag = h->dalloc<float>(n);
    }
#define HPR_ALLOC_N(f) v.f = h->dalloc<float>(n);
#define HPR_ALLOC_M(f) v.f = h->dalloc<float>(m);
    HPR_STATE_N(HPR_ALLOC_N) HPR_DATA_N(HPR_ALLOC_N) HPR_SCRATCH_N(HPR_ALLOC_N)
    HPR_STATE_M(HPR_ALLOC_M) HPR_DATA_M(HPR_ALLOC_M) HPR_SCRATCH_M(HPR_ALLOC_M)
#undef HPR_ALLOC_N
#undef HPR_ALLOC_M
    h->fp32 = true;
    h->fp32_dirty = true;
    h->images_valid = false;
}

// FP64 state and data -> FP32 mirrors (round to nearest).
void demote_state(hprsocp_handle *h) {
    IterView<float> &v = h->v32;
    const int64_t n = h->n, m = h->m;
    hipStream_t st = h->st;
#define HPR_DEM_N(f) launch_demote(st, n, h->f, v.f);
#define HPR_DEM_M(f) launch_demote(st, m, h->f, v.f);
    HPR_STATE_N(HPR_DEM_N) HPR_DATA_N(HPR_DEM_N) HPR_STATE_M(HPR_DEM_M) HPR_DATA_M(HPR_DEM_M)
#undef HPR_DEM_N
#undef HPR_DEM_M
    if (v.Fdiag) launch_demote(st, n, h->Fdiag, v.Fdiag);
    h->fp32_dirty = false;
}

// FP32 state -> FP64 (exact widening) of the active instances running in FP32,
// for the FP64 checks, restarts and reads.  The FP64 state of instances
// running in FP64 is left alone (their mirrors are stale).
void promote_state(hprsocp_handle *h) {
    IterView<float> &v = h->v32;
    hipStream_t st = h->st;
#define HPR_PRO_N(f) launch_promote_prec(st, h->seg_n, v.f, h->f, h->isc);
#define HPR_PRO_M(f) launch_promote_prec(st, h->seg_m, v.f, h->f, h->isc);
    HPR_STATE_N(HPR_PRO_N) HPR_STATE_M(HPR_PRO_M)
#undef HPR_PRO_N
#undef HPR_PRO_M
}

// Delta form: the instances run the FP32 kernels (f32 = 1) or the FP64 ones.
void delta_prec(hprsocp_handle *h, int f32) {
    bool change = false;
    for (int i = 0; i < h->nb; ++i)
        if (h->I(i).fp32 != f32) {
            h->I(i).fp32 = f32;
            change = true;
        }
    if (!change) return;
    if (h->dctl) launch_isc_fp32(h->st, h->isc, h->nb, f32);  // sigma, inner0, active live on the device
    else upload_isc(h);
}

// Delta form: FP64 state -> deltas from the anchors, the anchors' images
// (ra, rya) and the FP32 data of the float kernels.  Runs whenever the FP64
// state changed (iterations, restarts, sigma, initial points).
// Batches after retirements: the compacted tile lists of A, A^T (and Q) cover the
// running instances.  The FP64 lists hold every running instance unless
// FP32-state instances are split off into their own lists; the products they
// skip belong to retired instances, whose outputs nothing reads.
static bool act64(const hprsocp_handle *h) { return h->use_act && (!h->fp32 || h->delta); }

void delta_split(hprsocp_handle *h) {
    hipStream_t st = h->st;
    IterView<float> &V = h->v32;
    const int64_t n = h->n, m = h->m;
    const SegPlan &sn = h->seg_n, &sm = h->seg_m;
    launch_delta_split(st, sn, h->x, h->last_x, V.x, h->isc);
    launch_delta_split(st, sm, h->y, h->last_y, V.y, h->isc);
    if (h->has_quad) {
        launch_delta_split(st, sn, h->w, h->last_w, V.w, h->isc);
        launch_delta_split(st, sn, h->aty, h->last_aty, V.aty, h->isc);
        if (h->Q_is_diag) launch_demote(st, n, h->fact, V.fact);
        if (V.Fdiag) launch_demote(st, n, h->Fdiag, V.Fdiag);
    }
    h->fp32_dirty = false;
    if (h->images_valid) return;  // no restart, sigma or data change since they were formed
    if (h->has_quad) {
        qmap(h, h->last_w, h->d_img_n);  // Q last_w
        launch_delta_ra(st, sn, h->last_x, h->last_aty, h->d_img_n, h->c, h->isc, h->d_ra);
    } else {
        launch_spmv(st, act64(h) ? h->act_AT : h->plan_AT, h->AT, h->last_y, h->d_img_n);  // A^T last_y
        launch_delta_ra(st, sn, h->last_x, h->d_img_n, nullptr, h->c, h->isc, h->d_ra);
    }
    if (m > 0) {
        launch_spmv(st, act64(h) ? h->act_A : h->plan_A, h->A, h->last_x, h->d_img_m);  // A last_x
        launch_delta_rya(st, sm, h->m1, h->d_img_m, h->rhs, h->last_y, h->isc, h->d_rya);
    }
    h->images_valid = true;
}

// Delta form: deltas -> FP64 state (x = last_x + dx, ...), before anything
// reads or writes the FP64 state.
void delta_join(hprsocp_handle *h) {
    if (!h->delta || !h->delta_live) return;
    hipStream_t st = h->st;
    IterView<float> &V = h->v32;
    const SegPlan &sn = h->seg_n, &sm = h->seg_m;
    launch_delta_join(st, sn, h->last_x, V.x, h->x, h->isc);
    launch_delta_join(st, sm, h->last_y, V.y, h->y, h->isc);
    if (h->has_quad) {
        launch_delta_join(st, sn, h->last_w, V.w, h->w, h->isc);
        launch_delta_join(st, sn, h->last_aty, V.aty, h->aty, h->isc);
    }
    h->delta_live = false;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

void hprsocp_default_params(hprsocp_params *p) {
    memset(p, 0, sizeof(*p));
    p->stoptol = 1e-6;
    p->sigma = -1.0;
    p->max_iter = 2147483647LL;
    p->time_limit = 3600.0;
    p->check_iter = 150;
    p->warm_up = 0;
    p->print_frequency = -1;
    p->device_number = 0;
    p->use_Ruiz_scaling = 1;
    p->ruiz_iterations = 10;
    p->use_bc_scaling = 1;
    p->bc_scaling_norm_type = HPRSOCP_BC_L2;
    p->use_l2_scaling = 0;
    p->use_Pock_Chambolle_scaling = 1;
    p->soc_block_scaling_strategy = HPRSOCP_SOC_PHASE_TAPER;
    p->auto_save = 0;
    p->save_filename = kDefaultSaveName;
    p->verbose = 1;
    p->use_gpu = 1;
    p->precision = HPRSOCP_PREC_FP64;
    p->refine_tol = 1e-6;
    p->fp32_stall_checks = 40;
    p->use_graph = 1;
}

const char *hprsocp_last_error(const hprsocp_handle *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

void hprsocp_destroy(hprsocp_handle *h) { delete h; }

// `ranges`/`objc` describe the instances of a block-diagonal batch (null: the
// problem is one instance).
static int create_impl(const hprsocp_problem *p, const hprsocp_params *params, hprsocp_handle *h,
                       const std::vector<InstRange> *ranges = nullptr, const std::vector<double> *objc = nullptr) {
    const double t0 = now_s();
    if (params) h->params = *params;
    else hprsocp_default_params(&h->params);
    hprsocp_params &P = h->params;
    if (!P.use_gpu) throw Err{HPRSOCP_ERR_UNSUPPORTED, "use_gpu = false: this backend runs the HPR iteration on the GPU only"};
    // the handle keeps its own copies of what auto_save writes later: the
    // caller's strings and initial points need only live through this call
    h->save_path = P.save_filename && P.save_filename[0] ? P.save_filename : kDefaultSaveName;
    P.save_filename = h->save_path.c_str();
    if (P.auto_save && ranges) throw Err{HPRSOCP_ERR_UNSUPPORTED, "auto_save applies to single-problem handles, not batches"};
    if (P.precision != HPRSOCP_PREC_FP64 && P.precision != HPRSOCP_PREC_FP32_REFINE &&
        P.precision != HPRSOCP_PREC_FP32_DELTA)
        throw Err{HPRSOCP_ERR_INVALID, "unknown precision mode"};
    if (P.precision == HPRSOCP_PREC_FP32_REFINE && (!(P.refine_tol >= 0.0) || P.fp32_stall_checks <= 0))
        throw Err{HPRSOCP_ERR_INVALID, "FP32 mode needs refine_tol >= 0 and fp32_stall_checks > 0"};
    if (P.sigma != -1.0 && !(P.sigma > 0.0)) {
        char b[160];
        snprintf(b, sizeof b, "Invalid sigma value: %g. It should be a positive number or -1 for automatic.", P.sigma);
        throw Err{HPRSOCP_ERR_INVALID, b};
    }
    if (P.check_iter <= 0) throw Err{HPRSOCP_ERR_INVALID, "check_iter must be positive"};
    if (P.print_frequency != -1 && P.print_frequency <= 0)
        throw Err{HPRSOCP_ERR_INVALID, "Invalid print_frequency: it should be a positive integer or -1"};
    const int32_t m = p->m, n = p->n;
    if (m < 0 || n < 0) throw Err{HPRSOCP_ERR_INVALID, "negative dimensions"};
    if (P.auto_save) {
        if (P.initial_x) h->init_x.assign(P.initial_x, P.initial_x + n);
        if (P.initial_y) h->init_y.assign(P.initial_y, P.initial_y + m);
    }
    P.initial_x = P.initial_y = nullptr;
    std::string e;
    if (int rc = validate_csr(e, "A", m, n, p->A_nnz, p->A_rowptr, p->A_colidx, p->A_val)) throw Err{rc, e};
    if (p->number_linear_con < 0 || p->number_linear_con > m) throw Err{HPRSOCP_ERR_INVALID, "number_linear_con out of range"};
    if (p->number_lu_x < 0 || p->number_lu_x > n) throw Err{HPRSOCP_ERR_INVALID, "number_lu_x out of range"};
    if (!p->c || !p->l || !p->u || (m > 0 && (!p->AL || !p->AU))) throw Err{HPRSOCP_ERR_INVALID, "missing vector"};
    h->m = m;
    h->n = n;
    h->m1 = p->number_linear_con;
    h->n1 = p->number_lu_x;
    h->ncon = p->n_soc_con;
    h->nvar = p->n_soc_var;
    if (h->ncon < 0 || h->nvar < 0) throw Err{HPRSOCP_ERR_INVALID, "negative cone count"};
    h->h_con_idx.assign((size_t)h->ncon + 1, m);
    h->h_var_idx.assign((size_t)h->nvar + 1, n);
    // SOC index vectors must start at the first SOC row / variable and end at the
    // sentinel, also when there is no cone (builders.jl:55-67)
    if (h->ncon > 0 && !p->soc_con_idx) throw Err{HPRSOCP_ERR_INVALID, "missing soc_con_idx"};
    if (p->soc_con_idx) std::copy(p->soc_con_idx, p->soc_con_idx + h->ncon + 1, h->h_con_idx.begin());
    else h->h_con_idx[0] = m;
    if (h->nvar > 0 && !p->soc_var_idx) throw Err{HPRSOCP_ERR_INVALID, "missing soc_var_idx"};
    if (p->soc_var_idx) std::copy(p->soc_var_idx, p->soc_var_idx + h->nvar + 1, h->h_var_idx.begin());
    else h->h_var_idx[0] = n;
    if (h->h_con_idx[0] != h->m1 || h->h_con_idx[h->ncon] != m)
        throw Err{HPRSOCP_ERR_INVALID, "soc_con_idx must start at number_linear_con and end at m"};
    if (h->h_var_idx[0] != h->n1 || h->h_var_idx[h->nvar] != n)
        throw Err{HPRSOCP_ERR_INVALID, "soc_var_idx must start at number_lu_x and end at n"};
    for (int32_t i = 0; i < h->ncon; ++i)
        if (h->h_con_idx[i + 1] < h->h_con_idx[i]) throw Err{HPRSOCP_ERR_INVALID, "soc_con_idx not monotone"};
    for (int32_t i = 0; i < h->nvar; ++i)
        if (h->h_var_idx[i + 1] < h->h_var_idx[i]) throw Err{HPRSOCP_ERR_INVALID, "soc_var_idx not monotone"};

    // ---- instances
    if (ranges) {
        h->rng = *ranges;
    } else {
        h->rng.assign(1, InstRange{0, h->m1, h->m1, m, 0, h->n1, h->n1, n, 0, h->ncon, 0, h->nvar});
    }
    h->nb = (int32_t)h->rng.size();
    h->ih.assign(h->nb, InstHost{});
    for (int i = 0; i < h->nb; ++i) h->ih[i].obj_const = objc ? (*objc)[i] : p->obj_constant;
    h->h_row_inst.assign((size_t)m, 0);
    h->h_col_inst.assign((size_t)n, 0);
    for (int i = 0; i < h->nb; ++i) {
        const InstRange &R = h->rng[i];
        for (int32_t r = R.lin0; r < R.lin1; ++r) h->h_row_inst[r] = i;
        for (int32_t r = R.con0; r < R.con1; ++r) h->h_row_inst[r] = i;
        for (int32_t c = R.box0; c < R.box1; ++c) h->h_col_inst[c] = i;
        for (int32_t c = R.var0; c < R.var1; ++c) h->h_col_inst[c] = i;
    }

    // ---- device
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (P.device_number < 0 || P.device_number >= ndev) {
        char b[128];
        snprintf(b, sizeof b, "device_number %d not available (%d devices)", P.device_number, ndev);
        throw Err{HPRSOCP_ERR_DEVICE, b};
    }
    h->device = P.device_number;
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    setup_instances(h);

    auto phase = [&](const char *what) {
        if (P.verbose) printf("setup: %-28s %8.3f s\n", what, now_s() - t0);
    };
    phase("validation + instances");
    // ---- host-side matrices: A, A^T, Q (symmetrized, canonical), F, F^T
    HostCsr A;
    A.nrows = m;
    A.ncols = n;
    A.rp.assign(p->A_rowptr ? p->A_rowptr : nullptr, p->A_rowptr ? p->A_rowptr + m + 1 : nullptr);
    if (A.rp.empty()) A.rp.assign((size_t)m + 1, 0);
    par_assign(A.ci, p->A_colidx, p->A_nnz);
    par_assign(A.v, p->A_val, p->A_nnz);
    HostCsr AT = transpose(A);
    h->nnzA = p->A_nnz;
    h->q_kind = p->Q_kind;
    HostCsr Qh;
    if (p->Q_kind == HPRSOCP_Q_FACTORED) {
        if (int rc = validate_csr(e, "F", p->F_rows, n, p->F_nnz, p->F_rowptr, p->F_colidx, p->F_val)) throw Err{rc, e};
        h->has_quad = 1;
    } else if (p->Q_kind == HPRSOCP_Q_CSR) {
        if (int rc = validate_csr(e, "Q", n, n, p->Q_nnz, p->Q_rowptr, p->Q_colidx, p->Q_val)) throw Err{rc, e};
        if (p->Q_nnz > 0) {
            Qh.nrows = Qh.ncols = n;
            Qh.rp.assign(p->Q_rowptr, p->Q_rowptr + n + 1);
            Qh.ci.assign(p->Q_colidx, p->Q_colidx + p->Q_nnz);
            Qh.v.assign(p->Q_val, p->Q_val + p->Q_nnz);
            canonicalize(Qh);
            HostCsr Qt = transpose(Qh);
            bool sym = Qt.rp == Qh.rp && Qt.ci == Qh.ci && Qt.v == Qh.v;
            if (!sym) {  // (Q + Q^T)/2 up front for non-symmetric input
                HostCsr S;
                S.nrows = S.ncols = n;
                S.rp.push_back(0);
                for (int32_t i = 0; i < n; ++i) {
                    int32_t a = Qh.rp[i], ae = Qh.rp[i + 1], b = Qt.rp[i], be = Qt.rp[i + 1];
                    while (a < ae || b < be) {
                        const int32_t ca = a < ae ? Qh.ci[a] : INT32_MAX, cb = b < be ? Qt.ci[b] : INT32_MAX;
                        if (ca == cb) {
                            S.ci.push_back(ca);
                            S.v.push_back((Qh.v[a++] + Qt.v[b++]) / 2.0);
                        } else if (ca < cb) {
                            S.ci.push_back(ca);
                            S.v.push_back(Qh.v[a++] / 2.0);
                        } else {
                            S.ci.push_back(cb);
                            S.v.push_back(Qt.v[b++] / 2.0);
                        }
                    }
                    S.rp.push_back((int32_t)S.ci.size());
                }
                Qh = S;
            }
        }
        h->has_quad = Qh.v.size() > 0;
    } else {
        throw Err{HPRSOCP_ERR_INVALID, "unknown Q_kind"};
    }
    h->nnzQ = (int64_t)Qh.v.size();

    phase("host transpose / Q");
    // ---- upload
    int32_t *A_rp = h->upload<int32_t>(A.rp.data(), m + 1);
    int32_t *A_ci = h->upload<int32_t>(A.ci.data(), (int64_t)A.ci.size());
    h->A_v = h->upload<double>(A.v.data(), (int64_t)A.v.size());
    h->A = CsrView{A_rp, A_ci, h->A_v, m};
    int32_t *AT_rp = h->upload<int32_t>(AT.rp.data(), n + 1);
    int32_t *AT_ci = h->upload<int32_t>(AT.ci.data(), (int64_t)AT.ci.size());
    h->AT_v = h->upload<double>(AT.v.data(), (int64_t)AT.v.size());
    h->AT = CsrView{AT_rp, AT_ci, h->AT_v, n};
    h->A.c0 = contig_c0(h, A);
    h->AT.c0 = contig_c0(h, AT);
    h->h_A_rp = A.rp;
    h->con_idx = h->upload<int32_t>(h->h_con_idx.data(), h->ncon + 1);
    h->var_idx = h->upload<int32_t>(h->h_var_idx.data(), h->nvar + 1);
    if (h->nnzQ > 0) {
        int32_t *Q_rp = h->upload<int32_t>(Qh.rp.data(), n + 1);
        int32_t *Q_ci = h->upload<int32_t>(Qh.ci.data(), h->nnzQ);
        h->Q_v = h->upload<double>(Qh.v.data(), h->nnzQ);
        h->Q_v2 = h->dalloc<double>(h->nnzQ);
        h->Q = CsrView{Q_rp, Q_ci, h->Q_v, n};
        h->Q.c0 = contig_c0(h, Qh);
        // transpose partner of every entry (Qh is symmetric in pattern, rows sorted)
        HostCsr idx = Qh;
        for (int64_t k = 0; k < h->nnzQ; ++k) idx.v[k] = (double)k;
        HostCsr idxT = transpose(idx);
        std::vector<int32_t> tpos((size_t)h->nnzQ);
        for (int64_t k = 0; k < h->nnzQ; ++k) tpos[k] = (int32_t)idxT.v[k];
        h->Q_tpos = h->upload<int32_t>(tpos.data(), h->nnzQ);
        bool sd = true;
        for (int32_t i = 0; i < n && sd; ++i)
            for (int32_t k = Qh.rp[i]; k < Qh.rp[i + 1]; ++k)
                if (Qh.ci[k] != i) {
                    sd = false;
                    break;
                }
        h->q_struct_diag = sd;
        const PlanConfig qcfg = sd ? PlanConfig{}
                                   : tune_plan_config(h, h->Q, Qh.rp, n, n, n, {}, 0, &h->h_col_inst, "Q");
        h->plan_Q = make_plan(h, Qh.rp, n, n, {}, 0, &h->h_col_inst, &h->hp_Q, qcfg);
    }
    HostCsr F, FT;
    if (h->q_kind == HPRSOCP_Q_FACTORED) {
        F.nrows = p->F_rows;
        F.ncols = n;
        F.rp.assign(p->F_rowptr, p->F_rowptr + p->F_rows + 1);
        F.ci.assign(p->F_colidx, p->F_colidx + p->F_nnz);
        F.v.assign(p->F_val, p->F_val + p->F_nnz);
        FT = transpose(F);
        h->Frows = p->F_rows;
        h->nnzF = p->F_nnz;
        int32_t *F_rp = h->upload<int32_t>(F.rp.data(), F.nrows + 1);
        int32_t *F_ci = h->upload<int32_t>(F.ci.data(), (int64_t)F.ci.size());
        double *F_v = h->upload<double>(F.v.data(), (int64_t)F.v.size());
        h->F = CsrView{F_rp, F_ci, F_v, F.nrows};
        h->F.c0 = contig_c0(h, F);
        int32_t *FT_rp = h->upload<int32_t>(FT.rp.data(), n + 1);
        int32_t *FT_ci = h->upload<int32_t>(FT.ci.data(), (int64_t)FT.ci.size());
        double *FT_v = h->upload<double>(FT.v.data(), (int64_t)FT.v.size());
        h->FT = CsrView{FT_rp, FT_ci, FT_v, n};
        h->FT.c0 = contig_c0(h, FT);
        h->Fdiag = p->Q_diag ? h->upload<double>(p->Q_diag, n) : nullptr;
        h->Ftmp = h->dalloc<double>(F.nrows);
        const PlanConfig fcfg = tune_plan_config(h, h->F, F.rp, F.nrows, n, F.nrows, {}, 0, nullptr, "F");
        const PlanConfig ftcfg = tune_plan_config(h, h->FT, FT.rp, n, F.nrows, n, {}, 0, nullptr, "F^T");
        h->plan_F = make_plan(h, F.rp, F.nrows, F.nrows, {}, 0, nullptr, nullptr, fcfg);
        h->plan_FT = make_plan(h, FT.rp, n, n, {}, 0, nullptr, nullptr, ftcfg);
    }
    // plans: A rows (linear | constraint cones), A^T rows (box | variable cones)
    const PlanConfig acfg =
        tune_plan_config(h, h->A, A.rp, m, n, h->m1, h->h_con_idx, h->ncon, &h->h_row_inst, "A");
    const PlanConfig atcfg =
        tune_plan_config(h, h->AT, AT.rp, n, m, h->n1, h->h_var_idx, h->nvar, &h->h_col_inst, "A^T");
    h->plan_A = make_plan(h, A.rp, m, h->m1, h->h_con_idx, h->ncon, &h->h_row_inst, &h->hp_A, acfg);
    h->plan_AT = make_plan(h, AT.rp, n, h->n1, h->h_var_idx, h->nvar, &h->h_col_inst, &h->hp_AT, atcfg);
    if (h->nb > 1) {
        h->act_A = alloc_active(h, h->plan_A);
        h->act_AT = alloc_active(h, h->plan_AT);
        if (h->nnzQ > 0) h->act_Q = alloc_active(h, h->plan_Q);
        if (P.precision == HPRSOCP_PREC_FP32_REFINE) {
            h->act32_A = alloc_active(h, h->plan_A);
            h->act32_AT = alloc_active(h, h->plan_AT);
            if (h->nnzQ > 0) h->act32_Q = alloc_active(h, h->plan_Q);
        }
    }
    {
        std::vector<int32_t> rc((size_t)m, -1);
        for (int32_t k = 0; k < h->ncon; ++k)
            for (int32_t i = h->h_con_idx[k]; i < h->h_con_idx[k + 1]; ++i) rc[i] = k;
        h->row_cone = h->upload<int32_t>(rc.data(), m);
    }
    // vectors
    h->c = h->upload<double>(p->c, n);
    h->AL = h->upload<double>(p->AL, m);
    h->AU = h->upload<double>(p->AU, m);
    h->l = h->upload<double>(p->l, n);
    h->u = h->upload<double>(p->u, n);
    {
        std::vector<double> rf((size_t)m, 0.0);
        if (p->soc_rhs)
            for (int32_t i = h->m1; i < m; ++i) rf[i] = p->soc_rhs[i - h->m1];
        h->rhs = h->upload<double>(rf.data(), m);
    }
    double **vn[] = {&h->col_norm, &h->diag_Q, &h->x, &h->x_hat, &h->x_bar, &h->z_bar, &h->dx,
                     &h->last_x, &h->w, &h->w_bar, &h->last_w, &h->dw, &h->aty, &h->aty_bar, &h->last_aty,
                     &h->atdy, &h->qatdy, &h->qw, &h->dqw, &h->qx, &h->tempv, &h->rd, &h->fact1, &h->fact2,
                     &h->fact, &h->fact_m, &h->raw_n, &h->tmp_n, &h->tmp_n2};
    for (double **pp : vn) *pp = h->dalloc<double>(n);
    h->row_norm = h->dalloc<double>(m);
    double **vm[] = {&h->y, &h->y_bar, &h->dy, &h->last_y, &h->s, &h->ax, &h->rp, &h->raw_m, &h->tmp_m};
    for (double **pp : vm) *pp = h->dalloc<double>(m);
    if (P.auto_save) {  // saved_state of allocate_workspace (algorithm.jl:1574-1585)
        h->sv_x = h->dalloc<double>(n);
        h->sv_z = h->dalloc<double>(n);
        h->sv_w = h->dalloc<double>(n);
        h->sv_y = h->dalloc<double>(m);
    }
    h->part_n = h->dalloc<double>(h->plan_AT.nslots);
    h->part_m = h->dalloc<double>(h->plan_A.nslots);
    h->cone_t = h->dalloc<double>(h->ncon);
    h->cone_ns = h->dalloc<double>(h->ncon);
    {  // pieces of at most kConePiece rows per constraint cone (at least one per cone)
        std::vector<int32_t> pcone, prow, first(1, 0), multi;
        for (int32_t c = 0; c < h->ncon; ++c) {
            const int32_t s0 = h->h_con_idx[c], e0 = h->h_con_idx[c + 1];
            const int32_t np = std::max<int32_t>(1, (e0 - s0 + kConePiece - 1) / kConePiece);
            for (int32_t k = 0; k < np; ++k) {
                pcone.push_back(c);
                prow.push_back(s0 + k * kConePiece);
            }
            first.push_back((int32_t)pcone.size());
            if (np > 1) multi.push_back(c);
        }
        ConePieces &pc = h->cone_pieces;
        pc.npieces = (int32_t)pcone.size();
        pc.nmulti = (int32_t)multi.size();
        pc.cone = h->upload<int32_t>(pcone.data(), pc.npieces);
        pc.row = h->upload<int32_t>(prow.data(), pc.npieces);
        pc.first = h->upload<int32_t>(first.data(), (int64_t)first.size());
        pc.multi = h->upload<int32_t>(multi.data(), pc.nmulti);
        pc.part = h->dalloc<double>(pc.npieces);
    }

    phase("upload + plans + alloc");
    do_scaling(h);
    phase("scaling");
    // XCD column slices of every SpMV operator, kept where they measure faster
    // (values are final after scaling)
    // (batches gather from per-instance slices of the input already; their tile
    // lists get compacted as instances retire, which the sliced form cannot follow)
    if (h->nb == 1) {
        tune_slices(h, h->A, h->plan_A, A, h->A_v, h->m1, h->h_con_idx, h->ncon, "A", &h->h_row_inst);
        tune_slices(h, h->AT, h->plan_AT, AT, h->AT_v, h->n1, h->h_var_idx, h->nvar, "A^T", &h->h_col_inst);
    }
    if (h->q_kind == HPRSOCP_Q_FACTORED) {
        tune_slices(h, h->F, h->plan_F, F, h->F.val, F.nrows, {}, 0, "F", nullptr);
        tune_slices(h, h->FT, h->plan_FT, FT, h->FT.val, n, {}, 0, "F^T", nullptr);
    } else if (h->nnzQ > 0 && !h->q_struct_diag && h->nb == 1) {
        tune_slices(h, h->Q, h->plan_Q, Qh, h->Q_v, n, {}, 0, "Q", &h->h_col_inst);
    }
    phase("slice autotune");
    // check_Q_diagonal (scaling.jl:1312-1353), after scaling as in solve(); every
    // instance of a batch must land in the same class
    if (h->q_kind == HPRSOCP_Q_FACTORED) {
        h->Q_is_diag = 0;
    } else if (h->nnzQ == 0) {
        h->Q_is_diag = 1;
    } else {
        launch_q_diag(h->st, h->seg_n, h->Q.rowptr, h->Q.col, h->Q_v, h->diag_Q, region(h, 0), slot(h, SL_GEN),
                      SL_COUNT);
        read_slots(h);
        for (int i = 0; i < h->nb; ++i) {
            h->ih[i].Q_is_diag = hs(h, i, SL_GEN) == 0.0;
            h->ih[i].lambda_Q_diag = hs(h, i, SL_GEN + 2);
        }
        h->Q_is_diag = h->ih[0].Q_is_diag;
        for (int i = 1; i < h->nb; ++i)
            if (h->ih[i].Q_is_diag != h->Q_is_diag)
                throw Err{HPRSOCP_ERR_INVALID, "batch instances must all have a diagonal Q or all a general Q"};
    }
    {
        const char *env = getenv("HPRSOCP_FUSE_QW");
        h->fuse_qw = h->has_quad && h->q_kind == HPRSOCP_Q_FACTORED && h->nvar == 0 && h->nb == 1 &&
                     h->plan_FT.ntiles > 0 && !(env && atoi(env) == 0);
    }
    // sigma_0 (allocate_workspace, algorithm.jl:1538-1551)
    for (int i = 0; i < h->nb; ++i) {
        if (P.sigma == -1.0) {
            const double nbv = h->ih[i].norm_b, nc = h->ih[i].norm_c;
            h->I(i).sigma = (nc > 1e-16 && nbv > 1e-16 && nbv < 1e16 && nc < 1e16) ? nbv / nc : 1.0;
        } else {
            h->I(i).sigma = P.sigma;
        }
    }
    const HostOps ho{&A, &AT, &Qh, &F, &FT};
    if (P.precision == HPRSOCP_PREC_FP32_REFINE) {
        setup_fp32(h, ho);
        for (int i = 0; i < h->nb; ++i) h->I(i).fp32 = 1;
    } else if (P.precision == HPRSOCP_PREC_FP32_DELTA) {
        setup_fp32(h, ho);
        h->fp32 = false;  // the instance is an FP64 one outside the delta iterations
        h->delta = true;
        h->d_ra = h->dalloc<double>(n);
        h->d_img_n = h->dalloc<double>(n);
        h->d_rya = h->dalloc<double>(m);
        h->d_img_m = h->dalloc<double>(m);
    }
    upload_isc(h);
    if (h->Q_is_diag) update_q_factors(h);
    h->to_check = 1;
    h->sync();
    h->setup_time = now_s() - t0;
    return 0;
}

int hprsocp_create(const hprsocp_problem *prob, const hprsocp_params *params, hprsocp_handle **out) {
    *out = nullptr;
    g_create_error.clear();
    if (!prob) {
        g_create_error = "null problem";
        return HPRSOCP_ERR_INVALID;
    }
    hprsocp_handle *h = new (std::nothrow) hprsocp_handle();
    if (!h) return HPRSOCP_ERR_ALLOC;
    try {
        create_impl(prob, params, h);
    } catch (const Err &e) {
        g_create_error = e.msg;
        delete h;
        return e.code;
    }
    *out = h;
    return 0;
}

#define ABI_TRY(h, ...)                 \
    try {                               \
        __VA_ARGS__;                    \
    } catch (const Err &e_) {           \
        (h)->err = e_.msg;              \
        return e_.code;                 \
    }

int hprsocp_estimate_lambda(hprsocp_handle *h) {
    ABI_TRY(h, {
        delta_join(h);
        h->fp32_dirty = true;
        h->images_valid = false;
        const double t0 = now_s();
        const int nb = h->nb;
        std::vector<double> la(nb, 0.0), lq(nb, 0.0);
        if (h->m > 0) {
            la = power_iteration(h, true);
            for (double &v : la) v *= 1.05;
        }
        if (h->q_kind == HPRSOCP_Q_FACTORED || (h->nnzQ > 0 && !h->Q_is_diag)) {
            lq = power_iteration(h, false);
            for (double &v : lq) v *= 1.05;
        } else if (h->nnzQ > 0) {
            for (int i = 0; i < nb; ++i) lq[i] = std::isfinite(h->ih[i].lambda_Q_diag) ? h->ih[i].lambda_Q_diag : 0.0;
        }
        for (int i = 0; i < nb; ++i) {
            h->I(i).lambda_A = la[i];
            h->I(i).lambda_Q = lq[i];
        }
        upload_isc(h);
        h->lambda_set = true;
        h->power_time = now_s() - t0;
    });
    return 0;
}

}  // extern "C"

namespace {
// process_initial_points! (algorithm.jl:2612-2677) for the instances flagged in
// fx (x0 given) / fy (y0 given); x0/y0 are full-length (block-diagonal layout).
void set_initial_impl(hprsocp_handle *h, const double *x0, const double *y0, const std::vector<int32_t> &fx,
                      const std::vector<int32_t> &fy) {
    hipStream_t st = h->st;
    const int32_t m = h->m, n = h->n, nb = h->nb;
    delta_join(h);
    auto set_flags = [&](const std::vector<int32_t> &f) {
        memcpy(h->flag_h, f.data(), sizeof(int32_t) * nb);
        HIPCHK(hipMemcpyAsync(h->flag, h->flag_h, sizeof(int32_t) * nb, hipMemcpyHostToDevice, st));
        h->sync();
    };
    const InstVal bsv = isc_field(h, &h->isc->b_scale), csv = isc_field(h, &h->isc->c_scale);
    if (x0) {
        set_flags(fx);
        HIPCHK(hipMemcpyAsync(h->tmp_n, x0, sizeof(double) * n, hipMemcpyHostToDevice, st));
        launch_scale_vd(st, h->seg_n, h->tmp_n, h->col_norm, bsv, h->tmp_n2);
        double *dst[] = {h->x, h->x_bar, h->last_x, h->w, h->w_bar, h->last_w};
        for (int k = 0; k < (h->has_quad ? 6 : 3); ++k) launch_copy_flagged(st, h->seg_n, h->tmp_n2, dst[k], h->flag);
    }
    if (y0) {
        set_flags(fy);
        HIPCHK(hipMemcpyAsync(h->tmp_m, y0, sizeof(double) * m, hipMemcpyHostToDevice, st));
        launch_scale_vd(st, h->seg_m, h->tmp_m, h->row_norm, csv, h->rp);
        launch_copy_flagged(st, h->seg_m, h->rp, h->y, h->flag);
        launch_copy_flagged(st, h->seg_m, h->rp, h->y_bar, h->flag);
        launch_copy_flagged(st, h->seg_m, h->rp, h->last_y, h->flag);
        double *aty_dst = h->has_quad ? h->aty_bar : h->aty;
        if (m > 0) {
            launch_spmv(st, h->plan_AT, h->AT, h->y_bar, h->tmp_n2);
            launch_copy_flagged(st, h->seg_n, h->tmp_n2, aty_dst, h->flag);
            if (h->has_quad) {
                launch_copy_flagged(st, h->seg_n, h->tmp_n2, h->aty, h->flag);
                launch_copy_flagged(st, h->seg_n, h->tmp_n2, h->last_aty, h->flag);
            }
        }
        if (h->has_quad) {
            qmap(h, h->x_bar, h->qx);
            launch_init_zbar(st, h->seg_n, h->x_bar, h->qx, h->aty_bar, h->c, h->isc, h->tmp_n2);
        } else {
            launch_init_zbar(st, h->seg_n, h->x_bar, nullptr, h->aty, h->c, h->isc, h->tmp_n2);
        }
        launch_copy_flagged(st, h->seg_n, h->tmp_n2, h->z_bar, h->flag);
        if (m > 0) {
            launch_spmv(st, h->plan_A, h->A, h->x_bar, h->ax);
            launch_init_s(st, h->seg_m, h->ax, h->y, h->AL, h->AU, h->isc, h->rp);
            launch_copy_flagged(st, h->seg_m, h->rp, h->s, h->flag);
        }
    }
    h->fp32_dirty = true;
    h->images_valid = false;
    h->sync();
}
}  // namespace

extern "C" {

int hprsocp_set_initial(hprsocp_handle *h, const double *x0, const double *y0) {
    ABI_TRY(h, { set_initial_impl(h, x0, y0, std::vector<int32_t>(h->nb, 1), std::vector<int32_t>(h->nb, 1)); });
    return 0;
}

// main_update_gpu! (algorithm.jl:2284-2309)
extern "C++" {
// Q v on the update's precision (factored or general CSR Q).
template <typename T>
static void qmap_t(hprsocp_handle *h, IterView<T> &V, const T *v, T *out) {
    if (h->q_kind == HPRSOCP_Q_FACTORED) {
        launch_spmv<T>(h->st, h->plan_F, V.F, v, V.Ftmp);
        launch_spmv<T>(h->st, h->plan_FT, V.FT, V.Ftmp, out, V.Fdiag, v);
    } else {
        launch_spmv<T>(h->st, h->plan_Q, V.Q, v, out);
    }
}

// One HPR iteration (main_update_gpu!, algorithm.jl:2284-2309) on the
// matrices and vectors of V.  Scalars come per instance from h->isc (sigma,
// lambdas, Halpern counter inner0 + kstep), rounded to T inside the kernels.
template <typename T>
static void do_update_t(hprsocp_handle *h, IterView<T> &V, int64_t kstep, int full) {
    hipStream_t st = h->st;
    const int32_t m = h->m;
    const bool f32 = sizeof(T) == sizeof(float);
    const DevPlan &pA = h->use_act ? (f32 ? h->act32_A : h->act_A) : h->plan_A;
    const DevPlan &pAT = h->use_act ? (f32 ? h->act32_AT : h->act_AT) : h->plan_AT;
    const DevPlan &pQ = h->use_act ? (f32 ? h->act32_Q : h->act_Q) : h->plan_Q;
    auto ev = [&](int k) {
        if (h->profile) HIPCHK(hipEventRecord(h->ev[k], st));
    };
    if (!h->has_quad) {
        ev(0);
        NoQPrimalArgsT<T> a{V.y, V.x, V.x_hat, V.x_bar, V.z_bar, V.dx, V.last_x, V.c, V.l, V.u, h->isc, kstep, full};
        launch_noq_primal<T>(st, pAT, V.AT, h->var_idx, a, V.raw_n, h->part_n);
        ev(1);
        if (m > 0) {
            DualArgsT<T> d{V.x_hat, V.y, V.y_bar, V.s, V.dy, V.last_y, V.AL, V.AU, V.rhs, h->isc, kstep, full, 0};
            launch_dual<T>(st, pA, V.A, h->con_idx, d, V.raw_m, h->part_m);
        }
        ev(2);
        return;
    }
    const int sdiag = h->q_struct_diag && h->q_kind == HPRSOCP_Q_CSR;
    const bool fq = h->fuse_qw && !h->use_act;
    ev(0);
    if (fq) launch_spmv<T>(st, h->plan_F, V.F, V.w, V.Ftmp);  // F w; F^T runs with the primal step
    else if (!sdiag) qmap_t<T>(h, V, V.w, V.qw);
    ev(1);
    QPPrimalArgsT<T> a{V.x, V.x_hat, V.x_bar, V.z_bar, V.dx, V.w_bar, V.tempv, V.w, V.last_x, V.c,
                       V.l, V.u, V.aty, V.qw, V.diag_Q, V.fact1, V.fact2, h->isc, kstep,
                       full, sdiag, h->Q_is_diag};
    if (fq) launch_qp_primal_fq<T>(st, h->plan_FT, V.FT, V.Ftmp, V.Fdiag, V.qw, a);
    else launch_qp_primal<T>(st, pAT, h->var_idx, a, V.raw_n, h->part_n);
    ev(2);
    if (m > 0) {
        if (!sdiag) {
            if (h->q_kind == HPRSOCP_Q_FACTORED) {
                launch_spmv<T>(st, h->plan_F, V.F, V.w_bar, V.Ftmp);
                TempvArgsT<T> t{V.x_hat, V.qw, V.tempv, V.Fdiag, V.w_bar, h->isc};
                launch_spmv_tempv<T>(st, h->plan_FT, V.FT, V.Ftmp, t);
            } else {
                TempvArgsT<T> t{V.x_hat, V.qw, V.tempv, nullptr, nullptr, h->isc};
                launch_spmv_tempv<T>(st, pQ, V.Q, V.w_bar, t);
            }
        }
        ev(3);
        DualArgsT<T> d{V.tempv, V.y, V.y_bar, V.s, V.dy, V.last_y, V.AL, V.AU, V.rhs, h->isc, kstep, full, 1};
        launch_dual<T>(st, pA, V.A, h->con_idx, d, V.raw_m, h->part_m);
    } else {
        ev(3);
    }
    ev(4);
    W2ArgsT<T> w2{V.y_bar, V.w, V.w_bar, V.aty, V.dw, V.atdy, V.aty_bar, V.last_w, V.last_aty, V.fact,
                  h->isc, kstep, full, h->Q_is_diag};
    launch_w2<T>(st, pAT, V.AT, w2);
    ev(5);
}
// One partial HPR iteration in the delta form (precision FP32_DELTA): the same
// launch groups as do_update_t, with FP32 SpMVs on the FP32 deltas and the
// primal / dual epilogues forming raw values and projections in FP64 from the
// anchors (DPrimalArgs).  The w / aty update needs no anchor: OpW2 on the
// deltas with null anchors is the delta update.
static void do_update_delta_t(hprsocp_handle *h, int64_t kstep) {
    hipStream_t st = h->st;
    IterView<float> &V = h->v32;
    auto ev = [&](int k) {
        if (h->profile) HIPCHK(hipEventRecord(h->ev[k], st));
    };
    const int sdiag = h->q_struct_diag && h->q_kind == HPRSOCP_Q_CSR;
    // the running instances' tiles (batches after retirements), as do_update_t
    const DevPlan &pA = h->use_act ? h->act_A : h->plan_A;
    const DevPlan &pAT = h->use_act ? h->act_AT : h->plan_AT;
    const DevPlan &pQ = h->use_act ? h->act_Q : h->plan_Q;
    DPrimalArgs a{V.x, V.x_hat, V.w_bar, V.tempv, V.y, V.w, V.aty, V.qw, h->d_ra, h->last_x, h->last_w, h->l, h->u,
                  h->diag_Q, h->fact1, h->fact2, h->isc, kstep, sdiag, h->Q_is_diag};
    DDualArgs d{V.y, V.y_bar, h->d_rya, h->last_y, h->AL, h->AU, h->isc, kstep};
    if (!h->has_quad) {
        ev(0);
        launch_dprimal(st, pAT, V.AT, h->var_idx, a, false, h->raw_n, h->part_n);
        ev(1);
        if (h->m > 0) launch_ddual(st, pA, V.A, h->con_idx, V.x_hat, d, h->raw_m, h->part_m);
        ev(2);
        return;
    }
    const bool fq = h->fuse_qw && !h->use_act;
    ev(0);
    if (fq) launch_spmv<float>(st, h->plan_F, V.F, V.w, V.Ftmp);  // F dw; F^T runs with the primal step
    else if (!sdiag) qmap_t<float>(h, V, V.w, V.qw);              // Q dw
    ev(1);
    if (fq) launch_dprimal_fq(st, h->plan_FT, V.FT, V.Ftmp, V.Fdiag, V.qw, a);
    else launch_dprimal(st, pAT, V.AT, h->var_idx, a, true, h->raw_n, h->part_n);
    ev(2);
    if (h->m > 0) {
        if (!sdiag) {  // tempv - last_x = (x_hat - last_x) + sigma (Q dw - Q (w_bar - last_w))
            if (h->q_kind == HPRSOCP_Q_FACTORED) {
                launch_spmv<float>(st, h->plan_F, V.F, V.w_bar, V.Ftmp);
                TempvArgsT<float> t{V.x_hat, V.qw, V.tempv, V.Fdiag, V.w_bar, h->isc};
                launch_spmv_tempv<float>(st, h->plan_FT, V.FT, V.Ftmp, t);
            } else {
                TempvArgsT<float> t{V.x_hat, V.qw, V.tempv, nullptr, nullptr, h->isc};
                launch_spmv_tempv<float>(st, pQ, V.Q, V.w_bar, t);
            }
        }
        ev(3);
        launch_ddual(st, pA, V.A, h->con_idx, V.tempv, d, h->raw_m, h->part_m);
    } else {
        ev(3);
    }
    ev(4);
    W2ArgsT<float> w2{V.y_bar, V.w,    V.w_bar, V.aty,  nullptr, nullptr, nullptr, nullptr,
                      nullptr, V.fact, h->isc,  kstep,  0,       h->Q_is_diag};
    launch_w2<float>(st, pAT, V.AT, w2);
    ev(5);
}
}  // extern "C++"

// Sets every instance's Halpern counter (InstScal::inner0) and restarts the
// step count of the following updates.
static void set_inner(hprsocp_handle *h, const std::vector<int64_t> &inner) {
    for (int i = 0; i < h->nb; ++i) h->I(i).inner0 = (double)inner[i];
    upload_isc(h);
    h->kstep = 0;
}

// One iteration: the float kernels update the instances running in FP32, the
// double kernels the others.  The FP32 mirrors are refreshed from FP64 first if
// the FP64 state changed, and the FP64 state of the FP32 instances is
// refreshed from the mirrors after a full (check) iteration or when `sync64`.
static void do_update(hprsocp_handle *h, int full, bool sync64 = false) {
    h->to_check = full;
    const int64_t k = h->kstep++;
    if (h->delta) {  // check iterations in FP64 on the joined state, the others on the deltas
        if (full) {
            delta_join(h);
            delta_prec(h, 0);
            IterView<double> V = view64(h);
            do_update_t<double>(h, V, k, 1);
            h->fp32_dirty = true;
        } else {
            if (h->fp32_dirty) delta_split(h);
            delta_prec(h, 1);
            do_update_delta_t(h, k);
            h->delta_live = true;
            if (sync64) delta_join(h);
        }
        return;
    }
    if (h->n32 > 0) {
        if (h->fp32_dirty) demote_state(h);
        do_update_t<float>(h, h->v32, k, full);
        if (full || sync64) promote_state(h);
    }
    if (h->n64 > 0) {
        IterView<double> V = view64(h);
        do_update_t<double>(h, V, k, full);
    }
}

// Tile lists of the running instances, one set per precision, after
// retirements or precision switches; the captured graphs no longer fit.
static void compact_active(hprsocp_handle *h) {
    const int nb = h->nb;
    std::vector<char> k64(nb), k32(nb);
    for (int i = 0; i < nb; ++i) {  // the delta form runs every instance in both kernel sets
        k64[i] = h->I(i).active && (!h->I(i).fp32 || h->delta);
        k32[i] = h->I(i).active && h->I(i).fp32;
    }
    compact_plan(h, h->hp_A, h->act_A, k64);
    compact_plan(h, h->hp_AT, h->act_AT, k64);
    if (h->nnzQ > 0) compact_plan(h, h->hp_Q, h->act_Q, k64);
    if (h->fp32) {
        compact_plan(h, h->hp_A, h->act32_A, k32);
        compact_plan(h, h->hp_AT, h->act32_AT, k32);
        if (h->nnzQ > 0) compact_plan(h, h->hp_Q, h->act32_Q, k32);
    }
    h->use_act = true;
    h->clear_graphs();
}

// Stream capture under rocprofv3's kernel tracing has ended in a host SIGSEGV
// inside the HIP runtime on this ROCm image, with no capturable-API misuse in
// the captured code (kernel launches only).  A process that rocprofv3 starts
// (it sets ROCPROFILER_LIBRARY_CTOR / ROCP_TOOL_LIBRARIES) therefore launches
// every window eagerly: the same kernels with the same arguments, so the same
// results, and a kernel trace that covers every update.
static bool capture_refused() {
    static const bool refused = [] {
        const bool r = getenv("ROCPROFILER_LIBRARY_CTOR") || getenv("ROCP_TOOL_LIBRARIES");
        if (r) fprintf(stderr, "hprsocp: running under rocprofv3: HIP graph capture off, windows launched eagerly\n");
        return r;
    }();
    return refused;
}

// L consecutive partial updates.  With params.use_graph the run is captured
// once per (first step, length, precision) into a HIP graph and replayed: the
// kernels read sigma, the Halpern counter base and the retirement mask from
// device memory, so one capture serves every check window that starts at the
// same step.  Graphs are dropped when the launch shapes change (tile
// compaction) and captured again on demand.  In a batch, instances retire at
// many different checks and every retirement compacts the tile lists, so a
// window is captured only when it is requested a second time since the last
// compaction; the first time it runs eagerly (the same kernels, so the same
// results), which keeps capture costs out of the stretches between retirements.
static void run_partial_updates(hprsocp_handle *h, int64_t L) {
    const int fp = h->delta ? 4 : (h->n32 > 0 ? 1 : 0) | (h->n64 > 0 ? 2 : 0);
    const auto key = std::make_tuple(h->kstep, (int32_t)L, fp);
    bool eager = !h->params.use_graph || L < 2 || capture_refused();
    if (!eager && h->nb > 1 && !h->graphs.count(key)) eager = h->graph_seen[key]++ == 0;
    if (eager) {
        for (int64_t k = 0; k < L; ++k) do_update(h, 0);
        return;
    }
    h->to_check = 0;
    if (h->delta) {
        if (h->fp32_dirty) delta_split(h);
        delta_prec(h, 1);
    } else if (h->n32 > 0 && h->fp32_dirty) {
        demote_state(h);
    }
    auto it = h->graphs.find(key);
    if (it == h->graphs.end()) {
        if (h->graphs.size() >= 64) {
            if (h->dctl) h->retire_graphs();  // the stream may still run one of them
            else h->clear_graphs();
        }
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
        IterView<double> V = view64(h);
        for (int64_t k = 0; k < L; ++k) {
            if (fp & 1) do_update_t<float>(h, h->v32, h->kstep + k, 0);
            if (fp & 2) do_update_t<double>(h, V, h->kstep + k, 0);
            if (fp & 4) do_update_delta_t(h, h->kstep + k);
        }
        HIPCHK(hipStreamEndCapture(h->st, &g));
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
        it = h->graphs.emplace(key, ge).first;
    }
    HIPCHK(hipGraphLaunch(it->second, h->st));
    h->kstep += L;
    if (h->delta) h->delta_live = true;
}

int hprsocp_update(hprsocp_handle *h, int64_t inner, int32_t to_check) {
    if (!h->lambda_set) {
        h->err = "estimate_lambda must run before update";
        return HPRSOCP_ERR_STATE;
    }
    ABI_TRY(h, {
        set_inner(h, std::vector<int64_t>(h->nb, inner));
        do_update(h, to_check ? 1 : 0, true);
        HIPCHK(hipGetLastError());
    });
    return 0;
}

// compute_residuals! (algorithm.jl:113-196) for every instance; r has nb entries.
// Constants of instance i that the residual block needs.
static ctl::InstConst inst_const(hprsocp_handle *h, int i) {
    const InstHost &H = h->ih[i];
    const InstRange &R = h->rng[i];
    return ctl::InstConst{H.norm_b_org, H.norm_c_org, H.obj_const, R.m() > 0, R.lin1 > R.lin0, R.con1 > R.con0};
}

// The residual reductions of compute_residuals! (algorithm.jl:113-196) into the
// SL_RN / SL_RM / SL_BV slots, no host synchronisation.
static void residuals_enqueue(hprsocp_handle *h, int64_t iter) {
    hipStream_t st = h->st;
    delta_join(h);
    const int32_t m = h->m, n = h->n;
    const bool skip_q = !h->has_quad;
    // batches after retirements: the SpMVs over the running instances' tiles (the
    // compacted FP64 lists hold every running instance unless FP32-state
    // instances are split off into their own lists)
    const bool act = act64(h);
    if (!skip_q) qmap(h, h->x_bar, h->qx);
    if (m > 0) launch_spmv(st, act ? h->act_AT : h->plan_AT, h->AT, h->y_bar, h->atdy);
    else launch_fill(st, n, 0.0, h->atdy);
    ResidNArgs an{h->c, h->x_bar, h->z_bar, h->col_norm, skip_q ? nullptr : h->qx, h->atdy, h->rd, h->isc};
    launch_resid_n(st, h->seg_n, an, region(h, 0), slot(h, SL_RN), SL_COUNT);
    if (m > 0) {
        launch_spmv(st, act ? h->act_A : h->plan_A, h->A, h->x_bar, h->ax);
        launch_ax_to_org(st, h->seg_m, h->ax, h->row_norm, h->isc);
        launch_cone_rp_stats(st, h->ncon, h->con_idx, h->cinst_con, h->cone_pieces, h->ax, h->rhs, h->row_norm,
                             h->isc, h->cone_t, h->cone_ns);
        ResidMArgs am{h->y_bar, h->s, h->rhs, h->AL, h->AU, h->row_norm, h->ax, h->rp, h->row_cone, h->con_idx,
                      h->cone_t, h->cone_ns, h->isc, h->m1};
        launch_resid_m(st, h->seg_m, am, region(h, 1), slot(h, SL_RM), SL_COUNT);
    }
    if (iter == 0)
        launch_bounds_violation(st, h->seg_n, h->nvar, h->var_idx, h->cinst_var, h->x_bar, h->l, h->u, h->col_norm,
                                h->isc, h->dx, region(h, 2), slot(h, SL_BV), SL_COUNT);
}

static void residuals_impl(hprsocp_handle *h, int64_t iter, hprsocp_residuals *r) {
    const bool skip_q = !h->has_quad;
    residuals_enqueue(h, iter);
    read_slots(h);
    for (int i = 0; i < h->nb; ++i) {
        const double *v = h->hscal + (int64_t)i * SL_COUNT;
        ctl::residuals(v + SL_RN, v + SL_RM, v[SL_BV], h->I(i).b_scale, h->I(i).c_scale, inst_const(h, i), skip_q,
                       iter == 0, r[i]);
    }
}

int hprsocp_compute_residuals(hprsocp_handle *h, int64_t iter, hprsocp_residuals *out) {
    ABI_TRY(h, {
        std::vector<hprsocp_residuals> r(h->nb);
        residuals_impl(h, iter, r.data());
        *out = r[0];
    });
    return 0;
}

int hprsocp_probe_copy(hprsocp_handle *h, int64_t bytes, double *ms) {
    ABI_TRY(h, {
        const int64_t n = std::max<int64_t>(bytes / (int64_t)sizeof(double), 1);
        void *a = nullptr, *b = nullptr;
        if (hipMalloc(&a, (size_t)n * sizeof(double)) != hipSuccess) throw Err{HPRSOCP_ERR_ALLOC, "probe alloc"};
        if (hipMalloc(&b, (size_t)n * sizeof(double)) != hipSuccess) {
            (void)hipFree(a);
            throw Err{HPRSOCP_ERR_ALLOC, "probe alloc"};
        }
        launch_fill(h->st, n, 1.0, (double *)a);
        h->sync();
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, h->st));
        launch_copy(h->st, n, (const double *)a, (double *)b);
        HIPCHK(hipEventRecord(e1, h->st));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        *ms = t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(a);
        (void)hipFree(b);
    });
    return 0;
}

// Median HIP-event time of `reps` launches of one operator's SpMV as the
// iteration runs it (sliced or plain form, stored-product epilogue) on
// device-resident operands: A x (trans = 0) or A^T y (trans = 1).
int hprsocp_time_spmv(hprsocp_handle *h, int32_t trans, int32_t reps, double *ms) {
    ABI_TRY(h, {
        if (reps < 1 || !ms) throw Err{HPRSOCP_ERR_INVALID, "time_spmv: reps >= 1 and an output are required"};
        if (h->nb != 1) throw Err{HPRSOCP_ERR_UNSUPPORTED, "time_spmv applies to single-problem handles"};
        if (trans & 2) {  // the FP32 operator of the FP32 forms
            if (!h->fp32 && !h->delta)
                throw Err{HPRSOCP_ERR_STATE, "time_spmv: FP32 operators exist only in the FP32 precisions"};
            const bool t1 = trans & 1;
            const CsrT<float> &M32 = t1 ? h->v32.AT : h->v32.A;
            const DevPlan &pl = t1 ? h->plan_AT : h->plan_A;
            const int64_t lin = t1 ? h->m : h->n, lout = t1 ? h->n : h->m;
            float *din = h->dalloc<float>(lin), *dout = h->dalloc<float>(lout);
            launch_spmv<float>(h->st, pl, M32, din, dout);  // warm-up
            h->sync();
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            std::vector<float> t((size_t)reps);
            for (int r = 0; r < reps; ++r) {
                HIPCHK(hipEventRecord(e0, h->st));
                launch_spmv<float>(h->st, pl, M32, din, dout);
                HIPCHK(hipEventRecord(e1, h->st));
                HIPCHK(hipEventSynchronize(e1));
                HIPCHK(hipEventElapsedTime(&t[r], e0, e1));
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            h->dfree(din);
            h->dfree(dout);
            std::sort(t.begin(), t.end());
            *ms = t[t.size() / 2];
            return 0;
        }
        const CsrView &M = trans ? h->AT : h->A;
        const DevPlan &pl = trans ? h->plan_AT : h->plan_A;
        double *din = trans ? h->tmp_m : h->tmp_n;
        double *dout = trans ? h->tmp_n2 : h->ax;
        launch_fill(h->st, trans ? h->m : h->n, 1.0, din);
        launch_spmv(h->st, pl, M, din, dout);  // warm-up
        h->sync();
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        std::vector<float> t((size_t)reps);
        for (int r = 0; r < reps; ++r) {
            HIPCHK(hipEventRecord(e0, h->st));
            launch_spmv(h->st, pl, M, din, dout);
            HIPCHK(hipEventRecord(e1, h->st));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&t[r], e0, e1));
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
    });
    return 0;
}

// Gather-ceiling probe: nnz synthetic entries (rounded down to whole chunks)
// gathered from `slices` slices of `width` doubles, in random or column-sorted
// chunk order; median time of 5 launches after a warm-up.
}  // extern "C"

template <typename T>
static void probe_gather_t(hprsocp_handle *h, int64_t nnz, int32_t width, int32_t slices, int32_t sorted, double *ms) {
    const int C = sorted ? SortSpec<T>::chunk : kStreamNnz;
    const int bits = sorted ? SortSpec<T>::pos_bits : kProbePosBits;
    if (!ms || slices < 1 || width < 1 || nnz < C || (int64_t)width > (1ll << (32 - bits)))
        throw Err{HPRSOCP_ERR_INVALID, "probe_gather: need nnz >= one chunk, 1 <= width <= 2^(32-position bits), slices >= 1"};
    const int NT = sorted ? kSortBlock : kBlock;
    const int64_t nch = nnz / C;
    if (nch > (int64_t)INT32_MAX) throw Err{HPRSOCP_ERR_INVALID, "probe_gather: too many chunks"};
    void *pk = nullptr, *val = nullptr, *x = nullptr, *out = nullptr;
    auto release = [&]() {
        (void)hipFree(pk);
        (void)hipFree(val);
        (void)hipFree(x);
        (void)hipFree(out);
    };
    if (hipMalloc(&pk, (size_t)(nch * C) * 4) != hipSuccess || hipMalloc(&val, (size_t)(nch * C) * sizeof(T)) != hipSuccess ||
        hipMalloc(&x, (size_t)slices * width * sizeof(T)) != hipSuccess ||
        hipMalloc(&out, (size_t)(nch * NT) * sizeof(T)) != hipSuccess) {
        release();
        throw Err{HPRSOCP_ERR_ALLOC, "probe_gather alloc"};
    }
    if constexpr (sizeof(T) == 8) launch_fill(h->st, (int64_t)slices * width, 1.0, (double *)x);
    else HIPCHK(hipMemsetAsync(x, 0, (size_t)slices * width * sizeof(T), h->st));  // values do not matter
    launch_probe_fill<T>(h->st, nch, width, sorted, (uint32_t *)pk, (T *)val);
    launch_probe_gather<T>(h->st, nch, slices, width, sorted, (const uint32_t *)pk, (const T *)val, (const T *)x,
                           (T *)out);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    std::vector<float> t(5);
    for (int r = 0; r < 5; ++r) {
        HIPCHK(hipEventRecord(e0, h->st));
        launch_probe_gather<T>(h->st, nch, slices, width, sorted, (const uint32_t *)pk, (const T *)val,
                               (const T *)x, (T *)out);
        HIPCHK(hipEventRecord(e1, h->st));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&t[r], e0, e1));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->sync();
    release();
    std::sort(t.begin(), t.end());
    *ms = t[2];
}

extern "C" {

int hprsocp_probe_gather(hprsocp_handle *h, int64_t nnz, int32_t width, int32_t slices, int32_t sorted, double *ms) {
    ABI_TRY(h, { probe_gather_t<double>(h, nnz, width, slices, sorted, ms); });
    return 0;
}

int hprsocp_probe_gather32(hprsocp_handle *h, int64_t nnz, int32_t width, int32_t slices, int32_t sorted, double *ms) {
    ABI_TRY(h, { probe_gather_t<float>(h, nnz, width, slices, sorted, ms); });
    return 0;
}

int hprsocp_probe_ctl(hprsocp_handle *h, const double *in, int32_t n, double *dev_out, double *host_out) {
    ABI_TRY(h, {
        if (n < 0 || (n > 0 && (!in || !dev_out || !host_out)))
            throw Err{HPRSOCP_ERR_INVALID, "hprsocp_probe_ctl: n < 0 or a null array"};
        if (n == 0) return 0;
        double *din = h->upload(in, n);
        double *dout = h->dalloc<double>(5 * (int64_t)n);
        launch_ctl_probe(h->st, din, n, dout);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(dev_out, dout, sizeof(double) * 5 * (size_t)n, hipMemcpyDeviceToHost, h->st));
        h->sync();
        h->dfree(din);
        h->dfree(dout);
        for (int32_t k = 0; k < n; ++k) ctl_probe_eval(in[k], host_out + 5 * (int64_t)k);
    });
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ M-norm
namespace {

// Enqueues the reductions of compute_M_norm! (algorithm.jl:1302-1359) into 8
// device slots per instance starting at `base`; no host synchronisation.
struct MnormCtx {
    std::vector<double> sigma, lambda_A, lambda_Q;
    bool pending = false;
    void snap(hprsocp_handle *h) {
        sigma.resize(h->nb);
        lambda_A.resize(h->nb);
        lambda_Q.resize(h->nb);
        for (int i = 0; i < h->nb; ++i) {
            sigma[i] = h->I(i).sigma;
            lambda_A[i] = h->I(i).lambda_A;
            lambda_Q[i] = h->I(i).lambda_Q;
        }
        pending = true;
    }
};

void mnorm_enqueue(hprsocp_handle *h, int base) {
    const int32_t m = h->m;
    const bool q = h->has_quad;
    hipStream_t st = h->st;
    if (q) qmap(h, h->dw, h->dqw);
    if (m > 0) {
        launch_spmv(st, act64(h) ? h->act_AT : h->plan_AT, h->AT, h->dy, h->atdy);
        if (q) qmap(h, h->atdy, h->qatdy);
    }
    std::vector<RedTerm> tn;
    tn.push_back(RedTerm{h->dx, h->dx, nullptr, 0});                               // d0 = |dx|^2
    tn.push_back(q ? RedTerm{h->dqw, h->dx, nullptr, 0} : RedTerm{h->dx, h->dx, nullptr, 0});  // d1
    tn.push_back(m > 0 && q ? RedTerm{h->dqw, h->atdy, nullptr, 0} : RedTerm{h->dx, h->dx, nullptr, 0});  // d3
    tn.push_back(m > 0 ? RedTerm{h->atdy, h->dx, nullptr, 0} : RedTerm{h->dx, h->dx, nullptr, 0});       // d4
    if (q && h->Q_is_diag) {
        tn.push_back(m > 0 ? RedTerm{h->atdy, h->fact_m, h->qatdy, 0} : RedTerm{h->dx, h->dx, nullptr, 0});  // d5
        tn.push_back(RedTerm{h->dqw, h->dqw, nullptr, 0});                                                  // d6
    } else if (q) {
        tn.push_back(m > 0 ? RedTerm{h->atdy, h->qatdy, nullptr, 0} : RedTerm{h->dx, h->dx, nullptr, 0});
        tn.push_back(RedTerm{h->dw, h->dqw, nullptr, 0});
    }
    launch_reduce(st, h->seg_n, tn.data(), (int)tn.size(), region(h, 0), slot(h, base), SL_COUNT);
    RedTerm ty{h->dy, h->dy, nullptr, 0};
    launch_reduce(st, h->seg_m, &ty, 1, region(h, 1), slot(h, base + 7), SL_COUNT);  // d2
}

// M-norm of instance i from its 8 reduced dots d (host copy)
double mnorm_finish(hprsocp_handle *h, int i, const double *d, const MnormCtx &c) {
    return ctl::mnorm(d, c.sigma[i], c.lambda_A[i], c.lambda_Q[i], h->has_quad, h->rng[i].m() > 0, h->Q_is_diag);
}

// ------------------------------------------------------------------ restart
// HPRSOCP_restart (structs.jl:600-623) is the public hprsocp_restart_info.
using Restart = hprsocp_restart_info;

// initialize_restart, algorithm.jl:1886-1908
void init_restart(Restart &r) { ctl::init_restart(r); }

// check_restart, algorithm.jl:1016-1099
void check_restart(Restart &r, int64_t iter, int64_t check_iter, double sigma, double kkt) {
    ctl::check_restart(r, iter, check_iter, sigma, kkt);
}

// The golden-section walks of the instances whose walk is active, on the
// device: kGoldenPasses evaluations of g(x) = sum_j ATdy_j * QATdy_j / (1 + x d_j)
// at each walk's points (unified_golden_Q_diag, unified_operations.jl:411-449),
// one batched pass over the searching instances' vectors per 2^L - 1 points.
void golden_passes_enqueue(hprsocp_handle *h) {
    if (!h->gwalk) {
        h->gwalk = h->dalloc<GoldenWalk>(h->nb);  // zeroed: no walk active
        h->gnp = h->dalloc<int32_t>(h->nb);
    }
    double *xd = region(h, 2);  // nb * kMaxGolden points (region 2 is free here)
    for (int p = 0;; ++p) {
        launch_gold_step(h->st, h->gwalk, h->nb, h->dscal, SL_COUNT, SL_GOLD, xd, h->gnp);
        if (p == kGoldenPasses) break;
        launch_golden_eval(h->st, h->seg_n, h->atdy, h->qatdy, h->diag_Q, xd, (1 << kGoldenL) - 1, h->gnp,
                           region(h, 0), slot(h, SL_GOLD), SL_COUNT);
    }
}

// unified_golden_Q_diag for the listed instances at once, from the host.
std::vector<double> golden_q_diag(hprsocp_handle *h, const std::vector<int> &ids, const std::vector<double> &a,
                                  const std::vector<double> &b, double lo, double hi, double tol) {
    const int nb = h->nb;
    std::vector<GoldenWalk> W((size_t)nb);
    memset(W.data(), 0, sizeof(GoldenWalk) * W.size());
    for (size_t k = 0; k < ids.size(); ++k) gw_init(W[ids[k]], a[k], b[k], lo, hi, tol, kGoldenL);
    if (!h->gwalk) {
        h->gwalk = h->dalloc<GoldenWalk>(nb);
        h->gnp = h->dalloc<int32_t>(nb);
    }
    HIPCHK(hipMemcpyAsync(h->gwalk, W.data(), sizeof(GoldenWalk) * nb, hipMemcpyHostToDevice, h->st));
    golden_passes_enqueue(h);
    HIPCHK(hipMemcpyAsync(W.data(), h->gwalk, sizeof(GoldenWalk) * nb, hipMemcpyDeviceToHost, h->st));
    h->sync();
    std::vector<double> r(ids.size());
    for (size_t k = 0; k < ids.size(); ++k) {
        if (!gw_done(W[ids[k]])) throw Err{HPRSOCP_ERR_STATE, "golden-section search did not finish"};
        r[k] = gw_result(W[ids[k]]);
    }
    return r;
}

// The dots update_sigma! needs (algorithm.jl:781-829), for every instance, into
// the SL_SIG slots: s[0] |x_bar - last_x|^2, s[7] |y_bar - last_y|^2 and, with
// a quadratic term, the dots of dw = w_bar - last_w, dQw, ATdy = ATy_bar -
// last_ATy and QATdy.  No host synchronisation.
void sigma_dots_enqueue(hprsocp_handle *h) {
    hipStream_t st = h->st;
    const int32_t m = h->m, n = h->n;
    const bool skip_q = !h->has_quad, any_m = m > 0;
    if (!skip_q) {
        launch_axpby(st, n, 1.0, h->w_bar, -1.0, h->last_w, h->dw);
        qmap(h, h->dw, h->dqw);
    }
    launch_axpby(st, n, 1.0, h->x_bar, -1.0, h->last_x, h->dx);
    if (any_m) {
        launch_axpby(st, m, 1.0, h->y_bar, -1.0, h->last_y, h->dy);
        if (!skip_q) {
            launch_axpby(st, n, 1.0, h->aty_bar, -1.0, h->last_aty, h->atdy);
            qmap(h, h->atdy, h->qatdy);
        }
    }
    // n-side dots: b=|dx|^2, <dQw,ATdy>, |dQw|^2, <dw,dQw>, <ATdy,QATdy>
    std::vector<RedTerm> tn{RedTerm{h->dx, h->dx, nullptr, 0}};
    if (!skip_q) {
        tn.push_back(any_m ? RedTerm{h->dqw, h->atdy, nullptr, 0} : RedTerm{h->dx, h->dx, nullptr, 0});
        tn.push_back(RedTerm{h->dqw, h->dqw, nullptr, 0});
        tn.push_back(RedTerm{h->dw, h->dqw, nullptr, 0});
        tn.push_back(any_m ? RedTerm{h->atdy, h->qatdy, nullptr, 0} : RedTerm{h->dx, h->dx, nullptr, 0});
    }
    launch_reduce(st, h->seg_n, tn.data(), (int)tn.size(), region(h, 0), slot(h, SL_SIG), SL_COUNT);
    RedTerm ty{h->dy, h->dy, nullptr, 0};
    launch_reduce(st, h->seg_m, &ty, 1, region(h, 1), slot(h, SL_SIG + 7), SL_COUNT);
}

// update_sigma!, algorithm.jl:729-990, for every active instance with a restart
void update_sigma(hprsocp_handle *h, std::vector<Restart> &R, const std::vector<hprsocp_residuals> &RS,
                  const std::vector<char> &active) {
    std::vector<int> ids;
    for (int i = 0; i < h->nb; ++i)
        if (active[i] && R[i].restart_flag >= 1) ids.push_back(i);
    if (ids.empty()) return;
    const bool skip_q = !h->has_quad, qd = h->Q_is_diag;
    sigma_dots_enqueue(h);
    read_slots(h);
    // model sigma per instance
    const size_t K = ids.size();
    std::vector<double> sm(K), av(K), bv(K), cv(K), dv(K);
    std::vector<int> gold_ids;
    std::vector<double> ga, gb;
    for (size_t k = 0; k < K; ++k) {
        const int i = ids[k];
        const double *d = h->hscal + (int64_t)i * SL_COUNT + SL_SIG;
        const bool has_m = h->rng[i].m() > 0;
        double a, b, c, dl;
        ctl::sigma_terms(d, skip_q, has_m, qd, h->I(i).lambda_A, h->I(i).lambda_Q, a, b, c, dl);
        av[k] = a;
        bv[k] = b;
        cv[k] = c;
        dv[k] = dl;
        sm[k] = ctl::sigma_model(a, b, c, dl, skip_q, has_m, qd);
        if (!skip_q && has_m && qd) {  // the diagonal-Q search, over device vectors below
            gold_ids.push_back(i);
            ga.push_back(a);
            gb.push_back(b);
        }
    }
    if (!gold_ids.empty()) {
        const std::vector<double> g = golden_q_diag(h, gold_ids, ga, gb, 1e-12, 1e12, 1e-13);
        size_t q = 0;
        for (size_t k = 0; k < K; ++k)
            if (ids[k] == (q < gold_ids.size() ? gold_ids[q] : -1)) sm[k] = g[q++];
    }
    bool changed = false;
    for (size_t k = 0; k < K; ++k) {
        const int i = ids[k];
        const double sigma_old = h->I(i).sigma;
        const double sn = ctl::sigma_next(R[i], RS[i], sm[k], sigma_old);
        h->I(i).sigma = sn;
        if (fabs(sigma_old - sn) > 1e-15) changed = true;
    }
    upload_isc(h);
    if (qd && !skip_q && changed) update_q_factors(h);
}

void restart_copies_enqueue(hprsocp_handle *h);

// do_restart!, algorithm.jl:1130-1166, for the active instances with a restart flag
bool do_restart(hprsocp_handle *h, std::vector<Restart> &R, const std::vector<char> &active) {
    bool any = false;
    for (int i = 0; i < h->nb; ++i) {
        h->flag_h[i] = active[i] && R[i].restart_flag > 0;
        any = any || h->flag_h[i];
    }
    if (!any) return false;
    HIPCHK(hipMemcpyAsync(h->flag, h->flag_h, sizeof(int32_t) * h->nb, hipMemcpyHostToDevice, h->st));
    restart_copies_enqueue(h);
    h->sync();  // the flags' pinned mirror may change at the next check
    for (int i = 0; i < h->nb; ++i) {
        if (!h->flag_h[i]) continue;
        Restart &r = R[i];
        r.last_gap = r.current_gap;
        r.save_gap = INFINITY;
        r.times += 1;
        r.inner = 0;
    }
    return true;
}

// The copies of do_restart! (algorithm.jl:1130-1166) for the instances whose
// device flag is set: x, last_x <- x_bar; w, last_w <- w_bar; y, last_y <- y_bar;
// ATy, last_ATy, ATy_bar <- A^T y_bar (QP).
void restart_copies_enqueue(hprsocp_handle *h) {
    hipStream_t st = h->st;
    const SegPlan &sn = h->seg_n, &sm = h->seg_m;
    launch_copy_flagged(st, sn, h->x_bar, h->x, h->flag);
    launch_copy_flagged(st, sn, h->x_bar, h->last_x, h->flag);
    if (h->has_quad) {
        launch_copy_flagged(st, sn, h->w_bar, h->w, h->flag);
        launch_copy_flagged(st, sn, h->w_bar, h->last_w, h->flag);
    }
    if (h->m > 0) {
        launch_copy_flagged(st, sm, h->y_bar, h->y, h->flag);
        launch_copy_flagged(st, sm, h->y_bar, h->last_y, h->flag);
        if (h->has_quad) {
            launch_spmv(st, act64(h) ? h->act_AT : h->plan_AT, h->AT, h->y_bar, h->tmp_n2);
            launch_copy_flagged(st, sn, h->tmp_n2, h->aty_bar, h->flag);
            launch_copy_flagged(st, sn, h->tmp_n2, h->last_aty, h->flag);
            launch_copy_flagged(st, sn, h->tmp_n2, h->aty, h->flag);
        }
    }
}

// collect_results! (algorithm.jl:1222-1256) for instance i from the scaled
// device vectors sx, sw, sz (n) and sy (m): the maps x = b_scale D_C xbar,
// w = b_scale D_C wbar, z = c_scale D_C^{-1} zbar, y = c_scale D_R ybar,
// gathered from the instance's column / row ranges.  Any output may be null.
// Only instance i's chunks are unscaled, and the copies wait on one final
// synchronisation (stream order keeps each copy ahead of the next kernel that
// reuses the staging vector).
void fetch_unscaled(hprsocp_handle *h, int i, const double *sx, const double *sw, const double *sz, const double *sy,
                    double *x, double *y, double *z, double *w) {
    hipStream_t st = h->st;
    const InstRange &R = h->rng[i];
    const InstVal bsv = isc_field(h, &h->isc->b_scale), csv = isc_field(h, &h->isc->c_scale);
    auto sub = [&](const SegPlan &sp, const std::vector<int32_t> &first) {
        SegPlan s1 = sp;
        s1.chunks = sp.chunks + first[i];
        s1.nchunks = first[i + 1] - first[i];
        return s1;
    };
    const SegPlan sn = sub(h->seg_n, h->seg_n_first), sm = sub(h->seg_m, h->seg_m_first);
    auto fetch_n = [&](double *dst) {
        const int32_t nbx = R.box1 - R.box0, nvr = R.var1 - R.var0;
        if (nbx > 0) HIPCHK(hipMemcpyAsync(dst, h->tmp_n + R.box0, sizeof(double) * nbx, hipMemcpyDeviceToHost, st));
        if (nvr > 0)
            HIPCHK(hipMemcpyAsync(dst + nbx, h->tmp_n + R.var0, sizeof(double) * nvr, hipMemcpyDeviceToHost, st));
    };
    if (x) {
        launch_unscale(st, sn, sx, h->col_norm, bsv, 0, h->tmp_n);
        fetch_n(x);
    }
    if (w) {
        launch_unscale(st, sn, sw, h->col_norm, bsv, 0, h->tmp_n);
        fetch_n(w);
    }
    if (z) {
        launch_unscale(st, sn, sz, h->col_norm, csv, 1, h->tmp_n);
        fetch_n(z);
    }
    if (y && R.m() > 0) {
        launch_unscale(st, sm, sy, h->row_norm, csv, 0, h->tmp_m);
        const int32_t nl = R.lin1 - R.lin0, nc = R.con1 - R.con0;
        if (nl > 0) HIPCHK(hipMemcpyAsync(y, h->tmp_m + R.lin0, sizeof(double) * nl, hipMemcpyDeviceToHost, st));
        if (nc > 0) HIPCHK(hipMemcpyAsync(y + nl, h->tmp_m + R.con0, sizeof(double) * nc, hipMemcpyDeviceToHost, st));
    }
    h->sync();
}

// The best-iterate record of compute_residuals! (algorithm.jl:176-195): at
// iteration 0, or when the KKT value beats max(err_Rp, err_Rd, rel_gap) of the
// saved state, the bar iterate (scaled) and its residuals become the saved
// state.  w_bar is recorded only for problems with a quadratic term.
void track_best(hprsocp_handle *h, int64_t iter, const hprsocp_residuals &R) {
    hprsocp_handle::Saved &S = h->saved;
    if (!(iter == 0 || R.kkt < jmax(jmax(S.err_Rp, S.err_Rd), S.rel_gap))) return;
    hipStream_t st = h->st;
    const size_t bn = sizeof(double) * (size_t)h->n, bm = sizeof(double) * (size_t)h->m;
    if (bn > 0) {
        HIPCHK(hipMemcpyAsync(h->sv_x, h->x_bar, bn, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(h->sv_z, h->z_bar, bn, hipMemcpyDeviceToDevice, st));
        if (h->has_quad) HIPCHK(hipMemcpyAsync(h->sv_w, h->w_bar, bn, hipMemcpyDeviceToDevice, st));
    }
    if (bm > 0) HIPCHK(hipMemcpyAsync(h->sv_y, h->y_bar, bm, hipMemcpyDeviceToDevice, st));
    S.sigma = h->I(0).sigma;
    S.iter = iter;
    S.err_Rp = R.err_Rp;
    S.err_Rd = R.err_Rd;
    S.primal_obj = R.primal_obj;
    S.dual_obj = R.dual_obj;
    S.rel_gap = R.rel_gap;
}

// Local time as Julia's string(Dates.now()): yyyy-mm-ddTHH:MM:SS.sss
std::string timestamp_now() {
    using namespace std::chrono;
    const auto t = system_clock::now();
    const time_t tt = system_clock::to_time_t(t);
    const int ms = (int)(duration_cast<milliseconds>(t.time_since_epoch()).count() % 1000);
    struct tm tm_;
    localtime_r(&tt, &tm_);
    char b[40];
    snprintf(b, sizeof b, "%04d-%02d-%02dT%02d:%02d:%02d.%03d", tm_.tm_year + 1900, tm_.tm_mon + 1, tm_.tm_mday,
             tm_.tm_hour, tm_.tm_min, tm_.tm_sec, ms);
    return b;
}

// Names of the bc_scaling_norm_type and soc_block_scaling_strategy codes
// (the reference's Symbols, scaling.jl:181 and :625-632).
const char *bc_norm_name(int32_t k) {
    static const char *const names[] = {"l2", "rms", "linf"};
    return k >= 0 && k < 3 ? names[k] : "unknown";
}
const char *soc_strategy_name(int32_t k) {
    static const char *const names[] = {
        "max", "rms", "geom", "adaptive", "hybrid", "ruiz_geom", "role_split", "pock_safe_role_split",
        "size_split_role_split", "pre_pock_size_split", "pre_pock_huge_var_taper",
        "pre_pock_huge_var_taper_pock_geom", "pre_pock_band_split", "pre_pock_constraint_rms",
        "pre_pock_constraint_max", "pre_pock_constraint_taper", "small_cone_constraint_max",
        "count_aware_small_cone_bias", "phase_blend", "phase_taper"};
    return k >= 1 && k <= 20 ? names[k - 1] : "unknown";
}

// save_state_to_hdf5! (algorithm.jl:306-423): the current and the best
// iterate in original units, their residuals and the parameters, under the
// reference's dataset paths, in an .npz archive at params.save_filename.  A
// failed write is reported (verbose) and the solve goes on, as the
// reference's try/catch does (algorithm.jl:2872-2879).
void autosave(hprsocp_handle *h, int64_t iter, const hprsocp_residuals &R, double elapsed) {
    const hprsocp_params &p = h->params;
    const int32_t n = h->n, m = h->m;
    const bool q = h->has_quad != 0;
    std::vector<double> cx(n), cw(n), cz(n), cy(m), bx(n), bw(n), bz(n), by(m);
    fetch_unscaled(h, 0, h->x_bar, q ? h->w_bar : h->x_bar, h->z_bar, h->y_bar, cx.data(), cy.data(), cz.data(),
                   cw.data());
    fetch_unscaled(h, 0, h->sv_x, q ? h->sv_w : h->sv_x, h->sv_z, h->sv_y, bx.data(), by.data(), bz.data(),
                   bw.data());
    const hprsocp_handle::Saved &S = h->saved;
    hprsocp::NpzWriter f;
    f.add_i64("current/iteration", iter);
    f.add_f64("current/time_elapsed", elapsed);
    f.add_str("current/timestamp", timestamp_now());
    f.add_f64("current/x_org", cx.data(), n);
    f.add_f64("current/w_org", cw.data(), n);
    if (m > 0) f.add_f64("current/y_org", cy.data(), m);
    f.add_f64("current/z_org", cz.data(), n);
    f.add_f64("current/sigma", h->I(0).sigma);
    f.add_f64("current/err_Rp", R.err_Rp);
    f.add_f64("current/err_Rd", R.err_Rd);
    f.add_f64("current/primal_obj", R.primal_obj);
    f.add_f64("current/dual_obj", R.dual_obj);
    f.add_f64("current/rel_gap", R.rel_gap);
    f.add_f64("current/KKTx_and_gap", R.kkt);
    f.add_f64("best/x_org", bx.data(), n);
    f.add_f64("best/w_org", bw.data(), n);
    if (m > 0) f.add_f64("best/y_org", by.data(), m);
    f.add_f64("best/z_org", bz.data(), n);
    f.add_f64("best/sigma", S.sigma);
    f.add_i64("best/iteration", S.iter);
    f.add_f64("best/err_Rp", S.err_Rp);
    f.add_f64("best/err_Rd", S.err_Rd);
    f.add_f64("best/primal_obj", S.primal_obj);
    f.add_f64("best/dual_obj", S.dual_obj);
    f.add_f64("best/rel_gap", S.rel_gap);
    f.add_f64("best/KKTx_and_gap", jmax(jmax(S.err_Rp, S.err_Rd), S.rel_gap));
    f.add_f64("parameters/stoptol", p.stoptol);
    f.add_f64("parameters/sigma", p.sigma);
    f.add_i64("parameters/max_iter", p.max_iter);
    f.add_f64("parameters/time_limit", p.time_limit);
    f.add_i64("parameters/check_iter", p.check_iter);
    f.add_bool("parameters/warm_up", p.warm_up != 0);
    f.add_i64("parameters/print_frequency", p.print_frequency);
    f.add_i32("parameters/device_number", p.device_number);
    f.add_bool("parameters/use_Ruiz_scaling", p.use_Ruiz_scaling != 0);
    f.add_i64("parameters/ruiz_iterations", p.ruiz_iterations);
    f.add_bool("parameters/use_bc_scaling", p.use_bc_scaling != 0);
    f.add_str("parameters/bc_scaling_norm_type", bc_norm_name(p.bc_scaling_norm_type));
    f.add_bool("parameters/use_l2_scaling", p.use_l2_scaling != 0);
    f.add_bool("parameters/use_Pock_Chambolle_scaling", p.use_Pock_Chambolle_scaling != 0);
    f.add_str("parameters/soc_block_scaling_strategy", soc_strategy_name(p.soc_block_scaling_strategy));
    f.add_bool("parameters/auto_save", p.auto_save != 0);
    if (!h->init_x.empty()) f.add_f64("parameters/initial_x", h->init_x.data(), (int64_t)h->init_x.size());
    if (!h->init_y.empty()) f.add_f64("parameters/initial_y", h->init_y.data(), (int64_t)h->init_y.size());
    std::string err;
    if (!f.write(h->save_path, &err)) {
        if (p.verbose) printf("Warning: Failed to save the auto-save file: %s\n", err.c_str());
    } else if (p.verbose) {
        printf("State saved to %s at iteration %lld\n", h->save_path.c_str(), (long long)iter);
    }
}

double print_step(int64_t iter) {
    const double e = floor(log10((double)iter));
    return jmax(pow(10.0, e) / 10.0, 10.0);
}
bool rem_zero(int64_t iter, double step) { return fmod((double)iter, step) == 0.0; }

bool should_print(const hprsocp_params &p, int64_t iter, double t_start) {
    if (p.print_frequency == -1)
        return rem_zero(iter, print_step(iter)) || iter == p.max_iter || (now_s() - t_start > p.time_limit);
    return (iter % p.print_frequency == 0) || iter == p.max_iter || (now_s() - t_start > p.time_limit);
}

// The solve loop (algorithm.jl:2765-2905) over every instance of the handle in
// lockstep: one update launch per iteration serves all instances; residual
// checks, restarts and sigma updates are per instance; an instance that stops
// is retired (its tiles skip all work) and its result recorded.  res has nb
// entries.
}  // namespace

// State of the solve loop (algorithm.jl:2806-2904) between calls, so the
// loop can run in windows (hprsocp_iterate) as well as in one piece.
struct SolveLoop {
    bool live = false;  // a loop is in progress
    double t_start = 0;
    int64_t iter = 0;   // next loop body
    int nactive = 0;
    std::vector<Restart> R;
    std::vector<hprsocp_residuals> rs;
    std::vector<int64_t> it4, it6, inner, restarts32, iter32;
    std::vector<double> t4, t6, best32;
    std::vector<char> first4, first6, active, pend0, pend1;
    std::vector<int> stall32;
    // deferred M-norms: slot 0 -> last_gap candidate, slot 1 -> current_gap
    MnormCtx mn[2];
    std::vector<hprsocp_result> out;  // results of the retired instances
};

namespace {

// Everything solve() does between the setup and the first loop body.
void begin_loop(hprsocp_handle *h) {
    const hprsocp_params &p = h->params;
    if (!h->loop) h->loop = new SolveLoop();
    SolveLoop &S = *h->loop;
    S = SolveLoop{};
    S.t_start = now_s();
    delta_join(h);
    if (!h->lambda_set) hprsocp_estimate_lambda(h);
    const int nb = h->nb;
    S.R.resize(nb);
    for (int i = 0; i < nb; ++i) {
        init_restart(S.R[i]);
        S.R[i].best_sigma = h->I(i).sigma;
        h->I(i).active = 1;
    }
    h->saved = hprsocp_handle::Saved{};
    upload_isc(h);
    if (h->Q_is_diag) update_q_factors(h);
    const bool single = nb == 1;
    if (p.verbose) {
        printf("HPRSOCP SOLVER starts (MI355X HIP backend)%s...\n", single ? "" : " [batch]");
        if (single) printf(" iter     errRp     errRd         p_obj           d_obj          gap        sigma       time\n");
        else printf("batch of %d instances\n", nb);
    }
    S.rs.assign(nb, hprsocp_residuals{});
    S.it4.assign(nb, 0);
    S.it6.assign(nb, 0);
    S.inner.assign(nb, 0);
    S.t4.assign(nb, 0.0);
    S.t6.assign(nb, 0.0);
    S.first4.assign(nb, 1);
    S.first6.assign(nb, 1);
    S.active.assign(nb, 1);
    S.nactive = nb;
    S.pend0.assign(nb, 0);
    S.pend1.assign(nb, 0);
    S.out.assign(nb, hprsocp_result{});
    h->to_check = 1;
    // FP32 bulk: windows of FP32 iterations, FP64 checks; an instance continues
    // in FP64 once it reaches refine_tol or stops improving
    for (int i = 0; i < nb; ++i) h->I(i).fp32 = h->fp32 ? 1 : 0;
    upload_isc(h);
    h->fp32_dirty = true;
    h->images_valid = false;
    S.best32.assign(nb, INFINITY);
    S.stall32.assign(nb, 0);
    S.restarts32.assign(nb, 0);
    S.iter32.assign(nb, 0);
    S.live = true;
    // batches in FP64 or the FP32 delta form with no diagonal quadratic term take
    // their window-edge steps on the device (HPRSOCP_DEVICE_CONTROL=0: on the host)
    const char *env = getenv("HPRSOCP_DEVICE_CONTROL");
    h->dctl = nb > 1 && !h->fp32 && !p.auto_save && !(env && atoi(env) == 0);
    if (h->dctl) {
        if (!h->ci) {
            h->ci = h->dalloc<ctl::CtlInst>(nb);
            h->ic = h->dalloc<ctl::InstConst>(nb);
            h->ctl_status = h->dalloc<int32_t>(nb);
            h->noflag = h->dalloc<int32_t>(nb);
            HIPCHK(hipHostMalloc((void **)&h->ctl_status_h, sizeof(int32_t) * 2 * nb));
            for (auto &e : h->ctl_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&h->stage_ev, hipEventDisableTiming));
        }
        if (h->has_quad && h->Q_is_diag && !h->gwalk) {
            h->gwalk = h->dalloc<GoldenWalk>(nb);
            h->gnp = h->dalloc<int32_t>(nb);
        }
        std::vector<ctl::CtlInst> C((size_t)nb);
        std::vector<ctl::InstConst> K((size_t)nb);
        for (int i = 0; i < nb; ++i) {
            memset(&C[i], 0, sizeof(C[i]));
            ctl::init_restart(C[i].R);
            C[i].R.best_sigma = h->I(i).sigma;
            C[i].first4 = C[i].first6 = 1;
            C[i].status = HPRSOCP_STATUS_CONTINUE;
            K[i] = inst_const(h, i);
        }
        HIPCHK(hipMemcpyAsync(h->ci, C.data(), sizeof(ctl::CtlInst) * nb, hipMemcpyHostToDevice, h->st));
        HIPCHK(hipMemcpyAsync(h->ic, K.data(), sizeof(ctl::InstConst) * nb, hipMemcpyHostToDevice, h->st));
        h->sync();
    }
}

// A restart makes the body after a check a full update (to_check,
// algorithm.jl:2895-2900).  In a batch the update launches serve every
// instance, so the instances without a restart get a full update there too;
// their own solves run a partial one and keep the bar vectors of the previous
// full update.  Only a residual evaluation at a max_iter that is neither a
// check nor a print step reads those before the next full update; for that
// case the bar vectors are saved before the update and restored afterwards for
// the instances whose restart flag is 0.  y_bar is written by full updates only
// in LPs and in the delta form (FP64 QP updates write it every iteration).
static void bars_save(hprsocp_handle *h) {
    hipStream_t st = h->st;
    HIPCHK(hipMemcpyAsync(h->tmp_n, h->x_bar, sizeof(double) * h->n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(h->tmp_n2, h->z_bar, sizeof(double) * h->n, hipMemcpyDeviceToDevice, st));
    if (h->m > 0) {
        HIPCHK(hipMemcpyAsync(h->tmp_m, h->s, sizeof(double) * h->m, hipMemcpyDeviceToDevice, st));
        if (!h->has_quad || h->delta)
            HIPCHK(hipMemcpyAsync(h->rp, h->y_bar, sizeof(double) * h->m, hipMemcpyDeviceToDevice, st));
    }
}
static void bars_restore(hprsocp_handle *h) {
    hipStream_t st = h->st;
    if (!h->noflag) h->noflag = h->dalloc<int32_t>(h->nb);
    launch_flag_not(st, h->flag, h->noflag, h->nb);
    launch_copy_flagged(st, h->seg_n, h->tmp_n, h->x_bar, h->noflag);
    launch_copy_flagged(st, h->seg_n, h->tmp_n2, h->z_bar, h->noflag);
    if (h->m > 0) {
        launch_copy_flagged(st, h->seg_m, h->tmp_m, h->s, h->noflag);
        if (!h->has_quad || h->delta) launch_copy_flagged(st, h->seg_m, h->rp, h->y_bar, h->noflag);
    }
}

// compact_active for the device-controlled loop: staged copies in stream
// order, the captured windows retired (the stream may still run them).
static void compact_active_staged(hprsocp_handle *h, const std::vector<char> &keep) {
    compact_plan_staged(h, h->hp_A, h->act_A, keep, 0);
    compact_plan_staged(h, h->hp_AT, h->act_AT, keep, 1);
    if (h->nnzQ > 0) compact_plan_staged(h, h->hp_Q, h->act_Q, keep, 2);
    HIPCHK(hipEventRecord(h->stage_ev, h->st));
    h->stage_pending = true;
    h->use_act = true;
    h->retire_graphs();
}

// Print steps, and the loop bodies that start with a residual evaluation:
// checks, print steps and the last iteration (the time-limit print is
// CtlArgs::time_up).
static bool is_print(const hprsocp_params &p, int64_t t) {
    return p.print_frequency == -1 ? rem_zero(t, print_step(t)) : t % p.print_frequency == 0;
}
static bool is_event(const hprsocp_params &p, int64_t t) {
    return t % p.check_iter == 0 || t == p.max_iter || is_print(p, t);
}

// The solve loop of a batch with its window-edge steps on the device
// (k_ctl_event): per event the residual reductions (and at checks the sigma
// dots), the controller and the flag-driven restart copies are queued behind
// the previous window, the controller's status vector is copied to pinned
// memory, and the next window is queued before the host reads the previous
// event's status; the stream never waits for the host.  A check body always
// runs a full update followed by the M-norm reductions a restart would need
// (the controller consumes them only after a restart); full and partial
// updates give the same iterate, so every instance follows its host-controlled
// trajectory bit for bit in FP64.  Retirements seen one event late only delay
// the tile compaction: retired instances' tiles already skip their work.
bool run_loop_dev(hprsocp_handle *h) {
    struct DevControlOff {  // the host owns the per-instance scalars again, also after an error
        hprsocp_handle *h;
        ~DevControlOff() { h->dctl = false; }
    } off{h};
    const hprsocp_params &p = h->params;
    SolveLoop &S = *h->loop;
    const int nb = h->nb;
    const int64_t check_iter = p.check_iter;
    hipStream_t st = h->st;
    std::vector<int64_t> ev_iter;   // events whose status the host has read,
    std::vector<double> ev_time;    // and when it read them
    std::vector<std::pair<int64_t, int64_t>> windows;  // partial windows (first body, length), FP32 in the delta form
    std::vector<double> t_ret(nb, 0.0);
    // retired graphs, tagged with the event after which they were retired
    std::vector<std::pair<int64_t, size_t>> marks;
    auto poll = [&](int64_t e, int64_t it) {
        const int slot = (int)(e & 1);
        HIPCHK(hipEventSynchronize(h->ctl_ev[slot]));
        const double t = now_s() - S.t_start;
        ev_iter.push_back(it);
        ev_time.push_back(t);
        const int32_t *sv = h->ctl_status_h + (int64_t)slot * nb;
        bool retired = false;
        for (int i = 0; i < nb; ++i) {
            if (!S.active[i] || sv[i] == HPRSOCP_STATUS_CONTINUE) continue;
            S.active[i] = 0;
            h->I(i).active = 0;
            --S.nactive;
            t_ret[i] = t;
            retired = true;
        }
        // graphs retired after event e-1 ran at the latest in the window before event e
        size_t done = 0;
        while (!marks.empty() && marks.front().first <= e - 1) {
            done = marks.front().second;
            marks.erase(marks.begin());
        }
        if (done > 0) {
            for (size_t k = 0; k < done; ++k) (void)hipGraphExecDestroy(h->graphs_retired[k]);
            h->graphs_retired.erase(h->graphs_retired.begin(), h->graphs_retired.begin() + (int64_t)done);
            for (auto &mk : marks) mk.second -= done;
        }
        return retired;
    };
    CtlArgs a{};
    a.ci = h->ci;
    a.isc = h->isc;
    a.ic = h->ic;
    a.flag = h->flag;
    a.status = h->ctl_status;
    a.slots = h->dscal;
    a.slot_stride = SL_COUNT;
    a.sl_rn = SL_RN;
    a.sl_rm = SL_RM;
    a.sl_bv = SL_BV;
    a.sl_mn0 = SL_MN0;
    a.sl_mn1 = SL_MN1;
    a.sl_sig = SL_SIG;
    a.nb = nb;
    a.check_iter = check_iter;
    a.max_iter = p.max_iter;
    a.stoptol = p.stoptol;
    a.skip_q = !h->has_quad;
    a.q_diag = h->Q_is_diag;
    a.gw = h->gwalk;
    const bool gold = h->has_quad && h->Q_is_diag;  // checks run the golden-section passes between two phases
    int64_t t = S.iter, prev_t = 0, e = 0;
    bool last = false;
    while (true) {
        const bool check = t % check_iter == 0;
        last = t >= p.max_iter;
        int64_t N = t + 1;
        while (N < p.max_iter && !is_event(p, N)) ++N;
        h->kstep = 0;  // the controller sets InstScal::inner0 to the updates since the restart
        residuals_enqueue(h, t);
        if (check) sigma_dots_enqueue(h);
        a.iter = t;
        a.is_check = check;
        a.next_is_check = !last && N % check_iter == 0;
        a.time_up = now_s() - S.t_start > p.time_limit;
        if (check && gold) {
            a.phase = 1;
            launch_ctl_event(st, a);
            golden_passes_enqueue(h);
            a.phase = 2;
            launch_ctl_event(st, a);
            update_q_factors(h);  // the factors of the new sigma
        } else {
            a.phase = 0;
            launch_ctl_event(st, a);
        }
        if (check) {
            restart_copies_enqueue(h);
            h->images_valid = false;  // delta form: anchors or sigma may have changed
            h->fp32_dirty = true;
        }
        const int slot = (int)(e & 1);
        HIPCHK(hipMemcpyAsync(h->ctl_status_h + (int64_t)slot * nb, h->ctl_status, sizeof(int32_t) * nb,
                              hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(h->ctl_ev[slot], st));
        if (!last) {  // bodies t .. N-1
            // to_check (algorithm.jl:2895-2900): the body before a check or print
            // step is a full update, the body before a max_iter that is neither is
            // not, so that event's residuals read the bar vectors of the last full
            // update
            const bool n_full = N % check_iter == 0 || is_print(p, N);
            int64_t b = t;
            if (check) {  // the update a restart needs, bar vectors kept as above
                if (!n_full) bars_save(h);
                do_update(h, 1);
                if (!n_full) bars_restore(h);
                mnorm_enqueue(h, SL_MN0);
                b = t + 1;
            }
            const int64_t b_end = n_full ? N - 1 : N;  // partial bodies b .. b_end-1
            for (int64_t w = b; w < b_end;) {
                const int64_t L = std::min<int64_t>(b_end - w, check_iter);
                run_partial_updates(h, L);
                if (h->delta) windows.emplace_back(w, L);
                w += L;
            }
            if (n_full && b < N) do_update(h, 1);
            if (N % check_iter == 0) mnorm_enqueue(h, SL_MN1);
        }
        HIPCHK(hipGetLastError());
        if (e > 0 && poll(e - 1, prev_t) && S.nactive > 0) {
            compact_active_staged(h, S.active);
            marks.emplace_back(e, h->graphs_retired.size());
        }
        if (S.nactive == 0) break;
        if (last || a.time_up) {
            poll(e, t);
            break;
        }
        prev_t = t;
        t = N;
        ++e;
        if (p.verbose && check && t > 0 && (t / check_iter) % 20 == 0) {
            printf("iter %lld: %d active (device control), %.2f s\n", (long long)t, S.nactive, now_s() - S.t_start);
            fflush(stdout);
        }
    }
    h->sync();
    const double t_end = now_s() - S.t_start;
    std::vector<ctl::CtlInst> C((size_t)nb);
    HIPCHK(hipMemcpy(C.data(), h->ci, sizeof(ctl::CtlInst) * nb, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(h->isc_h, h->isc, sizeof(InstScal) * nb, hipMemcpyDeviceToHost));
    auto time_at = [&](int64_t it) {
        const auto k = std::lower_bound(ev_iter.begin(), ev_iter.end(), it) - ev_iter.begin();
        return k < (int64_t)ev_iter.size() ? ev_time[k] : t_end;
    };
    for (int i = 0; i < nb; ++i) {
        const ctl::CtlInst &c = C[i];
        if (c.status == HPRSOCP_STATUS_CONTINUE)
            throw Err{HPRSOCP_ERR_STATE, "device control: instance " + std::to_string(i) + " did not terminate"};
        const double tt = t_ret[i] > 0 ? t_ret[i] : t_end;
        const hprsocp_residuals &r = c.rs_end;
        hprsocp_result &o = S.out[i];
        memset(&o, 0, sizeof(o));
        o.iter = c.iter_end;
        o.time = tt;
        o.power_time = h->power_time;
        o.setup_time = h->setup_time;
        o.residuals = r.kkt;
        o.primal_obj = r.primal_obj;
        o.dual_obj = r.dual_obj;
        o.gap = r.rel_gap;
        o.err_Rp = r.err_Rp;
        o.err_Rp_linear = r.err_Rp_linear;
        o.err_Rp_soc = r.err_Rp_soc;
        o.err_Rd = r.err_Rd;
        o.status = c.status;
        o.time_4 = c.first4 ? tt : time_at(c.it4);
        o.iter_4 = c.it4 == 0 ? c.iter_end : c.it4;
        o.time_6 = c.first6 ? tt : time_at(c.it6);
        o.iter_6 = c.it6 == 0 ? c.iter_end : c.it6;
        o.sigma = c.sigma_end;
        o.lambda_max_A = h->I(i).lambda_A;
        o.lambda_max_Q = h->I(i).lambda_Q;
        o.restarts = (int32_t)c.restarts_end;
        int64_t n32 = 0;
        for (const auto &w : windows)
            if (w.first < c.iter_end) n32 += w.second;
        o.iter_fp32 = n32;
        S.rs[i] = r;
        S.R[i] = c.R;
    }
    if (p.verbose) {
        printf("batch done: %d instances, %lld iterations, %.3f s (device control)\n", nb, (long long)t, t_end);
        fflush(stdout);
    }
    h->use_act = false;
    h->clear_graphs();
    h->free_retired_graphs();
    h->dctl = false;  // before upload_isc below
    for (int i = 0; i < nb; ++i) {  // the handle stays usable after the solve
        h->I(i).active = 1;
        h->I(i).fp32 = 0;
    }
    upload_isc(h);
    h->images_valid = false;
    S.nactive = 0;
    S.live = false;
    return true;
}

// Runs loop bodies S.iter .. stop-1 (or to termination).  Returns true when
// every instance has terminated; S.out then holds the results.
bool run_loop(hprsocp_handle *h, int64_t stop) {
    if (h->dctl) return run_loop_dev(h);
    const hprsocp_params &p = h->params;
    SolveLoop &S = *h->loop;
    const double t_start = S.t_start;
    const int nb = h->nb;
    const bool single = nb == 1;
    const int64_t check_iter = p.check_iter;
    auto &R = S.R;
    auto &rs = S.rs;
    auto &it4 = S.it4;
    auto &it6 = S.it6;
    auto &inner = S.inner;
    auto &t4 = S.t4;
    auto &t6 = S.t6;
    auto &first4 = S.first4;
    auto &first6 = S.first6;
    auto &active = S.active;
    int &nactive = S.nactive;
    MnormCtx *mn = S.mn;
    auto &pend0 = S.pend0;
    auto &pend1 = S.pend1;
    auto &best32 = S.best32;
    auto &stall32 = S.stall32;
    auto &restarts32 = S.restarts32;
    auto &iter32 = S.iter32;
    hprsocp_result *res = S.out.data();
    for (; S.iter <= p.max_iter && S.iter < stop; ++S.iter) {
        int64_t &iter = S.iter;
        const bool print_yes = should_print(p, iter, t_start);
        const bool check = iter % check_iter == 0;
        if (check || print_yes) {
            residuals_impl(h, iter, rs.data());  // synchronises the stream
            if (p.auto_save) track_best(h, iter, rs[0]);
        }
        if (check && (mn[0].pending || mn[1].pending)) {
            // deferred M-norms (perform_iteration_step!, algorithm.jl:2731-2741); the gaps are
            // only consumed by check_restart at multiples of check_iter
            read_slots(h);
            for (int i = 0; i < nb; ++i) {
                if (!active[i]) continue;
                if (mn[0].pending && pend0[i])
                    R[i].last_gap = jmin(mnorm_finish(h, i, h->hscal + (int64_t)i * SL_COUNT + SL_MN0, mn[0]),
                                         R[i].last_gap);
                if (mn[1].pending && pend1[i])
                    R[i].current_gap = mnorm_finish(h, i, h->hscal + (int64_t)i * SL_COUNT + SL_MN1, mn[1]);
            }
            mn[0].pending = mn[1].pending = false;
        }
        const double tnow = now_s() - t_start;
        std::vector<int> status(nb, HPRSOCP_STATUS_CONTINUE);
        for (int i = 0; i < nb; ++i) {
            if (!active[i]) continue;
            if (rs[i].kkt < p.stoptol) status[i] = HPRSOCP_STATUS_OPTIMAL;
            else if (iter == p.max_iter) status[i] = HPRSOCP_STATUS_MAX_ITER;
            else if (tnow > p.time_limit) status[i] = HPRSOCP_STATUS_TIME_LIMIT;
        }
        if (h->n32 > 0 && !h->delta && iter > 0 && check) {
            // the FP64 state holds this check's iterate (the last update was a full one).
            // Stalled: no 1% improvement of the best KKT over fp32_stall_checks checks
            // that also saw two restarts (a long restart cycle alone is not a stall).
            int switched = 0;
            for (int i = 0; i < nb; ++i) {
                if (!active[i] || !h->I(i).fp32 || status[i] != HPRSOCP_STATUS_CONTINUE) continue;
                if (rs[i].kkt < 0.99 * best32[i]) {
                    best32[i] = rs[i].kkt;
                    stall32[i] = 0;
                    restarts32[i] = R[i].times;
                } else {
                    ++stall32[i];
                }
                const bool stalled = stall32[i] >= p.fp32_stall_checks && R[i].times - restarts32[i] >= 2;
                if (rs[i].kkt < p.refine_tol || stalled) {
                    h->I(i).fp32 = 0;
                    ++switched;
                    if (p.verbose && single)
                        printf("precision: FP64 iterations from iteration %lld (KKT %.3e)\n", (long long)iter, rs[i].kkt);
                }
            }
            if (switched) {
                upload_isc(h);
                if (nb > 1) compact_active(h);
                if (p.verbose && !single)
                    printf("precision: %d instance(s) continue in FP64 from iteration %lld; %d run in FP32\n", switched,
                           (long long)iter, h->n32);
            }
        }
        // as in the reference loop, restart logic runs also at an instance's last iteration
        for (int i = 0; i < nb; ++i)
            if (active[i]) check_restart(R[i], iter, check_iter, h->I(i).sigma, rs[i].kkt);
            else R[i].restart_flag = 0;
        std::vector<double> sig0(nb);
        for (int i = 0; i < nb; ++i) sig0[i] = h->I(i).sigma;
        update_sigma(h, R, rs, active);
        const bool restarted = do_restart(h, R, active);
        if (restarted) h->fp32_dirty = true;
        bool sig_changed = false;
        for (int i = 0; i < nb; ++i) sig_changed = sig_changed || h->I(i).sigma != sig0[i];
        if (restarted || sig_changed) h->images_valid = false;  // delta form: new anchors or sigma
        if (p.verbose && single && (print_yes || status[0] != HPRSOCP_STATUS_CONTINUE))
            printf("%5lld    %3.2e    %3.2e    %+7.6e    %+7.6e    %3.2e    %3.2e    %6.2f\n", (long long)iter,
                   rs[0].err_Rp, rs[0].err_Rd, rs[0].primal_obj, rs[0].dual_obj, rs[0].rel_gap, h->I(0).sigma,
                   now_s() - t_start);
        if (p.auto_save && (print_yes || status[0] != HPRSOCP_STATUS_CONTINUE))
            autosave(h, iter, rs[0], now_s() - t_start);
        bool retired = false;
        for (int i = 0; i < nb; ++i) {
            if (!active[i] || !(check || print_yes)) continue;
            if (rs[i].kkt < 1e-4 && first4[i]) {
                t4[i] = now_s() - t_start;
                it4[i] = iter;
                first4[i] = 0;
                if (p.verbose && single) printf("KKT < 1e-4 at iter = %lld\n", (long long)iter);
            }
            if (rs[i].kkt < 1e-6 && first6[i]) {
                t6[i] = now_s() - t_start;
                it6[i] = iter;
                first6[i] = 0;
                if (p.verbose && single) printf("KKT < 1e-6 at iter = %lld\n", (long long)iter);
            }
            if (status[i] == HPRSOCP_STATUS_CONTINUE) continue;
            const double tt = now_s() - t_start;
            hprsocp_result &o = res[i];
            memset(&o, 0, sizeof(o));
            o.iter = iter;
            o.time = tt;
            o.power_time = h->power_time;
            o.setup_time = h->setup_time;
            o.residuals = rs[i].kkt;
            o.primal_obj = rs[i].primal_obj;
            o.dual_obj = rs[i].dual_obj;
            o.gap = rs[i].rel_gap;
            o.err_Rp = rs[i].err_Rp;
            o.err_Rp_linear = rs[i].err_Rp_linear;
            o.err_Rp_soc = rs[i].err_Rp_soc;
            o.err_Rd = rs[i].err_Rd;
            o.status = status[i];
            o.time_4 = t4[i] == 0.0 ? tt : t4[i];
            o.iter_4 = it4[i] == 0 ? iter : it4[i];
            o.time_6 = t6[i] == 0.0 ? tt : t6[i];
            o.iter_6 = it6[i] == 0 ? iter : it6[i];
            o.sigma = h->I(i).sigma;
            o.lambda_max_A = h->I(i).lambda_A;
            o.lambda_max_Q = h->I(i).lambda_Q;
            o.restarts = (int32_t)R[i].times;
            o.iter_fp32 = iter32[i];
            active[i] = 0;
            h->I(i).active = 0;
            --nactive;
            retired = true;
            if (p.verbose && single) {
                printf("SOLUTION SUMMARY: status %d, iterations %lld, time %.3f s, primal %.12e, KKT %.3e\n",
                       status[i], (long long)iter, tt, rs[i].primal_obj, rs[i].kkt);
                fflush(stdout);
            }
        }
        if (nactive == 0) {
            h->sync();
            if (p.verbose && !single) {
                printf("batch done: %d instances, %lld iterations, %.3f s\n", nb, (long long)iter, now_s() - t_start);
                fflush(stdout);
            }
            h->use_act = false;
            h->clear_graphs();
            for (int i = 0; i < nb; ++i) {  // the handle stays usable after the solve
                h->I(i).active = 1;
                h->I(i).fp32 = h->fp32 ? 1 : 0;
            }
            upload_isc(h);
            S.live = false;
            return true;
        }
        if (p.verbose && !single && check && iter > 0 && (iter / check_iter) % 20 == 0) {
            double worst = 0;
            for (int i = 0; i < nb; ++i)
                if (active[i]) worst = jmax(worst, rs[i].kkt);
            printf("iter %lld: %d active, worst KKT %.3e, %.2f s\n", (long long)iter, nactive, worst, now_s() - t_start);
        }
        if (retired) upload_isc(h);
        if (retired && nb > 1) compact_active(h);  // launch only the tiles of instances still running
        // Halpern counters: set once per check window (restarts happen only at checks)
        if (check || retired) {
            for (int i = 0; i < nb; ++i) inner[i] = R[i].inner;
            set_inner(h, inner);
        }
        const int64_t nx = iter + 1;
        bool tc = (nx % check_iter == 0) || restarted;
        if (p.print_frequency == -1) tc = tc || rem_zero(nx, print_step(nx));
        else tc = tc || (nx % p.print_frequency == 0);
        // perform_iteration_step!, algorithm.jl:2722-2744.  A partial update is
        // followed by iterations whose bodies are plain (no residual check, no
        // print step): their partial updates run as one captured graph.
        if (!tc) {
            int64_t L = 1;
            auto plain = [&](int64_t t) {
                if (t >= p.max_iter || t % check_iter == 0) return false;
                return p.print_frequency == -1 ? !rem_zero(t, print_step(t)) : (t % p.print_frequency != 0);
            };
            while (plain(nx + L) && L < check_iter && iter + L < stop) ++L;
            for (int i = 0; i < nb; ++i)
                if (active[i] && (h->I(i).fp32 || h->delta)) iter32[i] += L;
            run_partial_updates(h, L);
            for (int i = 0; i < nb; ++i)
                if (active[i]) R[i].inner += L - 1;
            iter += L - 1;
        } else {
            for (int i = 0; i < nb; ++i)
                if (active[i] && h->I(i).fp32 && !h->delta) ++iter32[i];
            // a full update only for the restarted instances when the next residual
            // evaluation is a max_iter that is neither a check nor a print step
            bool keep = nb > 1 && restarted;
            for (int64_t q = nx; keep && q <= p.max_iter; ++q)
                if (q % check_iter == 0 || is_print(p, q)) keep = false;
            if (keep) bars_save(h);
            do_update(h, 1);
            if (keep) bars_restore(h);
        }
        const bool need_cur = (nx % check_iter) == 0;
        if (restarted) {
            mnorm_enqueue(h, SL_MN0);
            mn[0].snap(h);
            for (int i = 0; i < nb; ++i) pend0[i] = h->flag_h[i];
        }
        if (need_cur) {
            mnorm_enqueue(h, SL_MN1);
            mn[1].snap(h);
            for (int i = 0; i < nb; ++i) pend1[i] = active[i];
        }
        for (int i = 0; i < nb; ++i)
            if (active[i]) R[i].inner += 1;
    }
    return false;
}

// Result of a loop still in progress (single problem): the last check's
// residuals and the current counters, status CONTINUE.
void progress_result(hprsocp_handle *h, hprsocp_result *o) {
    const SolveLoop &S = *h->loop;
    const hprsocp_residuals &r = S.rs[0];
    memset(o, 0, sizeof(*o));
    o->iter = S.iter;
    o->time = now_s() - S.t_start;
    o->power_time = h->power_time;
    o->setup_time = h->setup_time;
    o->residuals = r.kkt;
    o->primal_obj = r.primal_obj;
    o->dual_obj = r.dual_obj;
    o->gap = r.rel_gap;
    o->err_Rp = r.err_Rp;
    o->err_Rp_linear = r.err_Rp_linear;
    o->err_Rp_soc = r.err_Rp_soc;
    o->err_Rd = r.err_Rd;
    o->status = HPRSOCP_STATUS_CONTINUE;
    o->iter_4 = S.it4[0];
    o->iter_6 = S.it6[0];
    o->time_4 = S.t4[0];
    o->time_6 = S.t6[0];
    o->sigma = h->I(0).sigma;
    o->lambda_max_A = h->I(0).lambda_A;
    o->lambda_max_Q = h->I(0).lambda_Q;
    o->restarts = (int32_t)S.R[0].times;
    o->iter_fp32 = S.iter32[0];
}

void solve_impl(hprsocp_handle *h, hprsocp_result *res) {
    begin_loop(h);
    run_loop(h, INT64_MAX);
    std::copy(h->loop->out.begin(), h->loop->out.end(), res);
}

}  // namespace

void free_loop(SolveLoop *s) { delete s; }

extern "C" {

int hprsocp_mnorm(hprsocp_handle *h, double *out) {
    ABI_TRY(h, {
        delta_join(h);
        mnorm_enqueue(h, SL_MN0);
        read_slots(h);
        MnormCtx c;
        c.snap(h);
        *out = mnorm_finish(h, 0, h->hscal + SL_MN0, c);
    });
    return 0;
}

int hprsocp_solve(hprsocp_handle *h, hprsocp_result *res) {
    ABI_TRY(h, { solve_impl(h, res); });
    return 0;
}

int hprsocp_iterate(hprsocp_handle *h, int64_t k, int32_t flags, hprsocp_result *res) {
    ABI_TRY(h, {
        if (h->nb != 1) throw Err{HPRSOCP_ERR_UNSUPPORTED, "hprsocp_iterate: single-problem handles only"};
        if (k < 0) throw Err{HPRSOCP_ERR_INVALID, "hprsocp_iterate: k must be >= 0"};
        if (!res) throw Err{HPRSOCP_ERR_INVALID, "hprsocp_iterate: res is NULL"};
        if ((flags & HPRSOCP_ITERATE_RESET) || !h->loop || !h->loop->live) begin_loop(h);
        SolveLoop &S = *h->loop;
        const int64_t stop = k > INT64_MAX - S.iter ? INT64_MAX : S.iter + k;
        if (run_loop(h, stop)) *res = S.out[0];
        else progress_result(h, res);
    });
    return 0;
}

void hprsocp_init_restart(hprsocp_restart_info *r) {
    if (r) init_restart(*r);
}

void hprsocp_check_restart(hprsocp_restart_info *r, int64_t iter, int64_t check_iter, double sigma, double kkt) {
    if (r && check_iter > 0) check_restart(*r, iter, check_iter, sigma, kkt);
}

int hprsocp_update_sigma(hprsocp_handle *h, hprsocp_restart_info *r, const hprsocp_residuals *res) {
    ABI_TRY(h, {
        if (h->nb != 1) throw Err{HPRSOCP_ERR_UNSUPPORTED, "hprsocp_update_sigma: single-problem handles only"};
        if (!r || !res) throw Err{HPRSOCP_ERR_INVALID, "hprsocp_update_sigma: NULL argument"};
        if (r->restart_flag >= 1) {
            delta_join(h);
            std::vector<Restart> R{*r};
            std::vector<hprsocp_residuals> RS{*res};
            const double s0 = h->I(0).sigma;
            update_sigma(h, R, RS, std::vector<char>{1});
            if (h->I(0).sigma != s0) h->images_valid = false;
            *r = R[0];
        }
    });
    return 0;
}

int hprsocp_restart(hprsocp_handle *h, hprsocp_restart_info *r) {
    ABI_TRY(h, {
        if (h->nb != 1) throw Err{HPRSOCP_ERR_UNSUPPORTED, "hprsocp_restart: single-problem handles only"};
        if (!r) throw Err{HPRSOCP_ERR_INVALID, "hprsocp_restart: NULL argument"};
        if (r->restart_flag > 0) {
            delta_join(h);
            std::vector<Restart> R{*r};
            if (do_restart(h, R, std::vector<char>{1})) {
                h->fp32_dirty = true;
                h->images_valid = false;
            }
            *r = R[0];
        }
    });
    return 0;
}

}  // extern "C"

namespace {
// collect_results! (algorithm.jl:1222-1256) for instance i: unscaled x_bar,
// y_bar, z_bar and w_bar (x_bar for problems without a quadratic term).
void get_solution_impl(hprsocp_handle *h, int i, double *x, double *y, double *z, double *w) {
    fetch_unscaled(h, i, h->x_bar, h->has_quad ? h->w_bar : h->x_bar, h->z_bar, h->y_bar, x, y, z, w);
}
}  // namespace

extern "C" {

int hprsocp_get_solution(hprsocp_handle *h, double *x, double *y, double *z, double *w) {
    ABI_TRY(h, { get_solution_impl(h, 0, x, y, z, w); });
    return 0;
}

int hprsocp_optimize(const hprsocp_problem *prob, const hprsocp_params *params, hprsocp_result *res, double *x,
                     double *y, double *z, double *w) {
    hprsocp_params p;
    if (params) p = *params;
    else hprsocp_default_params(&p);
    if (p.warm_up) {  // optimize(): a 200-iteration warm-up solve first (algorithm.jl:2540-2568)
        hprsocp_params q = p;
        q.max_iter = 200;
        q.verbose = 0;
        q.warm_up = 0;
        hprsocp_handle *hw = nullptr;
        int rc = hprsocp_create(prob, &q, &hw);
        if (rc) return rc;
        hprsocp_result tmp;
        rc = hprsocp_estimate_lambda(hw);
        if (!rc && (q.initial_x || q.initial_y)) rc = hprsocp_set_initial(hw, q.initial_x, q.initial_y);
        if (!rc) rc = hprsocp_solve(hw, &tmp);
        if (rc) g_create_error = hw->err;
        hprsocp_destroy(hw);
        if (rc) return rc;
    }
    hprsocp_handle *h = nullptr;
    int rc = hprsocp_create(prob, &p, &h);
    if (rc) return rc;
    rc = hprsocp_estimate_lambda(h);
    if (!rc && (p.initial_x || p.initial_y)) rc = hprsocp_set_initial(h, p.initial_x, p.initial_y);
    if (!rc) rc = hprsocp_solve(h, res);
    if (!rc) rc = hprsocp_get_solution(h, x, y, z, w);
    if (rc) g_create_error = h->err;
    hprsocp_destroy(h);
    return rc;
}

static double *vec_ptr(hprsocp_handle *h, int32_t id, int64_t *len) {
    const int64_t n = h->n, m = h->m;
    switch (id) {
    case HPRSOCP_V_X: *len = n; return h->x;
    case HPRSOCP_V_X_BAR: *len = n; return h->x_bar;
    case HPRSOCP_V_X_HAT: *len = n; return h->x_hat;
    case HPRSOCP_V_Z_BAR: *len = n; return h->z_bar;
    case HPRSOCP_V_DX: *len = n; return h->dx;
    case HPRSOCP_V_LAST_X: *len = n; return h->last_x;
    case HPRSOCP_V_Y: *len = m; return h->y;
    case HPRSOCP_V_Y_BAR: *len = m; return h->y_bar;
    case HPRSOCP_V_DY: *len = m; return h->dy;
    case HPRSOCP_V_LAST_Y: *len = m; return h->last_y;
    case HPRSOCP_V_S: *len = m; return h->s;
    case HPRSOCP_V_W: *len = n; return h->w;
    case HPRSOCP_V_W_BAR: *len = n; return h->w_bar;
    case HPRSOCP_V_LAST_W: *len = n; return h->last_w;
    case HPRSOCP_V_DW: *len = n; return h->dw;
    case HPRSOCP_V_ATY: *len = n; return h->aty;
    case HPRSOCP_V_ATY_BAR: *len = n; return h->aty_bar;
    case HPRSOCP_V_LAST_ATY: *len = n; return h->last_aty;
    case HPRSOCP_V_C: *len = n; return h->c;
    case HPRSOCP_V_L: *len = n; return h->l;
    case HPRSOCP_V_U: *len = n; return h->u;
    case HPRSOCP_V_AL: *len = m; return h->AL;
    case HPRSOCP_V_AU: *len = m; return h->AU;
    case HPRSOCP_V_SOC_RHS_FULL: *len = m; return h->rhs;
    case HPRSOCP_V_ROW_NORM: *len = m; return h->row_norm;
    case HPRSOCP_V_COL_NORM: *len = n; return h->col_norm;
    case HPRSOCP_V_DIAG_Q: *len = n; return h->diag_Q;
    case HPRSOCP_V_RP: *len = m; return h->rp;
    case HPRSOCP_V_AX: *len = m; return h->ax;
    case HPRSOCP_V_RD: *len = n; return h->rd;
    default: *len = 0; return nullptr;
    }
}

int hprsocp_get_vector(hprsocp_handle *h, int32_t id, double *out) {
    int64_t len;
    double *p = vec_ptr(h, id, &len);
    if (!p) {
        h->err = "unknown vector id";
        return HPRSOCP_ERR_INVALID;
    }
    ABI_TRY(h, {
        delta_join(h);
        if (len > 0) HIPCHK(hipMemcpyAsync(out, p, sizeof(double) * len, hipMemcpyDeviceToHost, h->st));
        h->sync();
    });
    return 0;
}

int hprsocp_set_vector(hprsocp_handle *h, int32_t id, const double *in) {
    int64_t len;
    double *p = vec_ptr(h, id, &len);
    if (!p) {
        h->err = "unknown vector id";
        return HPRSOCP_ERR_INVALID;
    }
    ABI_TRY(h, {
        delta_join(h);
        if (len > 0) HIPCHK(hipMemcpyAsync(p, in, sizeof(double) * len, hipMemcpyHostToDevice, h->st));
        h->fp32_dirty = true;
        h->images_valid = false;
        h->sync();
    });
    return 0;
}

int hprsocp_get_scalars(hprsocp_handle *h, hprsocp_scalars *o) {
    o->sigma = h->I(0).sigma;
    o->lambda_max_A = h->I(0).lambda_A;
    o->lambda_max_Q = h->I(0).lambda_Q;
    o->b_scale = h->I(0).b_scale;
    o->c_scale = h->I(0).c_scale;
    o->norm_b = h->ih[0].norm_b;
    o->norm_c = h->ih[0].norm_c;
    o->norm_b_org = h->ih[0].norm_b_org;
    o->norm_c_org = h->ih[0].norm_c_org;
    o->Q_is_diag = h->Q_is_diag;
    o->has_quadratic = h->has_quad;
    return 0;
}

int hprsocp_set_scalars(hprsocp_handle *h, double sigma, double la, double lq) {
    ABI_TRY(h, {
        delta_join(h);
        for (int i = 0; i < h->nb; ++i) {
            if (sigma >= 0) h->I(i).sigma = sigma;
            if (la >= 0) h->I(i).lambda_A = la;
            if (lq >= 0) h->I(i).lambda_Q = lq;
        }
        upload_isc(h);
        h->lambda_set = true;
        if (h->Q_is_diag) update_q_factors(h);
        h->fp32_dirty = true;
        h->images_valid = false;
        h->sync();
    });
    return 0;
}

int hprsocp_get_scaled_A(hprsocp_handle *h, double *val) {
    ABI_TRY(h, {
        if (h->nnzA > 0) HIPCHK(hipMemcpyAsync(val, h->A_v, sizeof(double) * h->nnzA, hipMemcpyDeviceToHost, h->st));
        h->sync();
    });
    return 0;
}

int hprsocp_plan_info(hprsocp_handle *h, char *buf, int64_t cap, int64_t *len) {
    ABI_TRY(h, {
        std::string j = std::string("{\"qw_primal_fused\": ") + (h->fuse_qw ? "1" : "0") + ", \"operators\": [";
        char b[1024];
        // forms forced by the environment are not timed (infinite times): reported as 0
        auto fin = [](float v) { return std::isfinite(v) ? (double)v : 0.0; };
        for (size_t k = 0; k < h->tune.size(); ++k) {
            const auto &r = h->tune[k];
            snprintf(b, sizeof b,
                     "%s{\"name\": \"%s\", \"form\": \"%s\", \"tuned\": %s, \"cached\": %s, "
                     "\"nnz_target\": %d, \"g_divisor\": %d, \"huge_target\": %d, \"huge_whole\": %d, "
                     "\"stream\": %d, \"tiles\": %d, \"candidates\": %d, \"ms_default\": %.5f, "
                     "\"ms\": %.5f, \"ms_form\": %.5f, \"block\": %d, \"form32\": \"%s\", \"block32\": %d, "
                     "\"ms32\": %.5f}",
                     k ? ", " : "", r.name.c_str(), r.form.c_str(), r.tuned ? "true" : "false",
                     r.cached ? "true" : "false", r.cfg.nnz_target, r.cfg.g_divisor, r.cfg.huge_target,
                     r.cfg.huge_whole, (int)r.cfg.stream, r.tiles, r.candidates, fin(r.ms_default), fin(r.ms),
                     fin(r.ms_form), r.block, r.form32.c_str(), r.block32, fin(r.ms32));
            j += b;
        }
        j += "]}";
        if (len) *len = (int64_t)j.size();
        if (buf && cap > 0) {
            const size_t c = std::min<size_t>((size_t)cap - 1, j.size());
            memcpy(buf, j.data(), c);
            buf[c] = 0;
        }
    });
    return 0;
}

int hprsocp_get_dims(hprsocp_handle *h, int32_t *m, int32_t *n, int64_t *nnz) {
    *m = h->m;
    *n = h->n;
    *nnz = h->nnzA;
    return 0;
}

int hprsocp_spmv(hprsocp_handle *h, int32_t trans, const double *in, double *out) {
    ABI_TRY(h, {
        const int64_t lin = trans ? h->m : h->n, lout = trans ? h->n : h->m;
        double *din = trans ? h->tmp_m : h->tmp_n;
        double *dout = trans ? h->tmp_n2 : h->ax;
        if (lin > 0) HIPCHK(hipMemcpyAsync(din, in, sizeof(double) * lin, hipMemcpyHostToDevice, h->st));
        if (trans) launch_spmv(h->st, h->plan_AT, h->AT, din, dout);
        else launch_spmv(h->st, h->plan_A, h->A, din, dout);
        if (lout > 0) HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * lout, hipMemcpyDeviceToHost, h->st));
        h->sync();
    });
    return 0;
}

// Runs `iters` partial HPR iterations (inner = 0, 1, ...) and times them with HIP
// events on the handle's stream; ms_kernel receives the mean time of the
// dominant fused SpMV+projection kernel(s) from a second, event-bracketed pass.
int hprsocp_time_iterations(hprsocp_handle *h, int32_t iters, double *ms_per_iter, double *ms_kernel) {
    if (!h->lambda_set) {
        h->err = "estimate_lambda must run before timing";
        return HPRSOCP_ERR_STATE;
    }
    ABI_TRY(h, {
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        for (int i = 0; i < h->nb; ++i) h->I(i).active = 1;
        set_inner(h, std::vector<int64_t>(h->nb, 0));
        // the FP32 copies of a changed FP64 state are made before the timed
        // region (a solve makes them once per check window)
        if (h->delta) {
            if (h->fp32_dirty) delta_split(h);
            delta_prec(h, 1);
        } else if (h->n32 > 0 && h->fp32_dirty) {
            demote_state(h);
        }
        HIPCHK(hipEventRecord(e0, h->st));
        for (int k = 0; k < iters; ++k) do_update(h, 0);
        HIPCHK(hipEventRecord(e1, h->st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_iter = iters > 0 ? ms / iters : 0.0;
        if (ms_kernel) {
            if (h->ev.empty()) {
                h->ev.resize(6);
                for (auto &e : h->ev) HIPCHK(hipEventCreate(&e));
            }
            struct ProfileOn {  // event records in do_update_t only for these eager launches
                hprsocp_handle *h;
                explicit ProfileOn(hprsocp_handle *hh) : h(hh) { h->profile = true; }
                ~ProfileOn() { h->profile = false; }
            } guard(h);
            double acc[6] = {0, 0, 0, 0, 0, 0};
            set_inner(h, std::vector<int64_t>(h->nb, 0));
            for (int k = 0; k < iters; ++k) {
                do_update(h, 0);
                HIPCHK(hipEventSynchronize(h->ev[h->has_quad ? 5 : 2]));
                for (int j = 0; j + 1 < (h->has_quad ? 6 : 3); ++j) {
                    float t = 0;
                    HIPCHK(hipEventElapsedTime(&t, h->ev[j], h->ev[j + 1]));
                    acc[j] += t;
                }
            }
            for (int j = 0; j < 6; ++j) ms_kernel[j] = iters > 0 ? acc[j] / iters : 0.0;
        }
        if (h->n32 > 0 && !h->delta) promote_state(h);  // (the delta form's mirrors hold differences)
        delta_join(h);
        h->sync();
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
    return 0;
}


// =================================================================== batch
struct hprsocp_batch {
    hprsocp_handle *h = nullptr;
};

}  // extern "C"

namespace {
// Block-diagonal union of independent problems in the canonical layout (all
// linear rows, then all SOC rows; all boxed columns, then all SOC columns).
struct Combined {
    hprsocp_problem p{};
    std::vector<int32_t> rp, qrp, con_idx, var_idx;
    RawVec<int32_t> ci, qci;
    RawVec<double> v, qv;
    std::vector<double> c, AL, AU, l, u, rhs;
    std::vector<InstRange> rng;
    std::vector<double> objc;
};

void concat_problems(const hprsocp_problem *const *probs, int32_t count, Combined &C) {
    int64_t M1 = 0, MC = 0, N1 = 0, NC = 0, KC = 0, KV = 0;
    const bool has_q = probs[0] && probs[0]->Q_kind == HPRSOCP_Q_CSR && probs[0]->Q_nnz > 0;
    // the CSR scans first, in parallel; the checks below report them in order
    std::vector<int> rcA(count, 0), rcQ(count, 0);
    std::vector<std::string> errA(count), errQ(count);
    host_parallel(count, [&](int64_t i) {
        const hprsocp_problem *q = probs[i];
        if (!q || q->m < 0 || q->n < 0) return;
        rcA[i] = validate_csr(errA[i], "A", q->m, q->n, q->A_nnz, q->A_rowptr, q->A_colidx, q->A_val);
        if (q->Q_kind == HPRSOCP_Q_CSR && q->Q_nnz > 0)
            rcQ[i] = validate_csr(errQ[i], "Q", q->n, q->n, q->Q_nnz, q->Q_rowptr, q->Q_colidx, q->Q_val);
    });
    for (int32_t i = 0; i < count; ++i) {
        const hprsocp_problem *q = probs[i];
        char b[200];
        if (!q) throw Err{HPRSOCP_ERR_INVALID, "null problem in batch"};
        snprintf(b, sizeof b, "batch instance %d: ", i);
        const std::string pre(b);
        if (q->m < 0 || q->n < 0) throw Err{HPRSOCP_ERR_INVALID, pre + "negative dimensions"};
        if (rcA[i]) throw Err{rcA[i], pre + errA[i]};
        if (q->Q_kind != HPRSOCP_Q_CSR) throw Err{HPRSOCP_ERR_UNSUPPORTED, pre + "batches take CSR Q only"};
        if ((q->Q_nnz > 0) != has_q)
            throw Err{HPRSOCP_ERR_INVALID, pre + "all instances of a batch must have a quadratic term, or none"};
        if (q->Q_nnz > 0 && rcQ[i]) throw Err{rcQ[i], pre + errQ[i]};
        if (q->number_linear_con < 0 || q->number_linear_con > q->m || q->number_lu_x < 0 || q->number_lu_x > q->n)
            throw Err{HPRSOCP_ERR_INVALID, pre + "number_linear_con / number_lu_x out of range"};
        if (!q->c || !q->l || !q->u || (q->m > 0 && (!q->AL || !q->AU)))
            throw Err{HPRSOCP_ERR_INVALID, pre + "missing vector"};
        if ((q->n_soc_con > 0 && !q->soc_con_idx) || (q->n_soc_var > 0 && !q->soc_var_idx) || q->n_soc_con < 0 ||
            q->n_soc_var < 0)
            throw Err{HPRSOCP_ERR_INVALID, pre + "missing SOC index vector"};
        if (q->soc_con_idx && (q->soc_con_idx[0] != q->number_linear_con || q->soc_con_idx[q->n_soc_con] != q->m))
            throw Err{HPRSOCP_ERR_INVALID, pre + "soc_con_idx must start at number_linear_con and end at m"};
        if (q->soc_var_idx && (q->soc_var_idx[0] != q->number_lu_x || q->soc_var_idx[q->n_soc_var] != q->n))
            throw Err{HPRSOCP_ERR_INVALID, pre + "soc_var_idx must start at number_lu_x and end at n"};
        M1 += q->number_linear_con;
        MC += q->m - q->number_linear_con;
        N1 += q->number_lu_x;
        NC += q->n - q->number_lu_x;
        KC += q->n_soc_con;
        KV += q->n_soc_var;
    }
    if (M1 + MC > INT32_MAX || N1 + NC > INT32_MAX) throw Err{HPRSOCP_ERR_INVALID, "batch too large for 32-bit indices"};
    const int32_t M = (int32_t)(M1 + MC), N = (int32_t)(N1 + NC);
    C.rng.resize(count);
    C.objc.resize(count);
    {
        int32_t l0 = 0, c0 = (int32_t)M1, b0 = 0, v0 = (int32_t)N1, k0 = 0, kv0 = 0;
        for (int32_t i = 0; i < count; ++i) {
            const hprsocp_problem *q = probs[i];
            InstRange &R = C.rng[i];
            R.lin0 = l0;
            R.lin1 = l0 += q->number_linear_con;
            R.con0 = c0;
            R.con1 = c0 += q->m - q->number_linear_con;
            R.box0 = b0;
            R.box1 = b0 += q->number_lu_x;
            R.var0 = v0;
            R.var1 = v0 += q->n - q->number_lu_x;
            R.kc0 = k0;
            R.kc1 = k0 += q->n_soc_con;
            R.kv0 = kv0;
            R.kv1 = kv0 += q->n_soc_var;
            C.objc[i] = q->obj_constant;
        }
    }
    auto gcol = [&](int32_t i, int32_t c) {
        const int32_t n1 = probs[i]->number_lu_x;
        return c < n1 ? C.rng[i].box0 + c : C.rng[i].var0 + (c - n1);
    };
    // Rows of instance i's pass 0 (linear rows of A / boxed columns of Q) and
    // pass 1 (SOC rows / SOC columns) land at precomputed offsets, so the
    // 2 x count segments fill in parallel.
    auto pack = [&](bool q_mat, int64_t nrows_total, std::vector<int32_t> &rp, RawVec<int32_t> &ci,
                    RawVec<double> &v) {
        const int64_t S = 2 * (int64_t)count;
        std::vector<int64_t> row_off(S + 1, 0), nnz_off(S + 1, 0);
        auto seg = [&](int64_t s, int32_t &r0, int32_t &r1, const int32_t *&crp, const int32_t *&cci,
                       const double *&cv) {
            const hprsocp_problem *q = probs[s % count];
            const bool pass = s >= count;
            const int32_t split = q_mat ? q->number_lu_x : q->number_linear_con, end = q_mat ? q->n : q->m;
            r0 = pass ? split : 0;
            r1 = pass ? end : split;
            crp = q_mat ? q->Q_rowptr : q->A_rowptr;
            cci = q_mat ? q->Q_colidx : q->A_colidx;
            cv = q_mat ? q->Q_val : q->A_val;
        };
        for (int64_t s = 0; s < S; ++s) {
            int32_t r0, r1;
            const int32_t *crp, *cci;
            const double *cv;
            seg(s, r0, r1, crp, cci, cv);
            row_off[s + 1] = row_off[s] + (r1 - r0);
            nnz_off[s + 1] = nnz_off[s] + (crp ? crp[r1] - crp[r0] : 0);
        }
        if (nnz_off[S] > INT32_MAX) throw Err{HPRSOCP_ERR_INVALID, "batch too large for 32-bit indices"};
        rp.assign((size_t)nrows_total + 1, 0);
        ci.resize((size_t)nnz_off[S]);
        v.resize((size_t)nnz_off[S]);
        host_parallel(S, [&](int64_t s) {
            int32_t r0, r1;
            const int32_t *crp, *cci;
            const double *cv;
            seg(s, r0, r1, crp, cci, cv);
            const int32_t i = (int32_t)(s % count);
            const int64_t base = crp ? crp[r0] : 0, o = nnz_off[s] - base;
            for (int32_t r = r0; r < r1; ++r) rp[row_off[s] + (r - r0) + 1] = (int32_t)(o + (crp ? crp[r + 1] : 0));
            if (crp)
                for (int64_t k = crp[r0]; k < crp[r1]; ++k) {
                    ci[o + k] = gcol(i, cci[k]);
                    v[o + k] = cv[k];
                }
        });
    };
    pack(false, M, C.rp, C.ci, C.v);
    if (has_q) pack(true, N, C.qrp, C.qci, C.qv);
    C.c.resize(N);
    C.l.resize(N);
    C.u.resize(N);
    C.AL.resize(M);
    C.AU.resize(M);
    C.rhs.assign((size_t)MC, 0.0);
    for (int32_t i = 0; i < count; ++i) {
        const hprsocp_problem *q = probs[i];
        for (int32_t c = 0; c < q->n; ++c) {
            const int32_t g = gcol(i, c);
            C.c[g] = q->c[c];
            C.l[g] = q->l[c];
            C.u[g] = q->u[c];
        }
        for (int32_t r = 0; r < q->m; ++r) {
            const int32_t g = r < q->number_linear_con ? C.rng[i].lin0 + r : C.rng[i].con0 + (r - q->number_linear_con);
            C.AL[g] = q->AL[r];
            C.AU[g] = q->AU[r];
        }
        if (q->soc_rhs)
            for (int32_t r = q->number_linear_con; r < q->m; ++r)
                C.rhs[C.rng[i].con0 - M1 + (r - q->number_linear_con)] = q->soc_rhs[r - q->number_linear_con];
        for (int32_t k = 0; k < q->n_soc_con; ++k)
            C.con_idx.push_back(C.rng[i].con0 + (q->soc_con_idx[k] - q->number_linear_con));
        for (int32_t k = 0; k < q->n_soc_var; ++k)
            C.var_idx.push_back(C.rng[i].var0 + (q->soc_var_idx[k] - q->number_lu_x));
    }
    C.con_idx.push_back(M);
    C.var_idx.push_back(N);
    hprsocp_problem &P = C.p;
    P.m = M;
    P.n = N;
    P.A_nnz = (int64_t)C.ci.size();
    P.A_rowptr = C.rp.data();
    P.A_colidx = C.ci.data();
    P.A_val = C.v.data();
    P.Q_kind = HPRSOCP_Q_CSR;
    P.Q_nnz = (int64_t)C.qci.size();
    P.Q_rowptr = has_q ? C.qrp.data() : nullptr;
    P.Q_colidx = C.qci.data();
    P.Q_val = C.qv.data();
    P.c = C.c.data();
    P.AL = C.AL.data();
    P.AU = C.AU.data();
    P.l = C.l.data();
    P.u = C.u.data();
    P.soc_rhs = C.rhs.data();
    P.number_linear_con = (int32_t)M1;
    P.number_lu_x = (int32_t)N1;
    P.n_soc_con = (int32_t)KC;
    P.soc_con_idx = C.con_idx.data();
    P.n_soc_var = (int32_t)KV;
    P.soc_var_idx = C.var_idx.data();
    P.obj_constant = 0.0;
}
}  // namespace

extern "C" {

int hprsocp_batch_create(const hprsocp_problem *const *probs, int32_t count, const hprsocp_params *params,
                         hprsocp_batch **out) {
    *out = nullptr;
    g_create_error.clear();
    if (!probs || count <= 0) {
        g_create_error = "empty batch";
        return HPRSOCP_ERR_INVALID;
    }
    if (params && (params->initial_x || params->initial_y)) {
        g_create_error = "batches take initial points through hprsocp_batch_set_initial";
        return HPRSOCP_ERR_INVALID;
    }
    hprsocp_batch *b = new (std::nothrow) hprsocp_batch();
    if (!b) return HPRSOCP_ERR_ALLOC;
    b->h = new (std::nothrow) hprsocp_handle();
    if (!b->h) {
        delete b;
        return HPRSOCP_ERR_ALLOC;
    }
    try {
        Combined C;
        concat_problems(probs, count, C);
        create_impl(&C.p, params, b->h, &C.rng, &C.objc);
    } catch (const Err &e) {
        g_create_error = e.msg;
        delete b->h;
        delete b;
        return e.code;
    }
    *out = b;
    return 0;
}

void hprsocp_batch_destroy(hprsocp_batch *b) {
    if (!b) return;
    delete b->h;
    delete b;
}

const char *hprsocp_batch_last_error(const hprsocp_batch *b) {
    return b ? b->h->err.c_str() : g_create_error.c_str();
}

int32_t hprsocp_batch_size(const hprsocp_batch *b) { return b->h->nb; }

int hprsocp_batch_plan_info(hprsocp_batch *b, char *buf, int64_t cap, int64_t *len) {
    return hprsocp_plan_info(b->h, buf, cap, len);
}

int hprsocp_batch_dims(const hprsocp_batch *b, int32_t i, int32_t *m, int32_t *n) {
    if (i < 0 || i >= b->h->nb) return HPRSOCP_ERR_INVALID;
    *m = b->h->rng[i].m();
    *n = b->h->rng[i].n();
    return 0;
}

int hprsocp_batch_estimate_lambda(hprsocp_batch *b) { return hprsocp_estimate_lambda(b->h); }

int hprsocp_batch_set_initial(hprsocp_batch *b, const double *const *x0s, const double *const *y0s) {
    hprsocp_handle *h = b->h;
    ABI_TRY(h, {
        const int nb = h->nb;
        std::vector<int32_t> fx(nb, 0), fy(nb, 0);
        std::vector<double> X((size_t)h->n, 0.0), Y((size_t)h->m, 0.0);
        bool anyx = false, anyy = false;
        for (int i = 0; i < nb; ++i) {
            const InstRange &R = h->rng[i];
            if (x0s && x0s[i]) {
                const int32_t nbx = R.box1 - R.box0;
                std::copy(x0s[i], x0s[i] + nbx, X.begin() + R.box0);
                std::copy(x0s[i] + nbx, x0s[i] + R.n(), X.begin() + R.var0);
                fx[i] = 1;
                anyx = true;
            }
            if (y0s && y0s[i]) {
                const int32_t nl = R.lin1 - R.lin0;
                std::copy(y0s[i], y0s[i] + nl, Y.begin() + R.lin0);
                std::copy(y0s[i] + nl, y0s[i] + R.m(), Y.begin() + R.con0);
                fy[i] = 1;
                anyy = true;
            }
        }
        set_initial_impl(h, anyx ? X.data() : nullptr, anyy ? Y.data() : nullptr, fx, fy);
    });
    return 0;
}

int hprsocp_batch_solve(hprsocp_batch *b, hprsocp_result *results) {
    hprsocp_handle *h = b->h;
    ABI_TRY(h, { solve_impl(h, results); });
    return 0;
}

int hprsocp_batch_get_solution(hprsocp_batch *b, int32_t i, double *x, double *y, double *z, double *w) {
    hprsocp_handle *h = b->h;
    if (i < 0 || i >= h->nb) {
        h->err = "instance index out of range";
        return HPRSOCP_ERR_INVALID;
    }
    ABI_TRY(h, { get_solution_impl(h, i, x, y, z, w); });
    return 0;
}

int hprsocp_batch_time_iterations(hprsocp_batch *b, int32_t iters, double *ms_per_iter, double *ms_kernel) {
    return hprsocp_time_iterations(b->h, iters, ms_per_iter, ms_kernel);
}

}  // extern "C"
