/*
 * hprsocp.h — C ABI of the MI355X (gfx950) backend for HPR-SOCP's dual
 * Halpern Peaceman–Rachford (HPR) iteration.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an hprsocp_errc (0 = success).  No exception crosses the boundary;
 * the text of the last failure is available from hprsocp_last_error().
 *
 * Indexing is 0-based throughout (the reference is 1-based; the Julia shim
 * subtracts one from rowPtr/colVal/SOC_*_idx before calling in).
 *
 * Canonical layout (reference src/utils/builders.jl:24-113):
 *   rows    = [ equalities | linear inequalities | SOC row blocks ]
 *   columns = [ boxed variables | SOC variable blocks ]
 * soc_con_idx / soc_var_idx hold the start of every cone and end with the
 * sentinel m (resp. n), exactly like SOC_con_idx / SOC_var_idx minus one.
 *
 * Threading: one handle per host thread; a handle is not re-entrant.  Each
 * handle owns one HIP stream and all device memory it allocates.
 */
#ifndef HPRSOCP_H
#define HPRSOCP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPRSOCP_ABI_VERSION 1

/* ---------------------------------------------------------------- errors */
typedef enum {
    HPRSOCP_OK = 0,
    HPRSOCP_ERR_INVALID = 1,     /* bad dimensions / indices / parameters  */
    HPRSOCP_ERR_DEVICE = 2,      /* HIP runtime or kernel launch failure   */
    HPRSOCP_ERR_ALLOC = 3,       /* device or host allocation failed       */
    HPRSOCP_ERR_UNSUPPORTED = 4, /* option this backend does not provide   */
    HPRSOCP_ERR_STATE = 5        /* call out of order (e.g. no lambda yet) */
} hprsocp_errc;

/* Termination status, reference algorithm.jl:1169-1187. */
typedef enum {
    HPRSOCP_STATUS_OPTIMAL = 0,
    HPRSOCP_STATUS_MAX_ITER = 1,
    HPRSOCP_STATUS_TIME_LIMIT = 2,
    HPRSOCP_STATUS_CONTINUE = 3
} hprsocp_status;

/* Kind of quadratic term. */
typedef enum {
    HPRSOCP_Q_CSR = 0,      /* explicit sparse Q (reference SparseMatrixCSC Q)      */
    HPRSOCP_Q_FACTORED = 1  /* operator Q = F^T F + diag(Q_diag) (reference
                               AbstractQOperator seam, Q_operator_interface.jl:33) */
} hprsocp_q_kind;

/* bc_scaling_norm_type, reference scaling.jl:623-633. */
typedef enum { HPRSOCP_BC_L2 = 0, HPRSOCP_BC_RMS = 1, HPRSOCP_BC_LINF = 2 } hprsocp_bc_norm;

/* soc_block_scaling_strategy codes, reference scaling.jl:152-171. */
typedef enum {
    HPRSOCP_SOC_MAX = 1, HPRSOCP_SOC_RMS = 2, HPRSOCP_SOC_GEOM = 3,
    HPRSOCP_SOC_ADAPTIVE = 4, HPRSOCP_SOC_HYBRID = 5, HPRSOCP_SOC_RUIZ_GEOM = 6,
    HPRSOCP_SOC_ROLE_SPLIT = 7, HPRSOCP_SOC_POCK_SAFE_ROLE_SPLIT = 8,
    HPRSOCP_SOC_SIZE_SPLIT_ROLE_SPLIT = 9, HPRSOCP_SOC_PRE_POCK_SIZE_SPLIT = 10,
    HPRSOCP_SOC_PRE_POCK_HUGE_VAR_TAPER = 11,
    HPRSOCP_SOC_PRE_POCK_HUGE_VAR_TAPER_POCK_GEOM = 12,
    HPRSOCP_SOC_PRE_POCK_BAND_SPLIT = 13, HPRSOCP_SOC_PRE_POCK_CONSTRAINT_RMS = 14,
    HPRSOCP_SOC_PRE_POCK_CONSTRAINT_MAX = 15, HPRSOCP_SOC_PRE_POCK_CONSTRAINT_TAPER = 16,
    HPRSOCP_SOC_SMALL_CONE_CONSTRAINT_MAX = 17,
    HPRSOCP_SOC_COUNT_AWARE_SMALL_CONE_BIAS = 18, HPRSOCP_SOC_PHASE_BLEND = 19,
    HPRSOCP_SOC_PHASE_TAPER = 20
} hprsocp_soc_strategy;

/* Arithmetic of the iteration. */
typedef enum {
    HPRSOCP_PREC_FP64 = 0,         /* plain FP64 method (reference numerics)          */
    HPRSOCP_PREC_FP32_REFINE = 1,  /* FP32 bulk: SpMV, projections and Halpern
                                      averaging in FP32 (cone norms t^2 - |xi|^2 in
                                      FP64); every check window ends in FP64:
                                      residuals, gap, M-norm, restart and sigma run
                                      on the FP64 copy of the state; FP64 iterations
                                      take over once the KKT residual reaches
                                      refine_tol or stalls in FP32                    */
    HPRSOCP_PREC_FP32_DELTA = 2    /* FP32 deltas from FP64 anchors: the state is the
                                      FP64 Halpern anchor of the restart cycle plus an
                                      FP32 difference; SpMVs run in FP32 on the
                                      differences, raw values and projections in
                                      FP64, and every restart moves the anchor (an
                                      FP64 refinement).  Check iterations run in
                                      FP64                                            */
} hprsocp_precision;

/* ------------------------------------------------------------ problem */
/* Mirrors QP_info_cpu (reference src/structs.jl:97-147).  All arrays are
 * caller-owned and only read during hprsocp_create(). */
typedef struct {
    int32_t m;                 /* rows of A                                   */
    int32_t n;                 /* columns of A (variables)                    */
    /* A, m x n CSR, 0-based int32 indices, FP64 values.  A^T is built and
       stored by the backend (reference keeps both, builders.jl:104-105).   */
    int64_t A_nnz;
    const int32_t *A_rowptr;   /* m+1 */
    const int32_t *A_colidx;   /* A_nnz */
    const double *A_val;       /* A_nnz */
    /* Q */
    int32_t Q_kind;            /* hprsocp_q_kind                              */
    int64_t Q_nnz;             /* CSR Q: n x n; 0 => no quadratic term        */
    const int32_t *Q_rowptr;   /* n+1 (may be NULL when Q_nnz == 0)          */
    const int32_t *Q_colidx;
    const double *Q_val;
    int32_t F_rows;            /* factored Q: F is F_rows x n CSR             */
    int64_t F_nnz;
    const int32_t *F_rowptr;
    const int32_t *F_colidx;
    const double *F_val;
    const double *Q_diag;      /* factored Q: n entries (NULL => zero)        */
    /* vectors */
    const double *c;           /* n */
    const double *AL;          /* m  (SOC rows: -inf)                        */
    const double *AU;          /* m  (SOC rows: +inf)                        */
    const double *l;           /* n */
    const double *u;           /* n */
    const double *soc_rhs;     /* m - number_linear_con (SOC rows only)       */
    int32_t number_linear_con; /* rows before the first SOC row               */
    int32_t number_lu_x;       /* boxed variables before the first SOC block  */
    int32_t n_soc_con;         /* constraint cones                            */
    const int32_t *soc_con_idx;/* n_soc_con+1, starts, last entry == m        */
    int32_t n_soc_var;         /* variable cones                              */
    const int32_t *soc_var_idx;/* n_soc_var+1, starts, last entry == n        */
    double obj_constant;
} hprsocp_problem;

/* ------------------------------------------------------------ params */
/* Mirrors HPRSOCP_parameters (reference src/structs.jl:208-278) with the same
 * defaults (hprsocp_default_params).  Fields after `use_gpu` are extensions. */
typedef struct {
    double stoptol;            /* 1e-6                                        */
    double sigma;              /* -1 => automatic (algorithm.jl:1538-1551)    */
    int64_t max_iter;          /* typemax(Int32)                              */
    double time_limit;         /* 3600 s                                      */
    int32_t check_iter;        /* 150                                         */
    int32_t warm_up;           /* 0                                           */
    int32_t print_frequency;   /* -1 => adaptive                              */
    int32_t device_number;     /* 0                                           */
    int32_t use_Ruiz_scaling;  /* 1                                           */
    int32_t ruiz_iterations;   /* 10                                          */
    int32_t use_bc_scaling;    /* 1                                           */
    int32_t bc_scaling_norm_type;       /* HPRSOCP_BC_L2                      */
    int32_t use_l2_scaling;    /* 0 (accepted, unused, as in the reference)   */
    int32_t use_Pock_Chambolle_scaling; /* 1                                  */
    int32_t soc_block_scaling_strategy; /* HPRSOCP_SOC_PHASE_TAPER            */
    const double *initial_x;   /* NULL or n                                   */
    const double *initial_y;   /* NULL or m                                   */
    int32_t auto_save;         /* 0; 1 writes the current / best / parameters
                                  checkpoint of algorithm.jl:306-423 at every
                                  print step and at termination (single-
                                  problem handles; an .npz archive whose
                                  member names are the HDF5 dataset paths)    */
    const char *save_filename; /* "HPRSOCP_autosave.npz" (NULL => that name)  */
    int32_t verbose;           /* 1                                           */
    int32_t use_gpu;           /* 1 (0 is rejected: this backend is the GPU)  */
    /* extensions */
    int32_t precision;         /* hprsocp_precision, default FP64             */
    int32_t use_graph;         /* capture the non-check iterations of a
                                  check window in one hipGraph (default 1);
                                  batches capture a window on its second
                                  request after a retirement, eager before */
    double refine_tol;         /* FP32_REFINE: KKT below which the iteration
                                  continues in FP64 (default 1e-6)            */
    int32_t fp32_stall_checks; /* FP32_REFINE: checks without a 1% improvement
                                  of the best KKT (and with two restarts or
                                  more) before FP64 takes over (default 40)   */
    int32_t reserved[7];
} hprsocp_params;

/* ------------------------------------------------------------ results */
/* Mirrors HPRSOCP_results (reference src/structs.jl:297-314). */
typedef struct {
    int64_t iter, iter_4, iter_6;
    double time, time_4, time_6, power_time, setup_time;
    double primal_obj, dual_obj;
    double residuals;          /* KKT = max(err_Rp, err_Rd, rel_gap)          */
    double gap;                /* rel_gap                                     */
    double err_Rp, err_Rp_linear, err_Rp_soc, err_Rd;
    double sigma, lambda_max_A, lambda_max_Q;
    int32_t status;            /* hprsocp_status                              */
    int32_t restarts;
    int64_t iter_fp32;         /* iterations run on the FP32 kernels (FP32 forms) */
} hprsocp_result;

/* Residual block, HPRSOCP_residuals (reference src/structs.jl:281-294). */
typedef struct {
    double err_Rp, err_Rp_linear, err_Rp_soc, err_Rd;
    double kkt, primal_obj, rel_gap, dual_obj;
} hprsocp_residuals;

/* Solver state vectors addressable through get/set (fine-grained seams). */
typedef enum {
    HPRSOCP_V_X = 0, HPRSOCP_V_X_BAR, HPRSOCP_V_X_HAT, HPRSOCP_V_Z_BAR,
    HPRSOCP_V_DX, HPRSOCP_V_LAST_X,                                   /* n */
    HPRSOCP_V_Y, HPRSOCP_V_Y_BAR, HPRSOCP_V_DY, HPRSOCP_V_LAST_Y,
    HPRSOCP_V_S,                                                      /* m */
    HPRSOCP_V_W, HPRSOCP_V_W_BAR, HPRSOCP_V_LAST_W, HPRSOCP_V_DW,
    HPRSOCP_V_ATY, HPRSOCP_V_ATY_BAR, HPRSOCP_V_LAST_ATY,             /* n */
    HPRSOCP_V_C, HPRSOCP_V_L, HPRSOCP_V_U,                            /* n scaled */
    HPRSOCP_V_AL, HPRSOCP_V_AU, HPRSOCP_V_SOC_RHS_FULL,               /* m scaled */
    HPRSOCP_V_ROW_NORM,                                               /* m */
    HPRSOCP_V_COL_NORM, HPRSOCP_V_DIAG_Q,                             /* n */
    HPRSOCP_V_RP, HPRSOCP_V_AX,         /* m, original units after compute_residuals */
    HPRSOCP_V_RD,                       /* n, original units after compute_residuals */
    HPRSOCP_V_COUNT
} hprsocp_vector_id;

/* Restart bookkeeping, HPRSOCP_restart (reference src/structs.jl:600-623).
 * Field order and types follow the Julia mutable struct (Int = int64_t,
 * Bool = uint8_t), so a Julia host passes its own restart_info object by
 * reference.  Owned by the caller of the fine-grained seams below. */
typedef struct {
    int64_t restart_flag;      /* set by check_restart: 0 none, 1 sufficient or
                                  first restart, 2 necessary, 3 long          */
    uint8_t first_restart;     /* true until the check at iter == check_iter  */
    double last_gap, current_gap, save_gap;
    int64_t inner;             /* iterations since the last restart           */
    int64_t step, sufficient, necessary, long_restarts, ratio, times;
    double weighted_norm, best_gap, best_kkt, best_sigma;
    int64_t best_iter;
    uint8_t sigma_correction_active;
    int64_t sigma_correction_hold, sigma_correction_dir;
} hprsocp_restart_info;

/* Scalar block of a handle. */
typedef struct {
    double sigma, lambda_max_A, lambda_max_Q;
    double b_scale, c_scale, norm_b, norm_c, norm_b_org, norm_c_org;
    int32_t Q_is_diag, has_quadratic;
} hprsocp_scalars;

typedef struct hprsocp_handle hprsocp_handle;

/* ------------------------------------------------------------ lifecycle */
/* Fills the reference defaults, HPRSOCP_parameters() (structs.jl:277). */
void hprsocp_default_params(hprsocp_params *params);

/* Copies the problem to the device and performs solve()'s setup up to the
 * workspace: scaling! (scaling.jl:785), check_Q_diagonal (scaling.jl:1312),
 * allocate_workspace (algorithm.jl:1443) incl. sigma_0.
 * Replaces prepare_model/transfer_to_gpu (algorithm.jl:2314-2364). */
int hprsocp_create(const hprsocp_problem *prob, const hprsocp_params *params,
                   hprsocp_handle **out);
void hprsocp_destroy(hprsocp_handle *h);
/* Text of the last error on this handle (or of the last failed create when
 * h is NULL).  Valid until the next call on the same handle. */
const char *hprsocp_last_error(const hprsocp_handle *h);

/* ------------------------------------------------------------ setup seams */
/* estimate_eigenvalues (algorithm.jl:2434-2467): lambda_max(A A^T)*1.05 and
 * lambda_max(Q) (power iteration, eigenvalues.jl:68-196). */
int hprsocp_estimate_lambda(hprsocp_handle *h);
/* process_initial_points! (algorithm.jl:2612-2677) with params.initial_x/y. */
int hprsocp_set_initial(hprsocp_handle *h, const double *x0, const double *y0);

/* ------------------------------------------------------------ hot path */
/* One HPR step, main_update_gpu! (algorithm.jl:2284-2309) with
 * Halpern_fact1 = 1/(inner+2); to_check selects full (1) or partial (0)
 * outputs.  The per-instance scalars are published synchronously, the step
 * is queued on the handle's stream.  With an FP32 precision the FP64 state
 * holds the step's result afterwards (FP32_DELTA: partial steps run on the
 * FP32 differences, full steps in FP64). */
int hprsocp_update(hprsocp_handle *h, int64_t inner, int32_t to_check);
/* compute_residuals! (algorithm.jl:113-196) at iteration `iter`. */
int hprsocp_compute_residuals(hprsocp_handle *h, int64_t iter, hprsocp_residuals *out);
/* compute_M_norm! (algorithm.jl:1302-1359) from the current dx, dy, dw. */
int hprsocp_mnorm(hprsocp_handle *h, double *out);
/* The whole solve loop (algorithm.jl:2765-2905) from the current state;
 * requires hprsocp_estimate_lambda first. */
int hprsocp_solve(hprsocp_handle *h, hprsocp_result *res);

/* Flags of hprsocp_iterate. */
#define HPRSOCP_ITERATE_RESET 1   /* start a new solve loop from the current state */

/* A window of the solve loop (algorithm.jl:2806-2904): runs the loop bodies of
 * iterations iter .. iter+k-1 of the solve that the handle holds, with the
 * residual checks, restart tests, sigma updates and restarts of the reference
 * loop, and returns after them (or at termination).  The runs of plain partial
 * updates inside the window replay captured HIP graphs.  HPRSOCP_ITERATE_RESET
 * (or a first call, or a call after termination) starts a new loop from the
 * current state, exactly as hprsocp_solve does; hprsocp_solve is
 * hprsocp_iterate(h, max_iter + 1, HPRSOCP_ITERATE_RESET, res).  *res receives
 * the loop's result so far; res->status is HPRSOCP_STATUS_CONTINUE until the
 * loop ends.  Calls in windows of any length give the same iterates, bit for
 * bit, as one hprsocp_solve. */
int hprsocp_iterate(hprsocp_handle *h, int64_t k, int32_t flags, hprsocp_result *res);

/* initialize_restart (algorithm.jl:1886-1908). Host only. */
void hprsocp_init_restart(hprsocp_restart_info *r);
/* check_restart (algorithm.jl:1016-1099): restart tests at iteration `iter`
 * from the gaps in *r and the current KKT. Host only. */
void hprsocp_check_restart(hprsocp_restart_info *r, int64_t iter, int64_t check_iter, double sigma,
                           double kkt);
/* update_sigma! (algorithm.jl:729-990): when r->restart_flag >= 1, the model
 * proposal from the device dots of (x_bar - last_x, y_bar - last_y,
 * w_bar - last_w) and their A^T / Q images, blended with r->best_sigma and
 * corrected by the residual imbalance of *res; updates the handle's sigma (and
 * the diagonal-Q factors), r->best_iter and the sigma-correction state. */
int hprsocp_update_sigma(hprsocp_handle *h, hprsocp_restart_info *r, const hprsocp_residuals *res);
/* do_restart! (algorithm.jl:1130-1166): when r->restart_flag > 0, moves the
 * iterates and the Halpern anchors to the averaged point (x, last_x <- x_bar;
 * w, last_w <- w_bar; y, last_y <- y_bar; ATy, last_ATy, ATy_bar <- A^T y_bar
 * for QP) and updates the gaps and counters of *r. */
int hprsocp_restart(hprsocp_handle *h, hprsocp_restart_info *r);
/* Unscaled solution, collect_results! (algorithm.jl:1222-1256).
 * Any pointer may be NULL.  Sizes n, m, n, n. */
int hprsocp_get_solution(hprsocp_handle *h, double *x, double *y, double *z, double *w);

/* One-shot optimize(model, params) (algorithm.jl:2538-2581). */
int hprsocp_optimize(const hprsocp_problem *prob, const hprsocp_params *params,
                     hprsocp_result *res, double *x, double *y, double *z, double *w);

/* ------------------------------------------------------------ inspection */
int hprsocp_get_vector(hprsocp_handle *h, int32_t id, double *out);
int hprsocp_set_vector(hprsocp_handle *h, int32_t id, const double *in);
int hprsocp_get_scalars(hprsocp_handle *h, hprsocp_scalars *out);
/* Overrides sigma / lambda_max_A / lambda_max_Q (negative = keep) and
 * refreshes the diagonal-Q factors (unified_update_Q_factors!,
 * unified_operations.jl:468-482). */
int hprsocp_set_scalars(hprsocp_handle *h, double sigma, double lambda_A, double lambda_Q);
/* Scaled A values in CSR order (A_nnz entries). */
int hprsocp_get_scaled_A(hprsocp_handle *h, double *val);
/* Launch-plan choices of the autotuner for every operator of the handle, as a
 * JSON object {"qw_primal_fused": 0|1 (the primal step runs in the F^T
 * product's tile kernel), "operators": [{name, form (plain | slicesP | sortedP |
 * sortedP-fused | sortedP-rounds), tuned, cached, nnz_target, g_divisor, huge_target,
 * huge_whole, stream, tiles, candidates, ms_default, ms, ms_form, block,
 * form32, block32, ms32}, ...]}: the tile-plan configuration
 * chosen by the timed search (the reference's fixed launch shapes,
 * kernels.jl:349-357, and cone size classes, algorithm.jl:1914-1919), the SpMV
 * time of the default and the chosen plan, the SpMV form kept by the XCD
 * slicing tuner with its block size, and the layout of the FP32 copy (form32:
 * own | fp64, with -fused for a slice-fused layout and -rounds for one cut
 * into whole rounds of workgroup slots); times of forms forced
 * through the environment read 0.  Writes at most cap bytes (NUL-terminated)
 * to buf and the full length to *len (either may be NULL). */
int hprsocp_plan_info(hprsocp_handle *h, char *buf, int64_t cap, int64_t *len);
/* Dimensions: m, n, nnz(A). */
int hprsocp_get_dims(hprsocp_handle *h, int32_t *m, int32_t *n, int64_t *nnz);

/* ------------------------------------------------------------ kernels */
/* Stand-alone SpMV through the backend's tile plan: y = A x (trans=0) or
 * x = A^T y (trans=1) on the handle's scaled A; host buffers in and out.
 * (Replaces unified_mul!, unified_operations.jl:140-166.) */
int hprsocp_spmv(hprsocp_handle *h, int32_t trans, const double *in, double *out);
/* Time the hot loop: runs `iters` HPR iterations (partial outputs, no
 * residual checks, Halpern counter 0..iters-1) from the current state and
 * returns the device time per iteration measured with HIP events around the
 * batch.  When ms_kernel is not NULL a second, event-bracketed batch fills
 * ms_kernel[0..5] with the mean time per iteration of each launch group:
 *   noQ: [0] A^T-SpMV + box/SOC-variable projection + Halpern,
 *        [1] A-SpMV + linear/SOC-row projection + Halpern;
 *   QP : [0] Q w, [1] primal projection (zxw1), [2] Q w_bar -> tempv,
 *        [3] A-SpMV dual update, [4] A^T-SpMV w2 update.
 * (No counterpart in the reference; it is the measurement hook.) */
int hprsocp_time_iterations(hprsocp_handle *h, int32_t iters, double *ms_per_iter,
                            double *ms_kernel);

/* Copies `bytes` between two fresh device buffers with the backend's plain copy
 * kernel (one launch on a cold source) and returns its HIP-event time.  Used to
 * calibrate hardware byte counters (FETCH_SIZE / WRITE_SIZE) against a kernel
 * of known traffic.  (No counterpart in the reference; measurement hook.) */
int hprsocp_probe_copy(hprsocp_handle *h, int64_t bytes, double *ms);

/* Median HIP-event time (ms) of `reps` launches of the FP64 SpMV A x
 * (trans = 0) or A^T y (trans = 1) exactly as the FP64 iteration runs it
 * (trans | 2: the FP32 operator of the FP32 precisions, as their iterations
 * run it)
 * (sliced or plain form chosen by the autotuner), on device-resident
 * operands.  Single-problem handles.  (No counterpart in the reference;
 * measurement hook.) */
int hprsocp_time_spmv(hprsocp_handle *h, int32_t trans, int32_t reps, double *ms);

/* Gather ceiling of the sliced SpMV: times a synthetic kernel that streams
 * nnz entries of 12 bytes (8-byte value, 4-byte packed column/position word)
 * and gathers one double per entry from `slices` column slices of `width`
 * doubles (block b reads slice b % slices, as the sliced SpMV does), in random
 * column order (sorted = 0, the row-major layout's access pattern) or in
 * column-sorted chunks (sorted = 1).  The data is generated on the device.
 * Returns the median of 5 launches in ms.  (No counterpart in the reference;
 * measurement hook.) */
int hprsocp_probe_gather(hprsocp_handle *h, int64_t nnz, int32_t width, int32_t slices, int32_t sorted,
                         double *ms);
/* The same gather ceiling for the FP32 operators: 8-byte entries (4-byte
 * value, 4-byte packed word), one float gathered per entry, and the FP32
 * sorted layout's chunks of 16384 entries (sorted = 1).  (Measurement hook.) */
int hprsocp_probe_gather32(hprsocp_handle *h, int64_t nnz, int32_t width, int32_t slices, int32_t sorted,
                           double *ms);
/* The control arithmetic of batch solves (residual block, M-norm, restart
 * test, sigma update) runs on the device or on the host, from the same source
 * with no fused multiply-add and its own exp / log; this hook evaluates the
 * same functions of in[0..n) in both places: for each x, out[5k..5k+4] =
 * exp(x), log|x|, sqrt|x|, a general-Q golden-section search and the sigma
 * update of a fixed restart state.  dev_out and host_out are host arrays of 5n
 * doubles; they agree bit for bit.  (Measurement hook.) */
int hprsocp_probe_ctl(hprsocp_handle *h, const double *in, int32_t n, double *dev_out, double *host_out);

/* ------------------------------------------------------------ cone projection */
/* Stand-alone projection onto a product of second-order cones, on device
 * memory: for each k < ncones, rows [cone_idx[k], cone_idx[k+1]) of `in` are
 * read as (t, xi) and written to the same rows of `out` as Pi_Q(t, xi): zero
 * when |xi| <= -t, unchanged when |xi| <= t, else ((t+|xi|)/2)(1, xi/|xi|)
 * (the reference's projection, kernels.jl:7585-7663, as used by the variable
 * and constraint cone updates, :7748-7811 and :7895-7951).  Rows outside every
 * cone are not touched.  in, out (n doubles, in == out allowed) and cone_idx
 * (ncones + 1 int32, 0-based, nondecreasing, first >= 0, last <= n) are device
 * pointers; the work is queued on `stream` (a hipStream_t; NULL = the default
 * stream).  Long cones are split over several workgroups; the per-tile partial
 * norms live in a scratch buffer that this call allocates on the stream and
 * frees again.  The call waits for the stream, and returns
 * HPRSOCP_ERR_INVALID (out untouched) for cone offsets that break the rules
 * above.  Thread-safe; the text of the last failure on the calling thread is
 * available from hprsocp_project_cones_last_error(). */
int hprsocp_project_cones(const double *in, double *out, int64_t n, const int32_t *cone_idx, int32_t ncones,
                          void *stream);
const char *hprsocp_project_cones_last_error(void);
/* hprsocp_project_cones without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments and results; the scratch is a fixed device array of the library
 * (about 1.75 MiB per GPU, held until the process ends), so a call is limited
 * to ncones <= 65536 and ceil(n / 2048) + ncones <= 131072 (larger calls
 * return HPRSOCP_ERR_INVALID before queuing anything).  The call only queues
 * the work on `stream` and returns without waiting; cone offsets that break
 * the rules of hprsocp_project_cones leave `out` untouched, with no error
 * code (none can be read back without waiting).  Every call on one GPU uses
 * the same scratch: calls that overlap on one GPU (different streams, or
 * threads) race and give wrong results; calls queued on one stream are
 * ordered and safe. */
int hprsocp_project_cones_noalloc(const double *in, double *out, int64_t n, const int32_t *cone_idx,
                                  int32_t ncones, void *stream);

/* ------------------------------------------------------------ CBF input */
/* Reads a Conic Benchmark Format file into the canonical layout, as read_cbf
 * + build_from_cbf do (reference src/utils/cbf_io.jl:493-750, 899-1131,
 * src/utils/builders.jl:24-130): rows L= | L+,L- (L- negated) | Q, F rows
 * dropped; variables linear | Q; objective sense max negates c and the
 * constant.  Host code only (no device is touched). */
typedef struct hprsocp_cbf hprsocp_cbf;
int hprsocp_cbf_read(const char *path, hprsocp_cbf **out);
const char *hprsocp_cbf_last_error(void);
/* Fills `prob` with pointers into the reader's storage (valid until destroy);
 * number_eq (may be NULL) receives the count of equality rows. */
int hprsocp_cbf_problem(const hprsocp_cbf *m, hprsocp_problem *prob, int32_t *number_eq);
void hprsocp_cbf_destroy(hprsocp_cbf *m);

/* ------------------------------------------------------------ batch */
/* Many independent problems solved together: the instances are packed into one
 * block-diagonal CSR / cone-offset layout and iterated in lockstep by the same
 * fused kernels (one launch per step for all instances).  Every instance keeps
 * its own scaling, lambda, sigma, restart state, Halpern counter and
 * termination; a finished instance is retired by its per-instance mask and its
 * tiles skip all work.  Instances must all have a CSR Q or all none (the
 * operator form is single-problem only), and all a diagonal Q or all a general
 * one.  (No counterpart in the reference, which solves one model per call; each
 * instance follows optimize(), algorithm.jl:2538-2581.) */
typedef struct hprsocp_batch hprsocp_batch;
int hprsocp_batch_create(const hprsocp_problem *const *probs, int32_t count,
                         const hprsocp_params *params, hprsocp_batch **out);
void hprsocp_batch_destroy(hprsocp_batch *b);
const char *hprsocp_batch_last_error(const hprsocp_batch *b);
int32_t hprsocp_batch_size(const hprsocp_batch *b);
/* m and n of instance i */
int hprsocp_batch_dims(const hprsocp_batch *b, int32_t i, int32_t *m, int32_t *n);
/* hprsocp_plan_info of the batch's packed operators */
int hprsocp_batch_plan_info(hprsocp_batch *b, char *buf, int64_t cap, int64_t *len);
/* per-instance estimate_eigenvalues (algorithm.jl:2434-2467) */
int hprsocp_batch_estimate_lambda(hprsocp_batch *b);
/* per-instance process_initial_points! (algorithm.jl:2612-2677); x0s / y0s hold
 * `count` pointers (sizes n_i / m_i); a NULL array or entry means a cold start. */
int hprsocp_batch_set_initial(hprsocp_batch *b, const double *const *x0s, const double *const *y0s);
/* solves every instance; results has `count` entries */
int hprsocp_batch_solve(hprsocp_batch *b, hprsocp_result *results);
/* collect_results! (algorithm.jl:1222-1256) of instance i; sizes n_i, m_i, n_i, n_i */
int hprsocp_batch_get_solution(hprsocp_batch *b, int32_t i, double *x, double *y, double *z, double *w);
/* like hprsocp_time_iterations, for the whole batch (measurement hook) */
int hprsocp_batch_time_iterations(hprsocp_batch *b, int32_t iters, double *ms_per_iter, double *ms_kernel);

#ifdef __cplusplus
}
#endif
#endif /* HPRSOCP_H */
