/*
 * hprsocp_oracle.c — FP64 CPU restatement of the HPR-SOCP solver (reference
 * PolyU-IOR/HPR-SOCP, Julia).  TEST INFRASTRUCTURE ONLY: the checker that the
 * HIP backend is compared against, and the CPU baseline timed by bench.py.
 * Nothing in the product path links, loads or calls this file.
 *
 * Each function cites the reference file:line it restates.  Indexing is
 * 0-based; the reference is 1-based.  Julia's NaN-propagating max/min/clamp
 * are reproduced (jmax/jmin/jclamp) because the sigma heuristics rely on them.
 *
 * Pinning (what checks each part, see tests/test_oracle.py):
 *   - SOC projection, residual split, bounds violation, dual-objective rhs
 *     shift, M-norm sign: reference known-answer tests (test/runtests.jl
 *     297-723) held in tests/golden/reference_kats.json;
 *   - SpMV / transpose: scipy.sparse; projections: Moreau decomposition and
 *     idempotence; power iteration: numpy eigvalsh; golden sections: dense
 *     scan; scaling factors: dense numpy restatement on small matrices;
 *   - whole solves: closed-form optima (data/model.cbf, Clarabel test,
 *     trust-region secular equation) and an independent KKT certificate.
 *   - parity unpinned: the exact iteration trajectory of the reference
 *     (Julia MersenneTwister start vectors and cuSPARSE summation order
 *     cannot be reproduced without Julia).
 */
#include "oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ misc */
static int g_threads = 1;
void oracle_set_threads(int32_t nthreads) { g_threads = nthreads < 1 ? 1 : nthreads; }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Julia semantics: max/min propagate NaN; clamp(x,lo,hi) =
 * ifelse(x > hi, hi, ifelse(x < lo, lo, x)). */
static double jmax(double a, double b) {
    if (a != a || b != b) return NAN;
    if (a > b) return a;
    if (b > a) return b;
    return signbit(a) ? b : a;
}
static double jmin(double a, double b) {
    if (a != a || b != b) return NAN;
    if (a < b) return a;
    if (b < a) return b;
    return signbit(a) ? a : b;
}
static double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

int64_t oracle_abi_sizeof(int32_t which) {
    switch (which) {
    case 0: return (int64_t)sizeof(hprsocp_problem);
    case 1: return (int64_t)sizeof(hprsocp_params);
    case 2: return (int64_t)sizeof(hprsocp_result);
    case 3: return (int64_t)sizeof(hprsocp_residuals);
    case 4: return (int64_t)sizeof(hprsocp_scalars);
    case 5: return (int64_t)sizeof(hprsocp_restart_info);
    default: return -1;
    }
}

/* ------------------------------------------------------------------ CSR */
typedef struct {
    int32_t nrows, ncols;
    int64_t nnz;
    int32_t *rowptr;
    int32_t *col;
    double *val;
} csr_t;

static void csr_free(csr_t *a) {
    free(a->rowptr);
    free(a->col);
    free(a->val);
    memset(a, 0, sizeof(*a));
}

static int csr_alloc(csr_t *a, int32_t nrows, int32_t ncols, int64_t nnz) {
    a->nrows = nrows;
    a->ncols = ncols;
    a->nnz = nnz;
    a->rowptr = (int32_t *)calloc((size_t)nrows + 1, sizeof(int32_t));
    a->col = (int32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
    a->val = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    return (a->rowptr && a->col && a->val) ? 0 : -1;
}

static int csr_copy(csr_t *dst, int32_t nrows, int32_t ncols, int64_t nnz, const int32_t *rp,
                    const int32_t *ci, const double *v) {
    if (csr_alloc(dst, nrows, ncols, nnz)) return -1;
    if (nnz > 0 || rp) memcpy(dst->rowptr, rp, ((size_t)nrows + 1) * sizeof(int32_t));
    if (nnz > 0) {
        memcpy(dst->col, ci, (size_t)nnz * sizeof(int32_t));
        memcpy(dst->val, v, (size_t)nnz * sizeof(double));
    }
    return 0;
}

/* Counting-sort transpose; columns inside each output row come out sorted by
 * source row, i.e. A^T in canonical CSR (builders.jl:105 keeps A' stored). */
static int csr_transpose(const csr_t *a, csr_t *t) {
    if (csr_alloc(t, a->ncols, a->nrows, a->nnz)) return -1;
    for (int64_t k = 0; k < a->nnz; ++k) t->rowptr[a->col[k] + 1]++;
    for (int32_t j = 0; j < a->ncols; ++j) t->rowptr[j + 1] += t->rowptr[j];
    int32_t *next = (int32_t *)malloc(((size_t)a->ncols + 1) * sizeof(int32_t));
    if (!next) return -1;
    memcpy(next, t->rowptr, ((size_t)a->ncols + 1) * sizeof(int32_t));
    for (int32_t i = 0; i < a->nrows; ++i)
        for (int32_t k = a->rowptr[i]; k < a->rowptr[i + 1]; ++k) {
            int32_t dst = next[a->col[k]]++;
            t->col[dst] = i;
            t->val[dst] = a->val[k];
        }
    free(next);
    return 0;
}

/* y = A x, sequential summation in CSR order per row (SparseArrays mul!). */
static void csr_mv(const csr_t *a, const double *x, double *y) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int32_t i = 0; i < a->nrows; ++i) {
        double s = 0.0;
        for (int32_t k = a->rowptr[i]; k < a->rowptr[i + 1]; ++k) s += a->val[k] * x[a->col[k]];
        y[i] = s;
    }
}

static double dot(const double *a, const double *b, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double nrm2(const double *a, int64_t n) { return sqrt(dot(a, a, n)); }
/* unified_absmax (unified_operations.jl:76-90): maximum(abs, x), 0 if empty. */
static double absmax(const double *a, int64_t n) {
    double m = 0.0;
    for (int64_t i = 0; i < n; ++i) m = jmax(m, fabs(a[i]));
    return m;
}

/* -------------------------------------------------------- workspace type */
struct oracle_ws {
    char err[512];
    hprsocp_params params;
    int32_t m, n, m1, n1, ncon, nvar;
    int32_t *con_idx; /* ncon+1 */
    int32_t *var_idx; /* nvar+1 */
    csr_t A, AT, Q;   /* scaled */
    int32_t q_kind;
    csr_t F, FT;      /* factored Q */
    double *Fdiag;    /* factored Q diag */
    double *Ftmp;     /* F_rows */
    int has_quad;
    int Q_is_diag;
    double *diag_Q;
    double obj_const;
    /* scaled data */
    double *c, *AL, *AU, *l, *u, *soc_rhs, *soc_rhs_full;
    /* scaling info (Scaling_info_cpu, structs.jl:173-186) */
    double *row_norm, *col_norm;
    double b_scale, c_scale, norm_b, norm_c, norm_b_org, norm_c_org;
    /* iterates (HPRSOCP_workspace_cpu, structs.jl:530-597) */
    double *w, *w_bar, *dw, *x, *x_hat, *x_bar, *dx, *y, *y_bar, *dy, *z_bar;
    double *Rp, *Rd, *Ax, *ATy, *ATy_bar, *ATdy, *QATdy, *s, *Qw, *Qw_bar, *Qx, *dQw;
    double *last_x, *last_y, *last_w, *last_ATy, *tempv;
    double *fact1, *fact2, *fact, *fact_M;
    double sigma, lambda_A, lambda_Q;
    int lambda_set;
    int to_check;
    double power_time, setup_time;
};

static int set_err(oracle_ws *w, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(w->err, sizeof(w->err), fmt, ap);
    va_end(ap);
    return HPRSOCP_ERR_INVALID;
}

const char *oracle_last_error(const oracle_ws *w) { return w ? w->err : "null handle"; }

/* HPRSOCP_parameters() defaults, structs.jl:277. */
void oracle_default_params(hprsocp_params *p) {
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
    p->save_filename = "HPRSOCP_autosave.npz"; /* auto_save is ignored by the oracle */
    p->verbose = 1;
    p->use_gpu = 1;
    p->precision = HPRSOCP_PREC_FP64;
    p->use_graph = 1;
    p->refine_tol = 1e-6;
    p->fp32_stall_checks = 40;
}

/* Q apply, Qmap! (sparse_matrix_operator.jl:47-49) or the factored operator
 * F^T(F v) + diag .* v. */
static void qmap(oracle_ws *w, const double *v, double *out) {
    if (w->q_kind == HPRSOCP_Q_FACTORED) {
        csr_mv(&w->F, v, w->Ftmp);
        csr_mv(&w->FT, w->Ftmp, out);
        if (w->Fdiag)
            for (int32_t i = 0; i < w->n; ++i) out[i] += w->Fdiag[i] * v[i];
    } else if (w->Q.nnz > 0) {
        csr_mv(&w->Q, v, out);
    } else {
        memset(out, 0, sizeof(double) * (size_t)w->n);
    }
}

/* ------------------------------------------------------- SOC projection */
void oracle_soc_project(double *v, int32_t len) {
    if (len <= 0) return;
    double t = v[0], ns = 0.0;
    for (int32_t j = 1; j < len; ++j) ns += v[j] * v[j];
    ns = sqrt(ns);
    if (ns <= -t) {
        for (int32_t j = 0; j < len; ++j) v[j] = 0.0;
    } else if (ns <= t) {
        /* inside: keep */
    } else {
        double pt = (ns + t) / 2.0, alpha = pt / ns;
        v[0] = pt;
        for (int32_t j = 1; j < len; ++j) v[j] = alpha * v[j];
    }
}

/* ------------------------------------------------------------- scaling */
enum { PH_GENERIC = 0, PH_RUIZ = 1, PH_PRE_POCK = 2, PH_POCK = 3 };
enum { LOC_GENERIC = 0, LOC_CON = 1, LOC_VAR = 2 };
#define ADAPTIVE_T 8
#define POCK_RMS_T 512
#define MANY_CONES_T 100000

static double max_biased(double bm, double rms) { return sqrt(sqrt(bm * bm * bm * rms)); }
static double huge_taper(double bm, double rms) { return sqrt(sqrt(bm * rms * rms * rms)); }

/* _soc_block_scale_value_scalar, scaling.jl:253-477. */
double oracle_soc_block_scale_value(int32_t L, double bm, double rms, int32_t s, int32_t ph,
                                    int32_t loc, int32_t cone_count) {
    if (L <= 0) return 1.0;
    if (s == HPRSOCP_SOC_MAX) return bm;
    double geom = sqrt(bm * rms);
    switch (s) {
    case HPRSOCP_SOC_RMS: return rms;
    case HPRSOCP_SOC_GEOM: return geom;
    case HPRSOCP_SOC_ADAPTIVE: return L <= ADAPTIVE_T ? bm : rms;
    case HPRSOCP_SOC_HYBRID:
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_POCK) return rms;
        return geom;
    case HPRSOCP_SOC_RUIZ_GEOM:
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : geom;
        if (ph == PH_POCK) return rms;
        return geom;
    case HPRSOCP_SOC_ROLE_SPLIT:
        if (loc == LOC_CON) return geom;
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_POCK) return rms;
        return geom;
    case HPRSOCP_SOC_POCK_SAFE_ROLE_SPLIT:
        if (ph == PH_POCK) return bm;
        if (loc == LOC_CON) return geom;
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        return geom;
    case HPRSOCP_SOC_SIZE_SPLIT_ROLE_SPLIT:
        if (loc == LOC_CON) return ph == PH_POCK ? bm : geom;
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_POCK) return L <= POCK_RMS_T ? rms : bm;
        return geom;
    case HPRSOCP_SOC_PRE_POCK_SIZE_SPLIT:
        if (loc == LOC_CON) return ph == PH_POCK ? bm : geom;
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_PRE_POCK) return L <= POCK_RMS_T ? rms : bm;
        if (ph == PH_POCK) return bm;
        return geom;
    case HPRSOCP_SOC_PRE_POCK_HUGE_VAR_TAPER:
        if (loc == LOC_CON) return ph == PH_POCK ? bm : geom;
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_PRE_POCK) return L <= POCK_RMS_T ? rms : huge_taper(bm, rms);
        if (ph == PH_POCK) return bm;
        return geom;
    case HPRSOCP_SOC_PRE_POCK_HUGE_VAR_TAPER_POCK_GEOM:
        if (loc == LOC_CON) return ph == PH_POCK ? bm : geom;
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_PRE_POCK) return L <= POCK_RMS_T ? rms : huge_taper(bm, rms);
        if (ph == PH_POCK) return L <= POCK_RMS_T ? bm : geom;
        return geom;
    case HPRSOCP_SOC_PRE_POCK_BAND_SPLIT:
        if (loc == LOC_CON) return ph == PH_POCK ? bm : geom;
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_PRE_POCK) return L <= ADAPTIVE_T ? bm : (L <= POCK_RMS_T ? rms : bm);
        if (ph == PH_POCK) return bm;
        return geom;
    case HPRSOCP_SOC_PRE_POCK_CONSTRAINT_RMS:
        if (loc == LOC_CON) {
            if (ph == PH_POCK) return bm;
            if (ph == PH_PRE_POCK) return rms;
            return geom;
        }
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_PRE_POCK) return L <= POCK_RMS_T ? rms : bm;
        if (ph == PH_POCK) return bm;
        return geom;
    case HPRSOCP_SOC_PRE_POCK_CONSTRAINT_MAX:
        if (loc == LOC_CON) {
            if (ph == PH_POCK || ph == PH_PRE_POCK) return bm;
            return geom;
        }
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_PRE_POCK) return L <= POCK_RMS_T ? rms : bm;
        if (ph == PH_POCK) return bm;
        return geom;
    case HPRSOCP_SOC_PRE_POCK_CONSTRAINT_TAPER:
        if (loc == LOC_CON) {
            if (ph == PH_POCK) return bm;
            if (ph == PH_PRE_POCK) return huge_taper(bm, rms);
            return geom;
        }
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_PRE_POCK) return L <= POCK_RMS_T ? rms : bm;
        if (ph == PH_POCK) return bm;
        return geom;
    case HPRSOCP_SOC_SMALL_CONE_CONSTRAINT_MAX:
        if (loc == LOC_CON) {
            if (ph == PH_POCK) return bm;
            if (ph == PH_RUIZ || ph == PH_PRE_POCK) return L <= ADAPTIVE_T ? bm : geom;
            return geom;
        }
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_PRE_POCK) return L <= POCK_RMS_T ? rms : bm;
        if (ph == PH_POCK) return bm;
        return geom;
    case HPRSOCP_SOC_COUNT_AWARE_SMALL_CONE_BIAS:
        if (loc == LOC_CON) {
            int many_small = cone_count >= MANY_CONES_T && L <= ADAPTIVE_T;
            if (ph == PH_POCK) return bm;
            if (many_small) {
                if (ph == PH_RUIZ) return max_biased(bm, rms);
                if (ph == PH_PRE_POCK) return bm;
            }
            return geom;
        }
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_PRE_POCK) return L <= POCK_RMS_T ? rms : bm;
        if (ph == PH_POCK) return bm;
        return geom;
    case HPRSOCP_SOC_PHASE_BLEND:
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_POCK) return L <= ADAPTIVE_T ? rms : geom;
        return geom;
    case HPRSOCP_SOC_PHASE_TAPER:
        if (ph == PH_RUIZ) return L <= ADAPTIVE_T ? bm : rms;
        if (ph == PH_POCK) return L <= ADAPTIVE_T ? rms : geom;
        return geom;
    default: return bm;
    }
}

/* _apply_soc_block_scaling! CPU, scaling.jl:500-526 (soc_block_scale_value
 * scaling.jl:479-498: block max, rms = sqrt(sum(abs2)/len)). */
static void apply_soc_block_scaling(double *t, const int32_t *idx, int32_t ncones, int32_t strategy,
                                    int32_t phase, int32_t loc) {
    for (int32_t i = 0; i < ncones; ++i) {
        int32_t s = idx[i], e = idx[i + 1];
        if (s >= e) continue;
        double bm = t[s], ss = 0.0;
        for (int32_t j = s; j < e; ++j) {
            bm = jmax(bm, t[j]);
            ss += t[j] * t[j];
        }
        double rms = sqrt(ss / (double)(e - s));
        double v = oracle_soc_block_scale_value(e - s, bm, rms, strategy, phase, loc, ncones);
        for (int32_t j = s; j < e; ++j) t[j] = v;
    }
}

static void clip_norms(double *v, int32_t n) {
    for (int32_t i = 0; i < n; ++i) v[i] = jclamp(v[i], 1e-4, 1e4);
}

/* One equilibration pass of scaling! (CPU branches): Ruiz (max, scaling.jl
 * 908-1003), pre-Pock (max, :1007-1094) or Pock-Chambolle (sum, :1098-1189). */
static void equilibrate(oracle_ws *w, int use_sum, int phase, double *trow, double *tcol) {
    const int32_t m = w->m, n = w->n;
    /* column statistics of |A| and |Q| (maximum(abs, A, dims=1)) */
    for (int32_t j = 0; j < n; ++j) tcol[j] = 0.0;
    double *tq = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int32_t i = 0; i < m; ++i)
        for (int32_t k = w->A.rowptr[i]; k < w->A.rowptr[i + 1]; ++k) {
            int32_t j = w->A.col[k];
            double a = fabs(w->A.val[k]);
            tcol[j] = use_sum ? tcol[j] + a : jmax(tcol[j], a);
        }
    for (int32_t i = 0; i < n && w->Q.nnz > 0; ++i)
        for (int32_t k = w->Q.rowptr[i]; k < w->Q.rowptr[i + 1]; ++k) {
            int32_t j = w->Q.col[k];
            double a = fabs(w->Q.val[k]);
            tq[j] = use_sum ? tq[j] + a : jmax(tq[j], a);
        }
    for (int32_t j = 0; j < n; ++j) {
        tcol[j] = sqrt(use_sum ? tcol[j] + tq[j] : jmax(tcol[j], tq[j]));
        if (tcol[j] == 0.0) tcol[j] = 1.0;
    }
    free(tq);
    if (m > 0) {
        for (int32_t i = 0; i < m; ++i) {
            double r = 0.0;
            for (int32_t k = w->A.rowptr[i]; k < w->A.rowptr[i + 1]; ++k)
                r = use_sum ? r + fabs(w->A.val[k]) : jmax(r, fabs(w->A.val[k]));
            trow[i] = sqrt(r);
            if (trow[i] == 0.0) trow[i] = 1.0;
        }
        apply_soc_block_scaling(trow, w->con_idx, w->ncon, w->params.soc_block_scaling_strategy,
                                phase, LOC_CON);
        clip_norms(trow, m);
    }
    apply_soc_block_scaling(tcol, w->var_idx, w->nvar, w->params.soc_block_scaling_strategy, phase,
                            LOC_VAR);
    clip_norms(tcol, n);
    for (int32_t i = 0; i < m; ++i) w->row_norm[i] *= trow[i];
    for (int32_t j = 0; j < n; ++j) w->col_norm[j] *= tcol[j];
    /* Q = DC*Q*DC, A = DR*A*DC  ((d_i * v) * d_j, d = 1 ./ t) */
    for (int32_t i = 0; i < n && w->Q.nnz > 0; ++i)
        for (int32_t k = w->Q.rowptr[i]; k < w->Q.rowptr[i + 1]; ++k)
            w->Q.val[k] = ((1.0 / tcol[i]) * w->Q.val[k]) * (1.0 / tcol[w->Q.col[k]]);
    for (int32_t i = 0; i < m; ++i)
        for (int32_t k = w->A.rowptr[i]; k < w->A.rowptr[i + 1]; ++k)
            w->A.val[k] = ((1.0 / trow[i]) * w->A.val[k]) * (1.0 / tcol[w->A.col[k]]);
    for (int32_t j = 0; j < n; ++j) w->c[j] /= tcol[j];
    if (m > 0) {
        for (int32_t i = 0; i < m; ++i) {
            w->AL[i] /= trow[i];
            w->AU[i] /= trow[i];
            w->soc_rhs_full[i] /= trow[i];
        }
        for (int32_t i = w->m1; i < m; ++i) w->soc_rhs[i - w->m1] /= trow[i];
    }
    for (int32_t j = 0; j < n; ++j) {
        w->l[j] *= tcol[j];
        w->u[j] *= tcol[j];
    }
}

/* _finite_bound_norm / _vector_bc_norm, scaling.jl:635-697. */
static double finite_bound_norm(const double *v, int64_t n, int type) {
    double acc = 0.0;
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (!isfinite(v[i])) continue;
        if (type == HPRSOCP_BC_LINF) acc = jmax(acc, fabs(v[i]));
        else {
            acc += v[i] * v[i];
            cnt++;
        }
    }
    if (type == HPRSOCP_BC_LINF) return acc;
    if (cnt == 0) return 0.0;
    return type == HPRSOCP_BC_RMS ? sqrt(acc / (double)cnt) : sqrt(acc);
}
static double vector_bc_norm(const double *v, int64_t n, int type) {
    if (n == 0) return 0.0;
    if (type == HPRSOCP_BC_LINF) return absmax(v, n);
    double s = dot(v, v, n);
    return type == HPRSOCP_BC_RMS ? sqrt(s / (double)n) : sqrt(s);
}

/* norm of max(|AL_nInf|, |AU_nInf|) (scaling.jl:879-884, 1272-1277) */
static double bound_norm(const oracle_ws *w, int inf) {
    double acc = 0.0;
    for (int32_t i = 0; i < w->m; ++i) {
        double a = w->AL[i] == -INFINITY ? 0.0 : w->AL[i];
        double b = w->AU[i] == INFINITY ? 0.0 : w->AU[i];
        double v = jmax(fabs(a), fabs(b));
        acc = inf ? jmax(acc, v) : acc + v * v;
    }
    return inf ? acc : sqrt(acc);
}
/* _soc_rhs_norm, scaling.jl:607-612 (infinity norm). */
static double soc_rhs_inf(const oracle_ws *w) {
    return (w->ncon == 0 || w->m - w->m1 == 0) ? 0.0 : absmax(w->soc_rhs, w->m - w->m1);
}

/* Q = (Q + Q')/2 on CPU (scaling.jl:1265-1269): merge Q with its transpose. */
static int symmetrize_Q(oracle_ws *w) {
    if (w->Q.nnz == 0) return 0;
    csr_t T;
    if (csr_transpose(&w->Q, &T)) return -1;
    csr_t S;
    if (csr_alloc(&S, w->n, w->n, w->Q.nnz + T.nnz)) return -1;
    int64_t p = 0;
    for (int32_t i = 0; i < w->n; ++i) {
        S.rowptr[i] = (int32_t)p;
        int32_t a = w->Q.rowptr[i], ae = w->Q.rowptr[i + 1], b = T.rowptr[i], be = T.rowptr[i + 1];
        /* rows of Q may be unsorted: sort a copy by column (insertion) */
        int32_t len = ae - a;
        int32_t *ci = (int32_t *)malloc(sizeof(int32_t) * (size_t)(len > 0 ? len : 1));
        double *cv = (double *)malloc(sizeof(double) * (size_t)(len > 0 ? len : 1));
        for (int32_t k = 0; k < len; ++k) {
            int32_t cc = w->Q.col[a + k];
            double vv = w->Q.val[a + k];
            int32_t q = k;
            while (q > 0 && ci[q - 1] > cc) {
                ci[q] = ci[q - 1];
                cv[q] = cv[q - 1];
                --q;
            }
            ci[q] = cc;
            cv[q] = vv;
        }
        int32_t ia = 0;
        while (ia < len || b < be) {
            int32_t ca = ia < len ? ci[ia] : 2147483647, cb = b < be ? T.col[b] : 2147483647;
            double v;
            int32_t cc;
            if (ca == cb) {
                v = (cv[ia] + T.val[b]) / 2.0;
                cc = ca;
                ia++;
                b++;
            } else if (ca < cb) {
                v = (cv[ia] + 0.0) / 2.0;
                cc = ca;
                ia++;
            } else {
                v = (0.0 + T.val[b]) / 2.0;
                cc = cb;
                b++;
            }
            S.col[p] = cc;
            S.val[p] = v;
            p++;
        }
        free(ci);
        free(cv);
    }
    S.rowptr[w->n] = (int32_t)p;
    S.nnz = p;
    csr_free(&T);
    csr_free(&w->Q);
    w->Q = S;
    return 0;
}

/* scaling! CPU path, scaling.jl:785-1289. */
static int do_scaling(oracle_ws *w) {
    const int32_t m = w->m, n = w->n;
    for (int32_t i = 0; i < m; ++i) w->row_norm[i] = 1.0;
    for (int32_t j = 0; j < n; ++j) w->col_norm[j] = 1.0;
    w->b_scale = w->c_scale = 1.0;
    if (w->q_kind == HPRSOCP_Q_FACTORED) {
        /* operator Q: skip ALL scaling (scaling.jl:797-846) */
        w->norm_b_org = m > 0 ? jmax(bound_norm(w, 1), soc_rhs_inf(w)) : 0.0;
        w->norm_c_org = absmax(w->c, n);
        w->norm_b = m > 0 ? jmax(bound_norm(w, 0), soc_rhs_inf(w)) : 0.0;
        w->norm_c = nrm2(w->c, n);
        return 0;
    }
    w->norm_b_org = jmax(bound_norm(w, 1), soc_rhs_inf(w));
    w->norm_c_org = absmax(w->c, n);
    double *trow = (double *)malloc(sizeof(double) * ((size_t)m + 1));
    double *tcol = (double *)malloc(sizeof(double) * ((size_t)n + 1));
    for (int32_t i = 0; i < m; ++i) trow[i] = 1.0;
    for (int32_t j = 0; j < n; ++j) tcol[j] = 1.0;
    if (w->params.use_Ruiz_scaling)
        for (int it = 0; it < (w->params.ruiz_iterations > 0 ? w->params.ruiz_iterations : 0); ++it)
            equilibrate(w, 0, PH_RUIZ, trow, tcol);
    int s = w->params.soc_block_scaling_strategy;
    int pre_pock = (s >= HPRSOCP_SOC_PRE_POCK_SIZE_SPLIT && s <= HPRSOCP_SOC_COUNT_AWARE_SMALL_CONE_BIAS);
    if (pre_pock) /* uses_pre_pock_stage, scaling.jl:136-145 */
        for (int it = 0; it < 2; ++it) equilibrate(w, 0, PH_PRE_POCK, trow, tcol);
    if (w->params.use_Pock_Chambolle_scaling) equilibrate(w, 1, PH_POCK, trow, tcol);
    free(trow);
    free(tcol);
    if (w->params.use_bc_scaling) {
        /* compute_bc_scales, scaling.jl:721-760 */
        int t = w->params.bc_scaling_norm_type;
        double hb = 1.0;
        hb = jmax(hb, finite_bound_norm(w->AL, m, t));
        hb = jmax(hb, finite_bound_norm(w->AU, m, t));
        hb = jmax(hb, (w->ncon == 0 || m - w->m1 == 0) ? 0.0 : vector_bc_norm(w->soc_rhs, m - w->m1, t));
        hb = jclamp(hb, 1e-4, 1e4);
        double hc = jclamp(jmax(1.0, vector_bc_norm(w->c, n, t)), 1e-4, 1e4);
        double bs = hb, cs = jclamp(hc, hb / 1e2, hb / 1e-2);
        double f = bs / cs;
        for (int64_t k = 0; k < w->Q.nnz; ++k) w->Q.val[k] = w->Q.val[k] * f;
        if (m > 0) {
            for (int32_t i = 0; i < m; ++i) {
                w->AL[i] /= bs;
                w->AU[i] /= bs;
                w->soc_rhs_full[i] /= bs;
            }
            for (int32_t i = 0; i < m - w->m1; ++i) w->soc_rhs[i] /= bs;
        }
        for (int32_t j = 0; j < n; ++j) {
            w->c[j] /= cs;
            w->l[j] /= bs;
            w->u[j] /= bs;
        }
        w->b_scale = bs;
        w->c_scale = cs;
    }
    csr_free(&w->AT);
    if (csr_transpose(&w->A, &w->AT)) return HPRSOCP_ERR_ALLOC;
    if (symmetrize_Q(w)) return HPRSOCP_ERR_ALLOC;
    w->norm_b = jmax(bound_norm(w, 0), soc_rhs_inf(w));
    w->norm_c = nrm2(w->c, n);
    return 0;
}

/* check_Q_diagonal CPU, scaling.jl:1312-1325: Q_is_diag = all(colsum|Q| ≈ |diag|). */
static void check_Q_diagonal(oracle_ws *w) {
    const int32_t n = w->n;
    for (int32_t j = 0; j < n; ++j) w->diag_Q[j] = 0.0;
    if (w->q_kind == HPRSOCP_Q_FACTORED) {
        w->Q_is_diag = 0;
        return;
    }
    double *cs = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int32_t i = 0; i < n && w->Q.nnz > 0; ++i)
        for (int32_t k = w->Q.rowptr[i]; k < w->Q.rowptr[i + 1]; ++k) {
            cs[w->Q.col[k]] += fabs(w->Q.val[k]);
            if (w->Q.col[k] == i) w->diag_Q[i] += w->Q.val[k];
        }
    int ok = 1;
    const double rtol = 1.4901161193847656e-08; /* sqrt(eps), isapprox default */
    for (int32_t j = 0; j < n; ++j) {
        double a = cs[j], b = fabs(w->diag_Q[j]);
        if (!(a == b || fabs(a - b) <= rtol * jmax(fabs(a), fabs(b)))) {
            ok = 0;
            break;
        }
    }
    free(cs);
    w->Q_is_diag = ok;
}

/* unified_update_Q_factors!, unified_operations.jl:468-482. */
static void update_Q_factors(oracle_ws *w) {
    double s = w->sigma, s2 = s * s;
    for (int32_t i = 0; i < w->n; ++i) {
        w->fact2[i] = 1.0 / (1.0 + s * w->diag_Q[i]);
        w->fact[i] = s * w->fact2[i];
        w->fact1[i] = s * w->diag_Q[i] * w->fact2[i];
        w->fact_M[i] = s2 * w->fact2[i];
    }
}

/* ------------------------------------------------------------- create */
static double *vdup(const double *src, int64_t n, double fill) {
    double *d = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!d) return NULL;
    for (int64_t i = 0; i < n; ++i) d[i] = src ? src[i] : fill;
    return d;
}
static double *vzero(int64_t n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static int validate_csr(oracle_ws *w, const char *name, int32_t nr, int32_t nc, int64_t nnz,
                        const int32_t *rp, const int32_t *ci, const double *v) {
    if (nnz < 0) return set_err(w, "%s: negative nnz", name);
    if (nnz == 0 && rp == NULL) return 0;
    if (!rp) return set_err(w, "%s: missing rowptr", name);
    if (nnz > 0 && (!ci || !v)) return set_err(w, "%s: missing colidx/values", name);
    if (rp[0] != 0 || rp[nr] != nnz) return set_err(w, "%s: rowptr must start at 0 and end at nnz", name);
    for (int32_t i = 0; i < nr; ++i)
        if (rp[i + 1] < rp[i]) return set_err(w, "%s: rowptr not monotone at row %d", name, i);
    for (int64_t k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= nc) return set_err(w, "%s: column index out of range at %lld", name, (long long)k);
    return 0;
}

int oracle_create(const hprsocp_problem *p, const hprsocp_params *params, oracle_ws **out) {
    double t0 = now_s();
    *out = NULL;
    oracle_ws *w = (oracle_ws *)calloc(1, sizeof(oracle_ws));
    if (!w) return HPRSOCP_ERR_ALLOC;
    *out = w;
    if (params) w->params = *params;
    else oracle_default_params(&w->params);
    const int32_t m = p->m, n = p->n;
    if (m < 0 || n < 0) return set_err(w, "negative dimensions");
    int rc = validate_csr(w, "A", m, n, p->A_nnz, p->A_rowptr, p->A_colidx, p->A_val);
    if (rc) return rc;
    if (p->number_linear_con < 0 || p->number_linear_con > m) return set_err(w, "number_linear_con out of range");
    if (p->number_lu_x < 0 || p->number_lu_x > n) return set_err(w, "number_lu_x out of range");
    if (!p->c || !p->l || !p->u || (m > 0 && (!p->AL || !p->AU))) return set_err(w, "missing vector");
    w->m = m;
    w->n = n;
    w->m1 = p->number_linear_con;
    w->n1 = p->number_lu_x;
    w->ncon = p->n_soc_con;
    w->nvar = p->n_soc_var;
    w->con_idx = (int32_t *)malloc(sizeof(int32_t) * ((size_t)w->ncon + 1));
    w->var_idx = (int32_t *)malloc(sizeof(int32_t) * ((size_t)w->nvar + 1));
    /* SOC_con_idx must start at number_linear_con and end at m (builders.jl:57-58) */
    if (w->ncon > 0 && !p->soc_con_idx) return set_err(w, "missing soc_con_idx");
    if (p->soc_con_idx) memcpy(w->con_idx, p->soc_con_idx, sizeof(int32_t) * ((size_t)w->ncon + 1));
    else w->con_idx[0] = m;
    if (w->con_idx[0] != w->m1 || w->con_idx[w->ncon] != m)
        return set_err(w, "soc_con_idx must start at number_linear_con and end at m");
    if (w->nvar > 0 && !p->soc_var_idx) return set_err(w, "missing soc_var_idx");
    if (p->soc_var_idx) memcpy(w->var_idx, p->soc_var_idx, sizeof(int32_t) * ((size_t)w->nvar + 1));
    else w->var_idx[0] = n;
    if (w->var_idx[0] != w->n1 || w->var_idx[w->nvar] != n)
        return set_err(w, "soc_var_idx must start at number_lu_x and end at n");
    for (int32_t i = 0; i < w->ncon; ++i)
        if (w->con_idx[i + 1] < w->con_idx[i]) return set_err(w, "soc_con_idx not monotone");
    for (int32_t i = 0; i < w->nvar; ++i)
        if (w->var_idx[i + 1] < w->var_idx[i]) return set_err(w, "soc_var_idx not monotone");
    if (w->params.sigma != -1.0 && !(w->params.sigma > 0.0))
        return set_err(w, "Invalid sigma value: %g. It should be a positive number or -1 for automatic.", w->params.sigma);

    if (csr_copy(&w->A, m, n, p->A_nnz, p->A_rowptr, p->A_colidx, p->A_val)) return HPRSOCP_ERR_ALLOC;
    if (csr_transpose(&w->A, &w->AT)) return HPRSOCP_ERR_ALLOC;
    w->q_kind = p->Q_kind;
    if (p->Q_kind == HPRSOCP_Q_FACTORED) {
        rc = validate_csr(w, "F", p->F_rows, n, p->F_nnz, p->F_rowptr, p->F_colidx, p->F_val);
        if (rc) return rc;
        if (csr_copy(&w->F, p->F_rows, n, p->F_nnz, p->F_rowptr, p->F_colidx, p->F_val)) return HPRSOCP_ERR_ALLOC;
        if (csr_transpose(&w->F, &w->FT)) return HPRSOCP_ERR_ALLOC;
        w->Fdiag = p->Q_diag ? vdup(p->Q_diag, n, 0.0) : NULL;
        w->Ftmp = vzero(p->F_rows);
        w->has_quad = 1; /* has_quadratic_terms(::AbstractQOperator) = true */
        csr_alloc(&w->Q, n, n, 0);
    } else {
        rc = validate_csr(w, "Q", n, n, p->Q_nnz, p->Q_rowptr, p->Q_colidx, p->Q_val);
        if (rc) return rc;
        if (p->Q_nnz > 0) {
            if (csr_copy(&w->Q, n, n, p->Q_nnz, p->Q_rowptr, p->Q_colidx, p->Q_val)) return HPRSOCP_ERR_ALLOC;
        } else csr_alloc(&w->Q, n, n, 0);
        w->has_quad = p->Q_nnz > 0; /* has_quadratic_terms: nnz(Q) > 0 */
    }
    w->obj_const = p->obj_constant;
    w->c = vdup(p->c, n, 0);
    w->AL = vdup(p->AL, m, 0);
    w->AU = vdup(p->AU, m, 0);
    w->l = vdup(p->l, n, 0);
    w->u = vdup(p->u, n, 0);
    w->soc_rhs = vdup(p->soc_rhs, m - w->m1, 0.0);
    w->soc_rhs_full = vzero(m);
    for (int32_t i = w->m1; i < m; ++i) w->soc_rhs_full[i] = w->soc_rhs[i - w->m1];
    w->row_norm = vzero(m);
    w->col_norm = vzero(n);
    w->diag_Q = vzero(n);
    rc = do_scaling(w);
    if (rc) return rc;
    check_Q_diagonal(w);
    /* allocate_workspace, algorithm.jl:1443-1588 */
    double **vn[] = {&w->w, &w->w_bar, &w->dw, &w->x, &w->x_hat, &w->x_bar, &w->dx, &w->z_bar,
                     &w->Rd, &w->ATy, &w->ATy_bar, &w->ATdy, &w->QATdy, &w->Qw, &w->Qw_bar,
                     &w->Qx, &w->dQw, &w->last_x, &w->last_w, &w->last_ATy, &w->tempv,
                     &w->fact1, &w->fact2, &w->fact, &w->fact_M};
    for (size_t k = 0; k < sizeof(vn) / sizeof(vn[0]); ++k)
        if (!(*vn[k] = vzero(n))) return HPRSOCP_ERR_ALLOC;
    double **vm[] = {&w->y, &w->y_bar, &w->dy, &w->Rp, &w->Ax, &w->s, &w->last_y};
    for (size_t k = 0; k < sizeof(vm) / sizeof(vm[0]); ++k)
        if (!(*vm[k] = vzero(m))) return HPRSOCP_ERR_ALLOC;
    if (w->params.sigma == -1.0) {
        double nb = w->norm_b, nc = w->norm_c;
        w->sigma = (nc > 1e-16 && nb > 1e-16 && nb < 1e16 && nc < 1e16) ? nb / nc : 1.0;
    } else w->sigma = w->params.sigma;
    if (w->Q_is_diag) update_Q_factors(w); /* solve() refreshes them, algorithm.jl:2816-2821 */
    w->to_check = 1;
    w->setup_time = now_s() - t0;
    return 0;
}

void oracle_destroy(oracle_ws *w) {
    if (!w) return;
    csr_free(&w->A);
    csr_free(&w->AT);
    csr_free(&w->Q);
    csr_free(&w->F);
    csr_free(&w->FT);
    double *vs[] = {w->Fdiag, w->Ftmp, w->diag_Q, w->c, w->AL, w->AU, w->l, w->u, w->soc_rhs,
                    w->soc_rhs_full, w->row_norm, w->col_norm, w->w, w->w_bar, w->dw, w->x,
                    w->x_hat, w->x_bar, w->dx, w->y, w->y_bar, w->dy, w->z_bar, w->Rp, w->Rd,
                    w->Ax, w->ATy, w->ATy_bar, w->ATdy, w->QATdy, w->s, w->Qw, w->Qw_bar, w->Qx,
                    w->dQw, w->last_x, w->last_y, w->last_w, w->last_ATy, w->tempv, w->fact1,
                    w->fact2, w->fact, w->fact_M};
    for (size_t k = 0; k < sizeof(vs) / sizeof(vs[0]); ++k) free(vs[k]);
    free(w->con_idx);
    free(w->var_idx);
    free(w);
}

/* ------------------------------------------------------ power iteration */
static uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Start vector of power_iteration_A/Q (eigenvalues.jl:76,159): standard
 * normals + 1e-8.  Julia's MersenneTwister(1) stream is not reproducible
 * here; this is SplitMix64(seed 1) + Box-Muller, shared bit-for-bit with the
 * backend's host setup. */
void oracle_start_vector(int64_t len, double *out) {
    uint64_t st = 1;
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t i = 0; i < len; i += 2) {
        double u1 = ((double)(splitmix64(&st) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        double u2 = ((double)(splitmix64(&st) >> 11)) * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log(u1));
        out[i] = r * cos(two_pi * u2) + 1e-8;
        if (i + 1 < len) out[i + 1] = r * sin(two_pi * u2) + 1e-8;
    }
}

/* power_iteration_A, eigenvalues.jl:68-118 (max 5000 iterations, tol 1e-4). */
static double power_A(oracle_ws *w) {
    const int32_t m = w->m, n = w->n;
    double *z = vzero(m), *q = vzero(m), *atq = vzero(n);
    oracle_start_vector(m, z);
    double lam = 1.0;
    for (int it = 0; it < 5000; ++it) {
        double nq = nrm2(z, m);
        for (int32_t i = 0; i < m; ++i) q[i] = z[i] / nq;
        csr_mv(&w->AT, q, atq);
        csr_mv(&w->A, atq, z);
        lam = dot(q, z, m);
        for (int32_t i = 0; i < m; ++i) q[i] = z[i] - lam * q[i];
        double err = nrm2(q, m) / (nrm2(z, m) + lam);
        if (err < 1e-4) break;
    }
    free(z);
    free(q);
    free(atq);
    return lam;
}

/* power_iteration_Q, eigenvalues.jl:152-196. */
static double power_Q(oracle_ws *w) {
    const int32_t n = w->n;
    double *z = vzero(n), *q = vzero(n);
    oracle_start_vector(n, z);
    double lam = 1.0;
    for (int it = 0; it < 5000; ++it) {
        double nq = nrm2(z, n);
        for (int32_t i = 0; i < n; ++i) q[i] = z[i] / nq;
        qmap(w, q, z);
        lam = dot(q, z, n);
        for (int32_t i = 0; i < n; ++i) q[i] = z[i] - lam * q[i];
        double err = nrm2(q, n) / (nrm2(z, n) + lam);
        if (err < 1e-4) break;
    }
    free(z);
    free(q);
    return lam;
}

/* estimate_eigenvalues, algorithm.jl:2434-2467 + compute_lambda_max_Q,
 * unified_operations.jl:726-741. */
int oracle_estimate_lambda(oracle_ws *w) {
    double t0 = now_s();
    w->lambda_A = w->m > 0 ? power_A(w) * 1.05 : 0.0;
    if (w->q_kind == HPRSOCP_Q_FACTORED) w->lambda_Q = power_Q(w) * 1.05;
    else if (w->Q.nnz > 0) {
        if (!w->Q_is_diag) w->lambda_Q = power_Q(w) * 1.05;
        else {
            double mx = w->Q.val[0];
            for (int64_t k = 1; k < w->Q.nnz; ++k) mx = jmax(mx, w->Q.val[k]);
            w->lambda_Q = mx;
        }
    } else w->lambda_Q = 0.0;
    w->lambda_set = 1;
    w->power_time = now_s() - t0;
    return 0;
}

/* process_initial_points!, algorithm.jl:2612-2677. */
int oracle_set_initial(oracle_ws *w, const double *x0, const double *y0) {
    const int32_t m = w->m, n = w->n;
    if (x0) {
        for (int32_t j = 0; j < n; ++j) {
            double v = x0[j] * w->col_norm[j] / w->b_scale;
            w->x[j] = w->x_bar[j] = w->last_x[j] = v;
            if (w->has_quad) w->w[j] = w->w_bar[j] = w->last_w[j] = v;
        }
    }
    if (y0) {
        for (int32_t i = 0; i < m; ++i) {
            double v = y0[i];
            v = v * w->row_norm[i] / w->c_scale;
            w->y[i] = w->y_bar[i] = w->last_y[i] = v;
        }
        if (m > 0) {
            if (w->has_quad) {
                csr_mv(&w->AT, w->y_bar, w->ATy_bar);
                memcpy(w->ATy, w->ATy_bar, sizeof(double) * (size_t)n);
                memcpy(w->last_ATy, w->ATy_bar, sizeof(double) * (size_t)n);
            } else csr_mv(&w->AT, w->y_bar, w->ATy);
        }
        if (w->has_quad) qmap(w, w->x_bar, w->Qx);
        for (int32_t j = 0; j < n; ++j) {
            double tmp = w->has_quad ? (-w->Qx[j] + w->ATy_bar[j] - w->c[j]) : (w->ATy[j] - w->c[j]);
            double z_raw = w->x_bar[j] + w->sigma * tmp;
            w->z_bar[j] = (w->x_bar[j] - z_raw) / w->sigma;
        }
        if (m > 0) {
            csr_mv(&w->A, w->x_bar, w->Ax);
            double f1 = w->lambda_A * w->sigma;
            for (int32_t i = 0; i < m; ++i)
                w->s[i] = jmin(jmax(w->Ax[i] - f1 * w->y[i], w->AL[i]), w->AU[i]);
        }
    }
    return 0;
}

/* ------------------------------------------------------------- updates */
/* Cone part shared by the x-side updates (kernels.jl:7585-7663 and
 * 7748-7811): z_raw is computed by the caller into zr[start..end). */
static void x_cone_update(oracle_ws *w, int32_t s, int32_t e, const double *zr, double h1, double h2,
                          int qp, double f1s, double f2s) {
    double *x = w->x, *xb = w->x_bar, *xh = w->x_hat, *zb = w->z_bar, *dx = w->dx;
    const double sig = w->sigma;
    double t = zr[s], ns = 0.0;
    for (int32_t j = s + 1; j < e; ++j) ns += zr[j] * zr[j];
    ns = sqrt(ns);
    double alpha = 0.0;
    int cas;
    if (ns <= -t) cas = 0;
    else if (ns <= t) cas = 1;
    else {
        cas = 2;
        alpha = ((ns + t) / 2.0) / ns;
    }
    for (int32_t j = s; j < e; ++j) {
        double raw = zr[j];
        double p = cas == 0 ? 0.0 : (cas == 1 ? raw : (j == s ? (ns + t) / 2.0 : alpha * raw));
        xb[j] = p;
        dx[j] = p - x[j];
        xh[j] = 2.0 * p - x[j];
        zb[j] = (p - raw) / sig;
        x[j] = h2 * xh[j] + h1 * w->last_x[j];
        if (qp)
            w->w_bar[j] = w->Q_is_diag ? w->fact1[j] * w->w[j] + w->fact2[j] * xh[j]
                                        : f1s * w->w[j] + f2s * xh[j];
    }
}

/* unified_update_zx_cpu! (noQ), kernels.jl:7672-7812. */
static void update_zx_noq(oracle_ws *w, double h1, double h2) {
    const int32_t n = w->n;
    csr_mv(&w->AT, w->y, w->ATy);
    const double sig = w->sigma;
    int32_t ln = w->nvar > 0 ? w->n1 : n;
    for (int32_t i = 0; i < ln; ++i) {
        double xi = w->x[i];
        double z_raw = xi + sig * (w->ATy[i] - w->c[i]);
        double xb = jmin(jmax(z_raw, w->l[i]), w->u[i]);
        double xh = 2.0 * xb - xi;
        if (w->to_check) {
            w->dx[i] = xb - xi;
            w->x_bar[i] = xb;
            w->z_bar[i] = (xb - z_raw) / sig;
        }
        w->x[i] = h2 * xh + h1 * w->last_x[i];
        w->x_hat[i] = xh;
    }
    for (int32_t k = 0; k < w->nvar; ++k) {
        int32_t s = w->var_idx[k], e = w->var_idx[k + 1];
        for (int32_t j = s; j < e; ++j) w->tempv[j] = w->x[j] + sig * (w->ATy[j] - w->c[j]);
        x_cone_update(w, s, e, w->tempv, h1, h2, 0, 0, 0);
    }
}

/* Cone part of the y updates (kernels.jl:7895-7951, 8039-8095). */
static void y_cone_update(oracle_ws *w, int32_t s, int32_t e, double f1, double f2, double h1, double h2) {
    double *y = w->y, *yb = w->y_bar, *dy = w->dy, *sv = w->s;
    for (int32_t j = s; j < e; ++j) sv[j] = w->Ax[j] - w->soc_rhs_full[j] - f1 * y[j];
    double t = sv[s], ns = 0.0;
    for (int32_t j = s + 1; j < e; ++j) ns += sv[j] * sv[j];
    ns = sqrt(ns);
    int cas;
    double alpha = 0.0;
    if (ns <= -t) cas = 0;
    else if (ns <= t) cas = 1;
    else {
        cas = 2;
        alpha = ((ns + t) / 2.0) / ns;
    }
    for (int32_t j = s; j < e; ++j) {
        double raw = sv[j];
        double p = cas == 0 ? 0.0 : (cas == 1 ? raw : (j == s ? (ns + t) / 2.0 : alpha * raw));
        sv[j] = p;
        double ybv = f2 * (p - raw);
        yb[j] = ybv;
        dy[j] = ybv - y[j];
        y[j] = h2 * (2.0 * ybv - y[j]) + h1 * w->last_y[j];
    }
}

/* Linear rows of the y updates (kernels.jl:7841-7893, 8005-8037). */
static void y_linear_update(oracle_ws *w, double f1, double f2, double h1, double h2, int write_ybar) {
    for (int32_t i = 0; i < w->m1; ++i) {
        double yi = w->y[i];
        double s_raw = w->Ax[i] - f1 * yi;
        double sp = jmin(jmax(s_raw, w->AL[i]), w->AU[i]);
        double yb = f2 * (sp - s_raw);
        double yh = 2.0 * yb - yi;
        if (w->to_check) {
            w->s[i] = sp;
            w->dy[i] = yb - yi;
        }
        if (w->to_check || write_ybar) w->y_bar[i] = yb;
        w->y[i] = h2 * yh + h1 * w->last_y[i];
    }
}

/* unified_update_y_noQ_cpu!, kernels.jl:7817-7952. */
static void update_y_noq(oracle_ws *w, double h1, double h2) {
    if (w->m == 0) return;
    csr_mv(&w->A, w->x_hat, w->Ax);
    double f1 = w->lambda_A * w->sigma, f2 = 1.0 / f1;
    y_linear_update(w, f1, f2, h1, h2, 0);
    for (int32_t k = 0; k < w->ncon; ++k) y_cone_update(w, w->con_idx[k], w->con_idx[k + 1], f1, f2, h1, h2);
}

/* update_zxw1_cpu!, kernels.jl:7487-7664. */
static void update_zxw1(oracle_ws *w, double h1, double h2) {
    const int32_t n = w->n;
    qmap(w, w->w, w->Qw);
    double f2s = 1.0 / (1.0 + w->sigma * w->lambda_Q), f1s = 1.0 - f2s;
    const double sig = w->sigma;
    int32_t ln = w->nvar > 0 ? w->n1 : n;
    for (int32_t i = 0; i < ln; ++i) {
        double xi = w->x[i];
        double z_raw = xi + sig * (-w->Qw[i] + w->ATy[i] - w->c[i]);
        double xb = jmin(jmax(z_raw, w->l[i]), w->u[i]);
        double xh = 2.0 * xb - xi;
        double wb = w->Q_is_diag ? w->fact1[i] * w->w[i] + w->fact2[i] * xh : f1s * w->w[i] + f2s * xh;
        if (w->to_check) {
            w->dx[i] = xb - xi;
            w->x_bar[i] = xb;
            w->z_bar[i] = (xb - z_raw) / sig;
        }
        w->x[i] = h2 * xh + h1 * w->last_x[i];
        w->x_hat[i] = xh;
        w->w_bar[i] = wb;
    }
    for (int32_t k = 0; k < w->nvar; ++k) {
        int32_t s = w->var_idx[k], e = w->var_idx[k + 1];
        for (int32_t j = s; j < e; ++j) w->tempv[j] = w->x[j] + sig * (-w->Qw[j] + w->ATy[j] - w->c[j]);
        x_cone_update(w, s, e, w->tempv, h1, h2, 1, f1s, f2s);
    }
}

/* update_y_cpu!, kernels.jl:7976-8096. */
static void update_y_qp(oracle_ws *w, double h1, double h2) {
    if (w->m == 0) return;
    const int32_t n = w->n;
    qmap(w, w->w_bar, w->Qw_bar);
    for (int32_t j = 0; j < n; ++j) w->tempv[j] = w->x_hat[j] + w->sigma * (w->Qw[j] - w->Qw_bar[j]);
    csr_mv(&w->A, w->tempv, w->Ax);
    double f1 = w->lambda_A * w->sigma, f2 = 1.0 / f1;
    y_linear_update(w, f1, f2, h1, h2, 1);
    for (int32_t k = 0; k < w->ncon; ++k) y_cone_update(w, w->con_idx[k], w->con_idx[k + 1], f1, f2, h1, h2);
}

/* update_w2_cpu!, kernels.jl:8100-8185. */
static void update_w2(oracle_ws *w, double h1, double h2) {
    csr_mv(&w->AT, w->y_bar, w->ATy_bar);
    double fs = w->sigma / (1.0 + w->sigma * w->lambda_Q);
    for (int32_t i = 0; i < w->n; ++i) {
        double wi = w->w[i], atyi = w->ATy[i], atyb = w->ATy_bar[i];
        double f = w->Q_is_diag ? w->fact[i] : fs;
        double wbn = w->w_bar[i] + f * (atyb - atyi);
        double wn = h1 * w->last_w[i] + h2 * (2.0 * wbn - wi);
        double an = h1 * w->last_ATy[i] + h2 * (2.0 * atyb - atyi);
        w->w[i] = wn;
        w->ATy[i] = an;
        if (w->to_check) {
            w->w_bar[i] = wbn;
            w->dw[i] = wbn - wi;
            w->ATdy[i] = atyb - atyi;
        }
    }
}

/* main_update_cpu!, algorithm.jl:1642-1665. */
int oracle_update(oracle_ws *w, int64_t inner, int32_t to_check) {
    if (!w->lambda_set) {
        snprintf(w->err, sizeof(w->err), "estimate_lambda must run before update");
        return HPRSOCP_ERR_STATE;
    }
    w->to_check = to_check ? 1 : 0;
    double h1 = 1.0 / ((double)inner + 2.0), h2 = 1.0 - h1;
    if (w->has_quad) {
        update_zxw1(w, h1, h2);
        update_y_qp(w, h1, h2);
        update_w2(w, h1, h2);
    } else {
        update_zx_noq(w, h1, h2);
        update_y_noq(w, h1, h2);
    }
    return 0;
}

/* ------------------------------------------------------------ residuals */
/* compute_bounds_violation! CPU, algorithm.jl:218-256. */
static double bounds_violation(oracle_ws *w) {
    const int32_t n = w->n;
    for (int32_t j = 0; j < n; ++j) {
        double lo = jmax(w->l[j] - w->x_bar[j], 0.0), up = jmax(w->x_bar[j] - w->u[j], 0.0);
        w->dx[j] = (lo + up) * (w->b_scale / w->col_norm[j]);
    }
    for (int32_t k = 0; k < w->nvar; ++k) {
        int32_t s = w->var_idx[k], e = w->var_idx[k + 1];
        if (s >= e) continue;
        double t = w->x_bar[s], ns = 0.0;
        w->dx[s] = t;
        for (int32_t j = s + 1; j < e; ++j) {
            w->dx[j] = w->x_bar[j];
            ns += w->x_bar[j] * w->x_bar[j];
        }
        ns = sqrt(ns);
        if (ns <= -t) {
        } else if (ns <= t) {
            for (int32_t j = s; j < e; ++j) w->dx[j] = 0.0;
        } else {
            double f = (1 + t / ns) / 2;
            w->dx[s] = (ns + t) / 2 - t;
            for (int32_t j = s + 1; j < e; ++j) w->dx[j] = f * w->dx[j] - w->dx[j];
        }
        for (int32_t j = s; j < e; ++j) w->dx[j] *= w->b_scale / w->col_norm[j];
    }
    return absmax(w->dx, n);
}

/* compute_residuals!, algorithm.jl:113-196, with compute_Rd!/compute_Rp!
 * (kernels.jl:118-292) and _compute_primal_residual_components (:77-93). */
int oracle_compute_residuals(oracle_ws *w, int64_t iter, hprsocp_residuals *r) {
    const int32_t m = w->m, n = w->n;
    const double bs = w->b_scale, cs = w->c_scale;
    int skip_q = !w->has_quad;
    double qx_dot = 0.0;
    if (!skip_q) {
        qmap(w, w->x_bar, w->Qx);
        qx_dot = dot(w->x_bar, w->Qx, n);
    }
    r->primal_obj = bs * cs * (dot(w->c, w->x_bar, n) + 0.5 * qx_dot) + w->obj_const;
    r->dual_obj = bs * cs * (-0.5 * qx_dot + dot(w->z_bar, w->x_bar, n)) + w->obj_const;
    if (m > 0) r->dual_obj += bs * cs * (dot(w->y_bar, w->s, m) + dot(w->y_bar, w->soc_rhs_full, m));
    r->rel_gap = fabs(r->primal_obj - r->dual_obj) / (1.0 + jmax(fabs(r->primal_obj), fabs(r->dual_obj)));
    /* compute_Rd! */
    if (m > 0) csr_mv(&w->AT, w->y_bar, w->ATdy);
    for (int32_t j = 0; j < n; ++j) {
        double cn = w->col_norm[j];
        double c_org = w->c[j] * cn * cs, at = w->ATdy[j] * cn * cs, z = w->z_bar[j] * cn * cs;
        if (!skip_q) {
            double qx = w->Qx[j] * cn * cs;
            w->Rd[j] = qx + c_org - at - z;
            w->Qx[j] = qx;
        } else w->Rd[j] = c_org - at - z;
        w->ATdy[j] = at;
    }
    double qx_inf = skip_q ? 0.0 : absmax(w->Qx, n);
    r->err_Rd = absmax(w->Rd, n) / (1.0 + jmax(jmax(w->norm_c_org, absmax(w->ATdy, n)), qx_inf));
    if (m > 0) {
        csr_mv(&w->A, w->x_bar, w->Ax);
        for (int32_t i = 0; i < m; ++i) {
            double rn = w->row_norm[i];
            double ax = w->Ax[i] * rn * bs, al = w->AL[i] * rn * bs, au = w->AU[i] * rn * bs;
            double pr = jmin(jmax(ax, al), au);
            w->Rp[i] = pr - ax;
            w->Ax[i] = ax;
        }
        for (int32_t k = 0; k < w->ncon; ++k) {
            int32_t s = w->con_idx[k], e = w->con_idx[k + 1];
            if (s >= e) continue;
            double t = w->Ax[s] - w->soc_rhs_full[s] * w->row_norm[s] * bs, ns = 0.0;
            for (int32_t j = s + 1; j < e; ++j) {
                double sj = w->Ax[j] - w->soc_rhs_full[j] * w->row_norm[j] * bs;
                ns += sj * sj;
                w->Rp[j] = sj;
            }
            ns = sqrt(ns);
            w->Rp[s] = t;
            if (ns <= -t) {
            } else if (ns <= t) {
                for (int32_t j = s; j < e; ++j) w->Rp[j] = 0.0;
            } else {
                double f = (1 + t / ns) / 2;
                w->Rp[s] = (ns + t) / 2 - t;
                for (int32_t j = s + 1; j < e; ++j) w->Rp[j] = f * w->Rp[j] - w->Rp[j];
            }
        }
        double denom = 1.0 + jmax(w->norm_b_org, absmax(w->Ax, m));
        r->err_Rp = absmax(w->Rp, m) / denom;
        r->err_Rp_linear = w->m1 > 0 ? absmax(w->Rp, w->m1) / denom : 0.0;
        r->err_Rp_soc = w->m1 < m ? absmax(w->Rp + w->m1, m - w->m1) / denom : 0.0;
    } else {
        r->err_Rp = r->err_Rp_linear = r->err_Rp_soc = 0.0;
    }
    if (iter == 0) r->err_Rp = jmax(r->err_Rp, bounds_violation(w));
    r->kkt = jmax(jmax(r->err_Rp, r->err_Rd), r->rel_gap);
    return 0;
}

/* compute_M_norm!, algorithm.jl:1302-1359. */
int oracle_mnorm(oracle_ws *w, double *out) {
    const int32_t m = w->m, n = w->n;
    const double sig = w->sigma;
    double M1 = 0.0, M2 = (1.0 / sig) * dot(w->dx, w->dx, n), M3 = 0.0;
    int skip_q = !w->has_quad;
    if (!skip_q) {
        qmap(w, w->dw, w->dQw);
        M2 -= 2.0 * dot(w->dQw, w->dx, n);
    }
    if (m > 0) {
        M1 = sig * w->lambda_A * dot(w->dy, w->dy, m);
        csr_mv(&w->AT, w->dy, w->ATdy);
        if (!skip_q) {
            qmap(w, w->ATdy, w->QATdy);
            M1 -= 2.0 * sig * dot(w->dQw, w->ATdy, n);
        }
        M2 += 2.0 * dot(w->ATdy, w->dx, n);
        if (!skip_q) {
            if (w->Q_is_diag) {
                for (int32_t j = 0; j < n; ++j) w->ATdy[j] *= w->fact_M[j];
                M3 = dot(w->ATdy, w->QATdy, n);
                M1 += sig * dot(w->dQw, w->dQw, n);
            } else {
                M3 = (sig * sig) / (1.0 + sig * w->lambda_Q) * dot(w->ATdy, w->QATdy, n);
                M1 += sig * w->lambda_Q * dot(w->dw, w->dQw, n);
            }
        }
    } else if (!w->Q_is_diag && !skip_q) {
        M1 = sig * w->lambda_Q * dot(w->dw, w->dQw, n);
    }
    M2 += jmax(M1, 0.0);
    double Mn = jmax(M2, 0.0) + jmax(M3, 0.0);
    *out = sqrt(Mn);
    return 0;
}

/* ------------------------------------------------------ restart / sigma */
/* HPRSOCP_restart (structs.jl:600-623) is the public hprsocp_restart_info. */
typedef hprsocp_restart_info restart_t;

/* initialize_restart, algorithm.jl:1886-1908. */
static void init_restart(restart_t *r) {
    memset(r, 0, sizeof(*r));
    r->first_restart = 1;
    r->save_gap = r->current_gap = r->last_gap = r->best_gap = r->best_kkt = INFINITY;
    r->best_sigma = 1.0;
    r->weighted_norm = INFINITY;
}

/* should_refresh_best_sigma, algorithm.jl:682-693. */
static int refresh_best(double cg, double bg, double ck, double bk, int64_t bi) {
    int improved = ck < bk;
    int compat = cg < bg && ck <= 1.01 * bk;
    int periodic = bi >= 10 && ck <= 1.2 * bk;
    return improved || compat || periodic;
}

/* check_restart, algorithm.jl:1016-1099. */
static void check_restart(restart_t *r, int64_t iter, int64_t check_iter, double sigma, double kkt) {
    r->restart_flag = 0;
    if (r->first_restart) {
        if (iter == check_iter) {
            r->first_restart = 0;
            r->restart_flag = 1;
            r->weighted_norm = r->current_gap;
            r->best_gap = r->current_gap;
            r->best_kkt = kkt;
            r->best_sigma = sigma;
        }
        return;
    }
    if (iter % check_iter != 0) return;
    if (r->current_gap < 0) r->current_gap = 1e-6;
    if (r->current_gap <= 0.36 * r->last_gap) {
        r->sufficient++;
        r->restart_flag = 1;
    }
    if (r->current_gap <= 0.8 * r->last_gap && r->current_gap > 1.00 * r->save_gap) {
        r->necessary++;
        r->restart_flag = 2;
    }
    double fact = 0.15;
    if (((double)r->inner >= jmax(500.0, fact * (double)iter) ||
         (r->inner >= 10000 && r->current_gap <= 0.75 * r->last_gap)) &&
        r->current_gap > 0.9 * r->save_gap) {
        r->long_restarts++;
        r->restart_flag = 3;
    }
    if (refresh_best(r->current_gap, r->best_gap, kkt, r->best_kkt, r->best_iter)) {
        r->best_gap = r->current_gap;
        r->best_kkt = kkt;
        r->best_sigma = sigma;
        r->best_iter = 0;
    }
    r->save_gap = r->current_gap;
}

static double f_dev(double x, double a, double b, double c, double d) {
    return a * x + b / x + c * x * x / (1 + d * x);
}

/* golden, kernels.jl:6272-6305 (maxiter 200). */
double oracle_golden(double ap, double bp, double cp, double dp, double lo, double hi, double tol) {
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

/* unified_golden_Q_diag CPU, unified_operations.jl:371-409. */
static double golden_q_diag(oracle_ws *w, double a, double b, double lo, double hi, double tol) {
    const double phi = (sqrt(5.0) - 1.0) / 2.0;
    const int32_t n = w->n;
#define FQ(xx)                                                                        \
    ({                                                                                \
        double _x = (xx);                                                             \
        for (int32_t _j = 0; _j < n; ++_j) w->tempv[_j] = w->QATdy[_j] / (1.0 + _x * w->diag_Q[_j]); \
        a * _x + b / _x + _x * _x * dot(w->ATdy, w->tempv, n);                        \
    })
    double x1 = hi - phi * (hi - lo), x2 = lo + phi * (hi - lo);
    double f1 = FQ(x1), f2 = FQ(x2);
    int it = 0;
    while (fabs(hi - lo) > tol * jmax(1.0, fabs(lo)) && it < 200) {
        if (f1 > f2) {
            lo = x1;
            x1 = x2;
            f1 = f2;
            x2 = lo + phi * (hi - lo);
            f2 = FQ(x2);
        } else {
            hi = x2;
            x2 = x1;
            f2 = f1;
            x1 = hi - phi * (hi - lo);
            f1 = FQ(x1);
        }
        it++;
    }
#undef FQ
    return (f1 < f2) ? x1 : x2;
}

/* update_sigma!, algorithm.jl:729-990. */
static void update_sigma(oracle_ws *w, restart_t *r, const hprsocp_residuals *res) {
    if (r->restart_flag < 1) return;
    const int32_t m = w->m, n = w->n;
    int skip_q = !w->has_quad, has_m = m > 0, qd = w->Q_is_diag;
    double sigma_old = w->sigma;
    const double SMIN = 1e-12, SMAX = 1e12, THETA_MAX = 1.0, IMB = 1e6, FLOOR = 2.0, THR = 1e-8,
                 EPS = 1e-16, LOG100 = log(100.0);
    if (!skip_q) {
        for (int32_t j = 0; j < n; ++j) w->dw[j] = w->w_bar[j] - w->last_w[j];
        qmap(w, w->dw, w->dQw);
    }
    for (int32_t j = 0; j < n; ++j) w->dx[j] = w->x_bar[j] - w->last_x[j];
    double a = 0.0, b = dot(w->dx, w->dx, n), c = 0.0, d = 0.0;
    if (has_m) {
        for (int32_t i = 0; i < m; ++i) w->dy[i] = w->y_bar[i] - w->last_y[i];
        if (!skip_q) {
            for (int32_t j = 0; j < n; ++j) w->ATdy[j] = w->ATy_bar[j] - w->last_ATy[j];
            qmap(w, w->ATdy, w->QATdy);
            a = w->lambda_A * dot(w->dy, w->dy, m) - 2.0 * dot(w->dQw, w->ATdy, n);
        } else {
            csr_mv(&w->AT, w->dy, w->ATdy);
            a = w->lambda_A * dot(w->dy, w->dy, m);
        }
    }
    if (!skip_q) {
        if (qd) {
            double nq = nrm2(w->dQw, n);
            a += nq * nq;
        } else {
            a += w->lambda_Q * dot(w->dw, w->dQw, n);
            if (has_m) {
                c = dot(w->ATdy, w->QATdy, n);
                d = w->lambda_Q;
            }
        }
    }
    a = jmax(a, EPS);
    b = jmax(b, EPS);
    double sm;
    if (skip_q) sm = sqrt(b / a);
    else if (qd) sm = has_m ? golden_q_diag(w, a, b, 1e-12, 1e12, 1e-13) : sqrt(b / a);
    else sm = has_m ? oracle_golden(a, b, c, d, 1e-12, 1e12, 1e-13) : sqrt(b / a);

    double sref = r->best_sigma, ck = res->kkt, cg = r->current_gap, bk = r->best_kkt, bg = r->best_gap;
    int ref_ok = (ck <= 10.0 * bk) && (cg <= 10.0 * bg);
    double pr = sqrt((ck / bk) * (cg / bg));
    pr = jclamp(pr, 1.0, 10.0);
    double sbase = sm;
    if (ref_ok) {
        double theta = jclamp(exp(-0.3 * (pr - 1.0)), 0.0, THETA_MAX);
        sbase = exp(theta * log(sm) + (1.0 - theta) * log(sref));
        sbase = jclamp(sbase, SMIN, SMAX);
    }
    double ep = jmax(res->err_Rp, EPS), ed = jmax(res->err_Rd, EPS);
    double es = jmin(ep, ed), el = jmax(ep, ed), ir = el / es;
    int both = el <= THR;
    int one = (es <= 0.5 * THR) && (el >= 2.0 * THR);
    int extreme = ir >= IMB;
    int trig = !both && (one || extreme);
    int keep = !both && ((es <= 0.8 * THR) || (ir >= 0.1 * IMB));
    int cdir = ed >= ep ? 1 : -1;
    if (both) {
        r->sigma_correction_active = 0;
        r->sigma_correction_hold = -1;
        r->sigma_correction_dir = 0;
    } else if (trig || (r->sigma_correction_active && keep)) {
        r->sigma_correction_active = 1;
        r->sigma_correction_hold = 0;
        r->sigma_correction_dir = cdir;
    } else {
        r->sigma_correction_active = 0;
        r->sigma_correction_hold = -1;
        r->sigma_correction_dir = 0;
    }
    double beta = 0.0;
    if (!both) {
        if (one) beta = 0.5;
        if (extreme) beta = 1.0;
        else if (r->sigma_correction_active) beta = 0.5;
    }
    double kappa = 1.0;
    if (beta > 0.0) {
        double rk = ed / ep;
        if (r->sigma_correction_active) {
            if (r->sigma_correction_dir > 0) rk = jmax(rk, FLOOR);
            else if (r->sigma_correction_dir < 0) rk = jmin(rk, 1.0 / FLOOR);
        }
        kappa = exp(beta * jclamp(log(rk), -LOG100, LOG100));
    }
    double sn = sbase * kappa;
    if (r->sigma_correction_active) {
        if (r->sigma_correction_dir > 0) sn = jmax(sn, sigma_old * 1.5);
        else if (r->sigma_correction_dir < 0) sn = jmin(sn, sigma_old / 1.5);
    }
    sn = jclamp(sn, SMIN, SMAX);
    w->sigma = sn;
    r->best_iter += 1;
    if (qd && !skip_q && fabs(sigma_old - sn) > 1e-15) update_Q_factors(w);
}

/* do_restart!, algorithm.jl:1130-1166. */
static void do_restart(oracle_ws *w, restart_t *r) {
    if (r->restart_flag <= 0) return;
    const int32_t m = w->m, n = w->n;
    memcpy(w->x, w->x_bar, sizeof(double) * (size_t)n);
    memcpy(w->last_x, w->x_bar, sizeof(double) * (size_t)n);
    if (w->has_quad) {
        memcpy(w->w, w->w_bar, sizeof(double) * (size_t)n);
        memcpy(w->last_w, w->w_bar, sizeof(double) * (size_t)n);
    }
    if (m > 0) {
        memcpy(w->y, w->y_bar, sizeof(double) * (size_t)m);
        memcpy(w->last_y, w->y_bar, sizeof(double) * (size_t)m);
        if (w->has_quad) {
            csr_mv(&w->AT, w->y_bar, w->ATy_bar);
            memcpy(w->last_ATy, w->ATy_bar, sizeof(double) * (size_t)n);
            memcpy(w->ATy, w->ATy_bar, sizeof(double) * (size_t)n);
        }
    }
    r->last_gap = r->current_gap;
    r->save_gap = INFINITY;
    r->times += 1;
    r->inner = 0;
}

/* Seams of the solve loop for known-answer tests, same meaning as the
 * hprsocp_ entry points of the same names. */
void oracle_init_restart(hprsocp_restart_info *r) { init_restart(r); }

void oracle_check_restart(hprsocp_restart_info *r, int64_t iter, int64_t check_iter, double sigma, double kkt) {
    if (check_iter > 0) check_restart(r, iter, check_iter, sigma, kkt);
}

int oracle_update_sigma(oracle_ws *w, hprsocp_restart_info *r, const hprsocp_residuals *res) {
    update_sigma(w, r, res);
    return 0;
}

int oracle_restart(oracle_ws *w, hprsocp_restart_info *r) {
    do_restart(w, r);
    return 0;
}

/* print_step, algorithm.jl:1910-1912. */
static double print_step(int64_t iter) {
    double e = floor(log10((double)iter));
    double v = pow(10.0, e) / 10.0;
    return jmax(v, 10.0);
}
static int rem_zero(int64_t iter, double step) { return fmod((double)iter, step) == 0.0; }

static int should_print(const hprsocp_params *p, int64_t iter, double t_start) {
    if (p->print_frequency == -1)
        return rem_zero(iter, print_step(iter)) || iter == p->max_iter || (now_s() - t_start > p->time_limit);
    return (iter % p->print_frequency == 0) || iter == p->max_iter || (now_s() - t_start > p->time_limit);
}

/* solve loop, algorithm.jl:2765-2905 (after setup). */
int oracle_solve(oracle_ws *w, hprsocp_result *res) {
    const hprsocp_params *p = &w->params;
    if (p->print_frequency != -1 && p->print_frequency <= 0)
        return set_err(w, "Invalid print_frequency: %d", p->print_frequency);
    double t_start = now_s();
    if (!w->lambda_set) oracle_estimate_lambda(w);
    restart_t r;
    init_restart(&r);
    r.best_sigma = w->sigma;
    if (w->Q_is_diag) update_Q_factors(w);
    hprsocp_residuals rs;
    memset(&rs, 0, sizeof(rs));
    int64_t it4 = 0, it6 = 0;
    double t4 = 0, t6 = 0;
    int first4 = 1, first6 = 1;
    int64_t check_iter = p->check_iter;
    w->to_check = 1;
    for (int64_t iter = 0; iter <= p->max_iter; ++iter) {
        int print_yes = should_print(p, iter, t_start);
        if (iter % check_iter == 0 || print_yes) oracle_compute_residuals(w, iter, &rs);
        int status = HPRSOCP_STATUS_CONTINUE;
        if (rs.kkt < p->stoptol) status = HPRSOCP_STATUS_OPTIMAL;
        else if (iter == p->max_iter) status = HPRSOCP_STATUS_MAX_ITER;
        else if (now_s() - t_start > p->time_limit) status = HPRSOCP_STATUS_TIME_LIMIT;
        check_restart(&r, iter, check_iter, w->sigma, rs.kkt);
        update_sigma(w, &r, &rs);
        do_restart(w, &r);
        if ((print_yes || status != HPRSOCP_STATUS_CONTINUE) && p->verbose)
            printf("%5lld    %3.2e    %3.2e    %+7.6e    %+7.6e    %3.2e    %3.2e    %6.2f\n",
                   (long long)iter, rs.err_Rp, rs.err_Rd, rs.primal_obj, rs.dual_obj, rs.rel_gap,
                   w->sigma, now_s() - t_start);
        if (rs.kkt < 1e-4 && first4) {
            t4 = now_s() - t_start;
            it4 = iter;
            first4 = 0;
        }
        if (rs.kkt < 1e-6 && first6) {
            t6 = now_s() - t_start;
            it6 = iter;
            first6 = 0;
        }
        if (status != HPRSOCP_STATUS_CONTINUE) {
            double tt = now_s() - t_start;
            memset(res, 0, sizeof(*res));
            res->iter = iter;
            res->time = tt;
            res->power_time = w->power_time;
            res->setup_time = w->setup_time;
            res->residuals = rs.kkt;
            res->primal_obj = rs.primal_obj;
            res->dual_obj = rs.dual_obj;
            res->gap = rs.rel_gap;
            res->err_Rp = rs.err_Rp;
            res->err_Rp_linear = rs.err_Rp_linear;
            res->err_Rp_soc = rs.err_Rp_soc;
            res->err_Rd = rs.err_Rd;
            res->status = status;
            res->time_4 = t4 == 0.0 ? tt : t4;
            res->iter_4 = it4 == 0 ? iter : it4;
            res->time_6 = t6 == 0.0 ? tt : t6;
            res->iter_6 = it6 == 0 ? iter : it6;
            res->sigma = w->sigma;
            res->lambda_max_A = w->lambda_A;
            res->lambda_max_Q = w->lambda_Q;
            res->restarts = (int32_t)r.times;
            return 0;
        }
        int64_t nx = iter + 1;
        w->to_check = (nx % check_iter == 0) || (r.restart_flag > 0);
        if (p->print_frequency == -1) w->to_check = w->to_check || rem_zero(nx, print_step(nx));
        else w->to_check = w->to_check || (nx % p->print_frequency == 0);
        /* perform_iteration_step!, algorithm.jl:2722-2744 */
        oracle_update(w, r.inner, w->to_check);
        int need_last = r.restart_flag > 0, need_cur = ((iter + 1) % check_iter) == 0;
        if (need_last || need_cur) {
            double mn;
            oracle_mnorm(w, &mn);
            if (need_last) r.last_gap = jmin(mn, r.last_gap);
            if (need_cur) r.current_gap = mn;
        }
        r.inner += 1;
    }
    return 0;
}

/* collect_results!, algorithm.jl:1222-1256. */
int oracle_get_solution(oracle_ws *w, double *x, double *y, double *z, double *wv) {
    const int32_t m = w->m, n = w->n;
    for (int32_t j = 0; j < n; ++j) {
        if (x) x[j] = w->b_scale * (w->x_bar[j] / w->col_norm[j]);
        if (z) z[j] = w->c_scale * (w->z_bar[j] * w->col_norm[j]);
        if (wv) wv[j] = w->b_scale * ((w->has_quad ? w->w_bar[j] : w->x_bar[j]) / w->col_norm[j]);
    }
    for (int32_t i = 0; i < m && y; ++i) y[i] = w->c_scale * (w->y_bar[i] / w->row_norm[i]);
    return 0;
}

/* optimize, algorithm.jl:2538-2581 (warm-up = a 200-iteration solve). */
int oracle_optimize(const hprsocp_problem *prob, const hprsocp_params *params, hprsocp_result *res,
                    double *x, double *y, double *z, double *wv) {
    hprsocp_params p;
    if (params) p = *params;
    else oracle_default_params(&p);
    if (p.warm_up) {
        hprsocp_params q = p;
        q.max_iter = 200;
        q.verbose = 0;
        q.warm_up = 0;
        oracle_ws *ww = NULL;
        int rc = oracle_create(prob, &q, &ww);
        if (!rc) {
            oracle_estimate_lambda(ww);
            if (q.initial_x || q.initial_y) oracle_set_initial(ww, q.initial_x, q.initial_y);
            hprsocp_result tmp;
            rc = oracle_solve(ww, &tmp);
        }
        oracle_destroy(ww);
        if (rc) return rc;
    }
    oracle_ws *w = NULL;
    int rc = oracle_create(prob, &p, &w);
    if (rc) {
        oracle_destroy(w);
        return rc;
    }
    double t0 = now_s();
    oracle_estimate_lambda(w);
    if (p.initial_x || p.initial_y) oracle_set_initial(w, p.initial_x, p.initial_y);
    rc = oracle_solve(w, res);
    (void)t0;
    if (!rc) oracle_get_solution(w, x, y, z, wv);
    oracle_destroy(w);
    return rc;
}

/* ------------------------------------------------------------ inspection */
static double *vec_ptr(oracle_ws *w, int32_t id, int64_t *len) {
    const int64_t n = w->n, m = w->m;
    switch (id) {
    case HPRSOCP_V_X: *len = n; return w->x;
    case HPRSOCP_V_X_BAR: *len = n; return w->x_bar;
    case HPRSOCP_V_X_HAT: *len = n; return w->x_hat;
    case HPRSOCP_V_Z_BAR: *len = n; return w->z_bar;
    case HPRSOCP_V_DX: *len = n; return w->dx;
    case HPRSOCP_V_LAST_X: *len = n; return w->last_x;
    case HPRSOCP_V_Y: *len = m; return w->y;
    case HPRSOCP_V_Y_BAR: *len = m; return w->y_bar;
    case HPRSOCP_V_DY: *len = m; return w->dy;
    case HPRSOCP_V_LAST_Y: *len = m; return w->last_y;
    case HPRSOCP_V_S: *len = m; return w->s;
    case HPRSOCP_V_W: *len = n; return w->w;
    case HPRSOCP_V_W_BAR: *len = n; return w->w_bar;
    case HPRSOCP_V_LAST_W: *len = n; return w->last_w;
    case HPRSOCP_V_DW: *len = n; return w->dw;
    case HPRSOCP_V_ATY: *len = n; return w->ATy;
    case HPRSOCP_V_ATY_BAR: *len = n; return w->ATy_bar;
    case HPRSOCP_V_LAST_ATY: *len = n; return w->last_ATy;
    case HPRSOCP_V_C: *len = n; return w->c;
    case HPRSOCP_V_L: *len = n; return w->l;
    case HPRSOCP_V_U: *len = n; return w->u;
    case HPRSOCP_V_AL: *len = m; return w->AL;
    case HPRSOCP_V_AU: *len = m; return w->AU;
    case HPRSOCP_V_SOC_RHS_FULL: *len = m; return w->soc_rhs_full;
    case HPRSOCP_V_ROW_NORM: *len = m; return w->row_norm;
    case HPRSOCP_V_COL_NORM: *len = n; return w->col_norm;
    case HPRSOCP_V_DIAG_Q: *len = n; return w->diag_Q;
    case HPRSOCP_V_RP: *len = m; return w->Rp;
    case HPRSOCP_V_AX: *len = m; return w->Ax;
    case HPRSOCP_V_RD: *len = n; return w->Rd;
    default: *len = 0; return NULL;
    }
}

int oracle_get_vector(oracle_ws *w, int32_t id, double *out) {
    int64_t len;
    double *p = vec_ptr(w, id, &len);
    if (!p) return set_err(w, "unknown vector id %d", id);
    memcpy(out, p, sizeof(double) * (size_t)len);
    return 0;
}
int oracle_set_vector(oracle_ws *w, int32_t id, const double *in) {
    int64_t len;
    double *p = vec_ptr(w, id, &len);
    if (!p) return set_err(w, "unknown vector id %d", id);
    memcpy(p, in, sizeof(double) * (size_t)len);
    return 0;
}
int oracle_get_scalars(oracle_ws *w, hprsocp_scalars *o) {
    o->sigma = w->sigma;
    o->lambda_max_A = w->lambda_A;
    o->lambda_max_Q = w->lambda_Q;
    o->b_scale = w->b_scale;
    o->c_scale = w->c_scale;
    o->norm_b = w->norm_b;
    o->norm_c = w->norm_c;
    o->norm_b_org = w->norm_b_org;
    o->norm_c_org = w->norm_c_org;
    o->Q_is_diag = w->Q_is_diag;
    o->has_quadratic = w->has_quad;
    return 0;
}
int oracle_set_scalars(oracle_ws *w, double sigma, double la, double lq) {
    if (sigma >= 0) w->sigma = sigma;
    if (la >= 0) w->lambda_A = la;
    if (lq >= 0) w->lambda_Q = lq;
    w->lambda_set = 1;
    if (w->Q_is_diag) update_Q_factors(w);
    return 0;
}
int oracle_get_scaled_A(oracle_ws *w, double *val) {
    memcpy(val, w->A.val, sizeof(double) * (size_t)w->A.nnz);
    return 0;
}
int oracle_get_dims(oracle_ws *w, int32_t *m, int32_t *n, int64_t *nnz) {
    *m = w->m;
    *n = w->n;
    *nnz = w->A.nnz;
    return 0;
}
int oracle_spmv(oracle_ws *w, int32_t trans, const double *in, double *out) {
    csr_mv(trans ? &w->AT : &w->A, in, out);
    return 0;
}
