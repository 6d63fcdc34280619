/*
 * oracle.h — FP64 CPU restatement of HPR-SOCP's solver, used ONLY as the
 * checker in tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * It is test infrastructure, never part of the product path.
 *
 * The entry points mirror include/hprsocp.h one to one (same structs, same
 * argument meaning) with the prefix oracle_ instead of hprsocp_, so a parity
 * test can drive both through identical calls.
 */
#ifndef HPRSOCP_ORACLE_H
#define HPRSOCP_ORACLE_H

#include "../include/hprsocp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ws oracle_ws;

void oracle_default_params(hprsocp_params *p);
int oracle_create(const hprsocp_problem *prob, const hprsocp_params *params, oracle_ws **out);
void oracle_destroy(oracle_ws *w);
const char *oracle_last_error(const oracle_ws *w);
int oracle_estimate_lambda(oracle_ws *w);
int oracle_set_initial(oracle_ws *w, const double *x0, const double *y0);
int oracle_update(oracle_ws *w, int64_t inner, int32_t to_check);
int oracle_compute_residuals(oracle_ws *w, int64_t iter, hprsocp_residuals *out);
int oracle_mnorm(oracle_ws *w, double *out);
int oracle_solve(oracle_ws *w, hprsocp_result *res);
int oracle_get_solution(oracle_ws *w, double *x, double *y, double *z, double *wv);
int oracle_optimize(const hprsocp_problem *prob, const hprsocp_params *params,
                    hprsocp_result *res, double *x, double *y, double *z, double *wv);
int oracle_get_vector(oracle_ws *w, int32_t id, double *out);
int oracle_set_vector(oracle_ws *w, int32_t id, const double *in);
int oracle_get_scalars(oracle_ws *w, hprsocp_scalars *out);
int oracle_set_scalars(oracle_ws *w, double sigma, double lambda_A, double lambda_Q);
int oracle_get_scaled_A(oracle_ws *w, double *val);
int oracle_get_dims(oracle_ws *w, int32_t *m, int32_t *n, int64_t *nnz);
int oracle_spmv(oracle_ws *w, int32_t trans, const double *in, double *out);

/* Seams of the solve loop (initialize_restart, check_restart, update_sigma!,
 * do_restart!; algorithm.jl:1886-1908, 1016-1099, 729-990, 1130-1166). */
void oracle_init_restart(hprsocp_restart_info *r);
void oracle_check_restart(hprsocp_restart_info *r, int64_t iter, int64_t check_iter, double sigma,
                          double kkt);
int oracle_update_sigma(oracle_ws *w, hprsocp_restart_info *r, const hprsocp_residuals *res);
int oracle_restart(oracle_ws *w, hprsocp_restart_info *r);

/* Pieces exposed for known-answer tests of single functions. */
/* Pi_Q(t, xi) in place on v[0..len) (kernels.jl:7585-7662 restated). */
void oracle_soc_project(double *v, int32_t len);
/* _soc_block_scale_value_scalar (scaling.jl:253-477). */
double oracle_soc_block_scale_value(int32_t block_len, double block_max, double rms,
                                    int32_t strategy, int32_t phase, int32_t location,
                                    int32_t cone_count);
/* golden (kernels.jl:6272-6305). */
double oracle_golden(double a, double b, double c, double d, double lo, double hi, double tol);
/* Deterministic power-iteration start vector (standard normals + 1e-8). */
void oracle_start_vector(int64_t len, double *out);
/* sizeof of the ABI structs as this C compiler lays them out
 * (0 problem, 1 params, 2 result, 3 residuals, 4 scalars, 5 restart info);
 * -1 if unknown. */
int64_t oracle_abi_sizeof(int32_t which);
/* Thread count used by the parallel loops (1 = sequential). */
void oracle_set_threads(int32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
