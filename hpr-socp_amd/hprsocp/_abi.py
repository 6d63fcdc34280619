"""ctypes mirror of include/hprsocp.h.

The same structures serve both the HIP backend (libhprsocp.so, prefix
``hprsocp_``) and, in tests only, the CPU oracle (liboracle.so, prefix
``oracle_``), whose entry points have identical signatures.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ABI_VERSION = 1

# hprsocp_errc
OK, ERR_INVALID, ERR_DEVICE, ERR_ALLOC, ERR_UNSUPPORTED, ERR_STATE = range(6)
# hprsocp_status
STATUS_NAMES = {0: "OPTIMAL", 1: "MAX_ITER", 2: "TIME_LIMIT", 3: "CONTINUE"}
# hprsocp_q_kind
Q_CSR, Q_FACTORED = 0, 1
# hprsocp_precision
PREC_FP64, PREC_FP32_REFINE, PREC_FP32_DELTA = 0, 1, 2
BC_NORMS = {"l2": 0, "rms": 1, "linf": 2, "l_inf": 2, "inf": 2}
SOC_STRATEGIES = {
    "max": 1, "rms": 2, "geom": 3, "adaptive": 4, "hybrid": 5, "ruiz_geom": 6,
    "role_split": 7, "pock_safe_role_split": 8, "size_split_role_split": 9,
    "pre_pock_size_split": 10, "pre_pock_huge_var_taper": 11,
    "pre_pock_huge_var_taper_pock_geom": 12, "pre_pock_band_split": 13,
    "pre_pock_constraint_rms": 14, "pre_pock_constraint_max": 15,
    "pre_pock_constraint_taper": 16, "small_cone_constraint_max": 17,
    "count_aware_small_cone_bias": 18, "phase_blend": 19, "phase_taper": 20,
}

# hprsocp_vector_id
VEC = {name: i for i, name in enumerate([
    "x", "x_bar", "x_hat", "z_bar", "dx", "last_x",
    "y", "y_bar", "dy", "last_y", "s",
    "w", "w_bar", "last_w", "dw", "ATy", "ATy_bar", "last_ATy",
    "c", "l", "u", "AL", "AU", "soc_rhs_full", "row_norm", "col_norm", "diag_Q",
    "Rp", "Ax", "Rd",
])}
VEC_IS_M = {"y", "y_bar", "dy", "last_y", "s", "AL", "AU", "soc_rhs_full", "row_norm", "Rp", "Ax"}

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class Problem(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("n", C.c_int32),
        ("A_nnz", C.c_int64), ("A_rowptr", _i32p), ("A_colidx", _i32p), ("A_val", _f64p),
        ("Q_kind", C.c_int32),
        ("Q_nnz", C.c_int64), ("Q_rowptr", _i32p), ("Q_colidx", _i32p), ("Q_val", _f64p),
        ("F_rows", C.c_int32),
        ("F_nnz", C.c_int64), ("F_rowptr", _i32p), ("F_colidx", _i32p), ("F_val", _f64p),
        ("Q_diag", _f64p),
        ("c", _f64p), ("AL", _f64p), ("AU", _f64p), ("l", _f64p), ("u", _f64p),
        ("soc_rhs", _f64p),
        ("number_linear_con", C.c_int32), ("number_lu_x", C.c_int32),
        ("n_soc_con", C.c_int32), ("soc_con_idx", _i32p),
        ("n_soc_var", C.c_int32), ("soc_var_idx", _i32p),
        ("obj_constant", C.c_double),
    ]


class Params(C.Structure):
    _fields_ = [
        ("stoptol", C.c_double), ("sigma", C.c_double), ("max_iter", C.c_int64),
        ("time_limit", C.c_double), ("check_iter", C.c_int32), ("warm_up", C.c_int32),
        ("print_frequency", C.c_int32), ("device_number", C.c_int32),
        ("use_Ruiz_scaling", C.c_int32), ("ruiz_iterations", C.c_int32),
        ("use_bc_scaling", C.c_int32), ("bc_scaling_norm_type", C.c_int32),
        ("use_l2_scaling", C.c_int32), ("use_Pock_Chambolle_scaling", C.c_int32),
        ("soc_block_scaling_strategy", C.c_int32),
        ("initial_x", _f64p), ("initial_y", _f64p),
        ("auto_save", C.c_int32), ("save_filename", C.c_char_p),
        ("verbose", C.c_int32), ("use_gpu", C.c_int32),
        ("precision", C.c_int32), ("use_graph", C.c_int32), ("refine_tol", C.c_double),
        ("fp32_stall_checks", C.c_int32), ("reserved", C.c_int32 * 7),
    ]


class Result(C.Structure):
    _fields_ = [
        ("iter", C.c_int64), ("iter_4", C.c_int64), ("iter_6", C.c_int64),
        ("time", C.c_double), ("time_4", C.c_double), ("time_6", C.c_double),
        ("power_time", C.c_double), ("setup_time", C.c_double),
        ("primal_obj", C.c_double), ("dual_obj", C.c_double),
        ("residuals", C.c_double), ("gap", C.c_double),
        ("err_Rp", C.c_double), ("err_Rp_linear", C.c_double), ("err_Rp_soc", C.c_double),
        ("err_Rd", C.c_double),
        ("sigma", C.c_double), ("lambda_max_A", C.c_double), ("lambda_max_Q", C.c_double),
        ("status", C.c_int32), ("restarts", C.c_int32), ("iter_fp32", C.c_int64),
    ]


class Residuals(C.Structure):
    _fields_ = [
        ("err_Rp", C.c_double), ("err_Rp_linear", C.c_double), ("err_Rp_soc", C.c_double),
        ("err_Rd", C.c_double), ("kkt", C.c_double), ("primal_obj", C.c_double),
        ("rel_gap", C.c_double), ("dual_obj", C.c_double),
    ]


class Scalars(C.Structure):
    _fields_ = [
        ("sigma", C.c_double), ("lambda_max_A", C.c_double), ("lambda_max_Q", C.c_double),
        ("b_scale", C.c_double), ("c_scale", C.c_double), ("norm_b", C.c_double),
        ("norm_c", C.c_double), ("norm_b_org", C.c_double), ("norm_c_org", C.c_double),
        ("Q_is_diag", C.c_int32), ("has_quadratic", C.c_int32),
    ]


class RestartInfo(C.Structure):
    """hprsocp_restart_info: field for field the reference's HPRSOCP_restart."""
    _fields_ = [
        ("restart_flag", C.c_int64), ("first_restart", C.c_uint8),
        ("last_gap", C.c_double), ("current_gap", C.c_double), ("save_gap", C.c_double),
        ("inner", C.c_int64), ("step", C.c_int64), ("sufficient", C.c_int64),
        ("necessary", C.c_int64), ("long_restarts", C.c_int64), ("ratio", C.c_int64),
        ("times", C.c_int64),
        ("weighted_norm", C.c_double), ("best_gap", C.c_double), ("best_kkt", C.c_double),
        ("best_sigma", C.c_double), ("best_iter", C.c_int64),
        ("sigma_correction_active", C.c_uint8),
        ("sigma_correction_hold", C.c_int64), ("sigma_correction_dir", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> (restype, argtypes); "H" = opaque handle pointer
_SIGS = {
    "default_params": (None, [C.POINTER(Params)]),
    "create": (C.c_int, [C.POINTER(Problem), C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "destroy": (None, [C.c_void_p]),
    "last_error": (C.c_char_p, [C.c_void_p]),
    "estimate_lambda": (C.c_int, [C.c_void_p]),
    "set_initial": (C.c_int, [C.c_void_p, _f64p, _f64p]),
    "update": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32]),
    "compute_residuals": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Residuals)]),
    "mnorm": (C.c_int, [C.c_void_p, _f64p]),
    "solve": (C.c_int, [C.c_void_p, C.POINTER(Result)]),
    "get_solution": (C.c_int, [C.c_void_p, _f64p, _f64p, _f64p, _f64p]),
    "optimize": (C.c_int, [C.POINTER(Problem), C.POINTER(Params), C.POINTER(Result),
                           _f64p, _f64p, _f64p, _f64p]),
    "get_vector": (C.c_int, [C.c_void_p, C.c_int32, _f64p]),
    "set_vector": (C.c_int, [C.c_void_p, C.c_int32, _f64p]),
    "get_scalars": (C.c_int, [C.c_void_p, C.POINTER(Scalars)]),
    "set_scalars": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "get_scaled_A": (C.c_int, [C.c_void_p, _f64p]),
    "get_dims": (C.c_int, [C.c_void_p, _i32p, _i32p, C.POINTER(C.c_int64)]),
    "spmv": (C.c_int, [C.c_void_p, C.c_int32, _f64p, _f64p]),
    "init_restart": (None, [C.POINTER(RestartInfo)]),
    "check_restart": (None, [C.POINTER(RestartInfo), C.c_int64, C.c_int64, C.c_double, C.c_double]),
    "update_sigma": (C.c_int, [C.c_void_p, C.POINTER(RestartInfo), C.POINTER(Residuals)]),
    "restart": (C.c_int, [C.c_void_p, C.POINTER(RestartInfo)]),
}

# Entry points of the HIP backend only (the oracle runs its loop in one piece).
_LOOP_SIGS = {
    "iterate": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(Result)]),
    "plan_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, C.POINTER(C.c_int64)]),
}
ITERATE_RESET = 1

# Entry points every library exporting this ABI must provide.
ABI_FUNCTIONS = sorted(_SIGS)

# Batch entry points (HIP backend only).
_BATCH_SIGS = {
    "batch_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "batch_destroy": (None, [C.c_void_p]),
    "batch_last_error": (C.c_char_p, [C.c_void_p]),
    "batch_size": (C.c_int32, [C.c_void_p]),
    "batch_dims": (C.c_int, [C.c_void_p, C.c_int32, _i32p, _i32p]),
    "batch_estimate_lambda": (C.c_int, [C.c_void_p]),
    "batch_set_initial": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "batch_solve": (C.c_int, [C.c_void_p, C.POINTER(Result)]),
    "batch_get_solution": (C.c_int, [C.c_void_p, C.c_int32, _f64p, _f64p, _f64p, _f64p]),
    "batch_time_iterations": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double), _f64p]),
    "batch_plan_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, C.POINTER(C.c_int64)]),
}


# Stand-alone device-pointer kernels (HIP backend only).
_KERNEL_SIGS = {
    "project_cones": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_void_p]),
    "project_cones_last_error": (C.c_char_p, []),
    "project_cones_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_void_p]),
}


# Measurement hooks (HIP backend only).
_HOOK_SIGS = {
    "time_iterations": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double), _f64p]),
    "probe_copy": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]),
    "time_spmv": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "probe_gather": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "probe_gather32": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "probe_ctl": (C.c_int, [C.c_void_p, _f64p, C.c_int32, _f64p, _f64p]),
}


def bind_hooks(lib: C.CDLL, prefix: str = "hprsocp_"):
    for name, (res, args) in list(_HOOK_SIGS.items()) + list(_LOOP_SIGS.items()) + list(_KERNEL_SIGS.items()):
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
    return lib


def bind_batch(lib: C.CDLL, prefix: str = "hprsocp_"):
    for name, (res, args) in _BATCH_SIGS.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
    return lib


def bind(lib: C.CDLL, prefix: str):
    """Attach argtypes/restype for every ABI function under ``prefix``."""
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
    return lib


def f64(a):
    """Contiguous float64 view kept alive by the caller; None -> NULL."""
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.int32)


def ptr_f64(a):
    return None if a is None else a.ctypes.data_as(_f64p)


def ptr_i32(a):
    return None if a is None else a.ctypes.data_as(_i32p)


class ProblemBuffers:
    """Owns the numpy arrays a :class:`Problem` points into.

    ``csr`` arguments are scipy.sparse matrices (converted to canonical CSR
    with int32 indices).  Cone offsets are 0-based with a trailing sentinel.
    """

    def __init__(self, A, c, AL, AU, l, u, soc_rhs, number_linear_con, number_lu_x,
                 soc_con_idx, soc_var_idx, Q=None, F=None, Q_diag=None, obj_constant=0.0):
        import scipy.sparse as sp
        A = sp.csr_matrix(A)
        A.sum_duplicates()
        m, n = A.shape
        self.keep = []
        k = self.keep.append
        self.A_rowptr, self.A_colidx, self.A_val = i32(A.indptr), i32(A.indices), f64(A.data)
        k(self.A_rowptr); k(self.A_colidx); k(self.A_val)
        p = Problem()
        p.m, p.n = m, n
        p.A_nnz = int(A.nnz)
        p.A_rowptr, p.A_colidx, p.A_val = (ptr_i32(self.A_rowptr), ptr_i32(self.A_colidx),
                                           ptr_f64(self.A_val))
        if F is not None:
            F = sp.csr_matrix(F)
            F.sum_duplicates()
            arrs = (i32(F.indptr), i32(F.indices), f64(F.data))
            for a in arrs:
                k(a)
            p.Q_kind = Q_FACTORED
            p.F_rows = F.shape[0]
            p.F_nnz = int(F.nnz)
            p.F_rowptr, p.F_colidx, p.F_val = ptr_i32(arrs[0]), ptr_i32(arrs[1]), ptr_f64(arrs[2])
            if Q_diag is not None:
                qd = f64(Q_diag)
                k(qd)
                p.Q_diag = ptr_f64(qd)
        else:
            p.Q_kind = Q_CSR
            if Q is not None:
                Q = sp.csr_matrix(Q)
                Q.sum_duplicates()
                Q.eliminate_zeros()
            if Q is not None and Q.nnz > 0:
                arrs = (i32(Q.indptr), i32(Q.indices), f64(Q.data))
                for a in arrs:
                    k(a)
                p.Q_nnz = int(Q.nnz)
                p.Q_rowptr, p.Q_colidx, p.Q_val = (ptr_i32(arrs[0]), ptr_i32(arrs[1]),
                                                   ptr_f64(arrs[2]))
        vecs = {}
        for name, v, ln in (("c", c, n), ("AL", AL, m), ("AU", AU, m), ("l", l, n), ("u", u, n),
                            ("soc_rhs", soc_rhs, m - number_linear_con)):
            a = f64(np.zeros(0) if v is None else v)
            if a.shape != (ln,):
                raise ValueError(f"{name} must have length {ln}, got {a.shape}")
            k(a)
            vecs[name] = a
            setattr(p, name, ptr_f64(a))
        p.number_linear_con = int(number_linear_con)
        p.number_lu_x = int(number_lu_x)
        ci = i32(soc_con_idx if soc_con_idx is not None else [m])
        vi = i32(soc_var_idx if soc_var_idx is not None else [n])
        k(ci); k(vi)
        p.n_soc_con = len(ci) - 1
        p.soc_con_idx = ptr_i32(ci)
        p.n_soc_var = len(vi) - 1
        p.soc_var_idx = ptr_i32(vi)
        p.obj_constant = float(obj_constant)
        self.problem = p
        self.m, self.n = m, n
        self.vecs = vecs


def library_path(name: str) -> str:
    here = os.path.dirname(os.path.abspath(__file__))
    return os.path.join(here, name)
