"""hprsocp — MI355X (gfx950) backend for HPR-SOCP's dual Halpern Peaceman–Rachford
iteration, with a host API that mirrors the reference package's:

* ``HPRSOCP_parameters``   — reference src/structs.jl:208-278 (same names, defaults)
* ``build_from_SOCP_data`` — reference src/utils/builders.jl:24-113 (same checks)
* ``optimize``             — reference src/algorithm.jl:2538-2581
* ``HPRSOCP_results``      — reference src/structs.jl:297-314

Every solve runs through the C ABI of ``libhprsocp.so`` (include/hprsocp.h);
there is no CPU fallback: if the library or a GPU is missing the call raises.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import scipy.sparse as sp

from . import _abi

__all__ = [
    "HPRSOCP_parameters", "HPRSOCP_results", "QP_info", "build_from_SOCP_data", "build_qp",
    "optimize", "Solver", "BatchSolver", "load_library", "HPRSOCPError", "build_from_cbf",
    "load_autosave", "resume_parameters",
]

_LIB = None


class HPRSOCPError(RuntimeError):
    """Raised for any non-zero status returned through the C ABI."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[hprsocp error {code}] {msg}")
        self.code = code


def load_library(path: Optional[str] = None) -> C.CDLL:
    """Load libhprsocp.so (built in-tree by __graft_entry__.build())."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or _abi.library_path("libhprsocp.so")
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP backend library not found at {p}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (make -C hpr-socp_amd/csrc)")
    lib = _abi.bind_hooks(_abi.bind(C.CDLL(p), "hprsocp_"))
    if path is None:
        _LIB = lib
    return lib


# --------------------------------------------------------------------- params
@dataclass
class HPRSOCP_parameters:
    """Solver parameters; defaults equal HPRSOCP_parameters() of the reference."""
    stoptol: float = 1e-6
    sigma: float = -1.0
    max_iter: int = 2147483647
    time_limit: float = 3600.0
    check_iter: int = 150
    warm_up: bool = False
    print_frequency: int = -1
    device_number: int = 0
    use_Ruiz_scaling: bool = True
    ruiz_iterations: int = 10
    use_bc_scaling: bool = True
    bc_scaling_norm_type: str = "l2"
    use_l2_scaling: bool = False
    use_Pock_Chambolle_scaling: bool = True
    soc_block_scaling_strategy: str = "phase_taper"
    initial_x: Optional[np.ndarray] = None
    initial_y: Optional[np.ndarray] = None
    auto_save: bool = False
    save_filename: str = "HPRSOCP_autosave.npz"
    verbose: bool = True
    use_gpu: bool = True
    # extensions of this backend
    precision: str = "fp64"          # "fp64" | "fp32_refine" | "fp32_delta" (single problems)
    refine_tol: float = 1e-6         # FP32 mode: KKT at which FP64 iterations take over
    fp32_stall_checks: int = 40      # FP32 mode: stalled checks before FP64 takes over
    use_graph: bool = True

    def to_c(self, keep: list) -> _abi.Params:
        p = _abi.Params()
        p.stoptol = float(self.stoptol)
        p.sigma = float(self.sigma)
        p.max_iter = int(self.max_iter)
        p.time_limit = float(self.time_limit)
        p.check_iter = int(self.check_iter)
        p.warm_up = int(bool(self.warm_up))
        p.print_frequency = int(self.print_frequency)
        p.device_number = int(self.device_number)
        p.use_Ruiz_scaling = int(bool(self.use_Ruiz_scaling))
        p.ruiz_iterations = int(self.ruiz_iterations)
        p.use_bc_scaling = int(bool(self.use_bc_scaling))
        key = str(self.bc_scaling_norm_type).lstrip(":")
        if key not in _abi.BC_NORMS:
            raise ValueError(f"Unsupported bc scaling norm_type: {self.bc_scaling_norm_type}. "
                             "Use :l2, :rms, or :linf.")
        p.bc_scaling_norm_type = _abi.BC_NORMS[key]
        p.use_l2_scaling = int(bool(self.use_l2_scaling))
        p.use_Pock_Chambolle_scaling = int(bool(self.use_Pock_Chambolle_scaling))
        skey = str(self.soc_block_scaling_strategy).lstrip(":")
        if skey not in _abi.SOC_STRATEGIES:
            raise ValueError(f"Unsupported soc_block_scaling_strategy={self.soc_block_scaling_strategy}")
        p.soc_block_scaling_strategy = _abi.SOC_STRATEGIES[skey]
        if self.initial_x is not None:
            a = _abi.f64(self.initial_x)
            keep.append(a)
            p.initial_x = _abi.ptr_f64(a)
        if self.initial_y is not None:
            a = _abi.f64(self.initial_y)
            keep.append(a)
            p.initial_y = _abi.ptr_f64(a)
        p.auto_save = int(bool(self.auto_save))
        fname = os.fsencode(self.save_filename)
        keep.append(fname)
        p.save_filename = fname
        p.verbose = int(bool(self.verbose))
        p.use_gpu = int(bool(self.use_gpu))
        p.precision = {"fp64": _abi.PREC_FP64, "fp32_refine": _abi.PREC_FP32_REFINE,
                       "fp32_delta": _abi.PREC_FP32_DELTA}[self.precision]
        p.refine_tol = float(self.refine_tol)
        p.fp32_stall_checks = int(self.fp32_stall_checks)
        p.use_graph = int(bool(self.use_graph))
        return p


@dataclass
class HPRSOCP_results:
    """Mirrors HPRSOCP_results (structs.jl:297-314) plus a few diagnostics."""
    iter: int = 0
    iter_4: int = 0
    iter_6: int = 0
    time: float = 0.0
    time_4: float = 0.0
    time_6: float = 0.0
    power_time: float = 0.0
    primal_obj: float = 0.0
    residuals: float = 0.0
    gap: float = 0.0
    status: str = ""
    x: np.ndarray = field(default_factory=lambda: np.zeros(0))
    y: np.ndarray = field(default_factory=lambda: np.zeros(0))
    z: np.ndarray = field(default_factory=lambda: np.zeros(0))
    w: np.ndarray = field(default_factory=lambda: np.zeros(0))
    dual_obj: float = 0.0
    err_Rp: float = 0.0
    err_Rd: float = 0.0
    setup_time: float = 0.0
    sigma: float = 0.0
    lambda_max_A: float = 0.0
    lambda_max_Q: float = 0.0
    restarts: int = 0
    iter_fp32: int = 0

    @classmethod
    def from_c(cls, r: _abi.Result, x, y, z, w) -> "HPRSOCP_results":
        return cls(iter=r.iter, iter_4=r.iter_4, iter_6=r.iter_6, time=r.time, time_4=r.time_4,
                   time_6=r.time_6, power_time=r.power_time, primal_obj=r.primal_obj,
                   residuals=r.residuals, gap=r.gap, status=_abi.STATUS_NAMES[r.status],
                   x=x, y=y, z=z, w=w, dual_obj=r.dual_obj, err_Rp=r.err_Rp, err_Rd=r.err_Rd,
                   setup_time=r.setup_time, sigma=r.sigma, lambda_max_A=r.lambda_max_A,
                   lambda_max_Q=r.lambda_max_Q, restarts=r.restarts, iter_fp32=r.iter_fp32)


# --------------------------------------------------------------------- model
@dataclass
class QP_info:
    """Canonical problem, QP_info_cpu (structs.jl:97-147).

    ``SOC_con_idx`` / ``SOC_var_idx`` keep the reference's 1-based convention
    (start of each cone, sentinel m+1 / n+1).  ``Q`` is a scipy sparse matrix or
    ``("factored", F, diag)`` for the operator form Q = F'F + diag.
    """
    Q: object
    c: np.ndarray
    A: sp.csr_matrix
    soc_rhs: np.ndarray
    soc_rhs_full: np.ndarray
    AL: np.ndarray
    AU: np.ndarray
    SOC_con_idx: np.ndarray
    number_linear_con: int
    l: np.ndarray
    u: np.ndarray
    SOC_var_idx: np.ndarray
    number_lu_x: int
    obj_constant: float = 0.0

    @property
    def AT(self):
        return self.A.T.tocsr()

    @property
    def shape(self):
        return self.A.shape

    def buffers(self) -> _abi.ProblemBuffers:
        F = Qd = None
        Q = self.Q
        if isinstance(Q, tuple) and Q and Q[0] == "factored":
            F, Qd = Q[1], Q[2]
            Q = None
        return _abi.ProblemBuffers(
            self.A, self.c, self.AL, self.AU, self.l, self.u, self.soc_rhs, self.number_linear_con,
            self.number_lu_x, np.asarray(self.SOC_con_idx) - 1, np.asarray(self.SOC_var_idx) - 1,
            Q=Q, F=F, Q_diag=Qd, obj_constant=self.obj_constant)


def build_from_SOCP_data(Q, c, A, rhs, SOC_con_idx, number_eq, number_ineq, l, u, SOC_var_idx,
                         obj_constant: float = 0.0, verbose: bool = True) -> QP_info:
    """Canonical mixed linear/SOC model; same contract and errors as the
    reference builder (builders.jl:24-113): rows = [eq | ineq (>= rhs) | SOC],
    columns = [linear | SOC blocks]; SOC index vectors are 1-based."""
    Q = sp.csr_matrix(Q, dtype=np.float64)
    A = sp.csr_matrix(A, dtype=np.float64)
    c = np.array(c, dtype=np.float64)
    rhs = np.array(rhs, dtype=np.float64)
    l = np.array(l, dtype=np.float64)
    u = np.array(u, dtype=np.float64)
    soc_con_idx = np.array(SOC_con_idx, dtype=np.int64)
    soc_var_idx = np.array(SOC_var_idx, dtype=np.int64)
    m, n = A.shape
    if len(c) != n:
        raise ValueError("Dimension mismatch: length(c) must equal number of columns of A.")
    if len(rhs) != m:
        raise ValueError("Dimension mismatch: length(rhs) must equal number of rows of A.")
    if len(l) != n:
        raise ValueError("Dimension mismatch: length(l) must equal number of variables.")
    if len(u) != n:
        raise ValueError("Dimension mismatch: length(u) must equal number of variables.")
    if number_eq < 0:
        raise ValueError("number_eq must be nonnegative.")
    if number_ineq < 0:
        raise ValueError("number_ineq must be nonnegative.")
    if number_eq + number_ineq > m:
        raise ValueError("number_eq + number_ineq cannot exceed the number of rows of A.")
    if len(soc_con_idx) == 0:
        raise ValueError("SOC_con_idx must contain at least one sentinel entry.")
    if len(soc_var_idx) == 0:
        raise ValueError("SOC_var_idx must contain at least one sentinel entry.")
    if soc_con_idx[0] != number_eq + number_ineq + 1:
        raise ValueError("SOC_con_idx must start at number_eq + number_ineq + 1.")
    if soc_con_idx[-1] != m + 1:
        raise ValueError("SOC_con_idx must end at m + 1.")
    nl = number_eq + number_ineq
    soc_rhs = rhs[nl:].copy()
    soc_rhs_full = np.zeros(m)
    soc_rhs_full[nl:] = soc_rhs
    number_lu_x = int(soc_var_idx[0] - 1)
    if soc_var_idx[-1] != n + 1:
        raise ValueError("SOC_var_idx must end at n + 1.")
    if not (0 <= number_lu_x <= n):
        raise ValueError("Invalid SOC_var_idx start.")
    AL = np.full(m, -np.inf)
    AU = np.full(m, np.inf)
    AL[:number_eq] = rhs[:number_eq]
    AU[:number_eq] = rhs[:number_eq]
    AL[number_eq:nl] = rhs[number_eq:nl]
    return QP_info(Q=Q, c=c, A=A, soc_rhs=soc_rhs, soc_rhs_full=soc_rhs_full, AL=AL, AU=AU,
                   SOC_con_idx=soc_con_idx, number_linear_con=nl, l=l, u=u,
                   SOC_var_idx=soc_var_idx, number_lu_x=number_lu_x,
                   obj_constant=float(obj_constant))


def build_qp(A, c, AL, AU, l, u, soc_rhs, number_linear_con, number_lu_x, soc_con_idx0,
             soc_var_idx0, Q=None, obj_constant: float = 0.0) -> QP_info:
    """QP_info with two-sided linear rows [AL, AU] (the struct form the
    reference's builder fills).  Cone offsets here are 0-based with sentinel."""
    A = sp.csr_matrix(A, dtype=np.float64)
    m, n = A.shape
    soc_rhs = np.asarray(soc_rhs, dtype=np.float64)
    full = np.zeros(m)
    full[number_linear_con:] = soc_rhs
    if Q is None:
        Q = sp.csr_matrix((n, n))
    return QP_info(Q=Q, c=np.asarray(c, float), A=A, soc_rhs=soc_rhs, soc_rhs_full=full,
                   AL=np.asarray(AL, float), AU=np.asarray(AU, float),
                   SOC_con_idx=np.asarray(soc_con_idx0, np.int64) + 1,
                   number_linear_con=int(number_linear_con), l=np.asarray(l, float),
                   u=np.asarray(u, float), SOC_var_idx=np.asarray(soc_var_idx0, np.int64) + 1,
                   number_lu_x=int(number_lu_x), obj_constant=float(obj_constant))


# --------------------------------------------------------------------- solver
class Solver:
    """One backend handle (prepared, scaled problem resident on the GPU).

    Exposes the reference's per-step seams (estimate_eigenvalues, the update,
    compute_residuals!, compute_M_norm!) for tests and tooling.  ``lib`` and
    ``prefix`` default to the HIP backend.
    """

    def __init__(self, model: QP_info, params: Optional[HPRSOCP_parameters] = None,
                 lib: Optional[C.CDLL] = None, prefix: str = "hprsocp_"):
        self.lib = lib if lib is not None else load_library()
        self.prefix = prefix
        self.model = model
        self.params = params or HPRSOCP_parameters(verbose=False)
        self._keep = []
        self._buf = model.buffers()
        cp = self.params.to_c(self._keep)
        h = C.c_void_p()
        rc = self._fn("create")(C.byref(self._buf.problem), C.byref(cp), C.byref(h))
        if rc != 0:
            msg = self._fn("last_error")(None)
            raise HPRSOCPError(rc, msg.decode() if msg else "create failed")
        self.h = h
        self.m, self.n = model.A.shape

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc):
        if rc != 0:
            msg = self._fn("last_error")(self.h)
            raise HPRSOCPError(rc, msg.decode() if msg else "error")

    def close(self):
        if getattr(self, "h", None):
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # seams
    def estimate_lambda(self):
        self._check(self._fn("estimate_lambda")(self.h))
        return self.scalars()

    def set_initial(self, x0=None, y0=None):
        xa, ya = _abi.f64(x0), _abi.f64(y0)
        self._check(self._fn("set_initial")(self.h, _abi.ptr_f64(xa), _abi.ptr_f64(ya)))

    def update(self, inner: int, to_check: bool = True):
        self._check(self._fn("update")(self.h, int(inner), int(bool(to_check))))

    def residuals(self, it: int) -> dict:
        r = _abi.Residuals()
        self._check(self._fn("compute_residuals")(self.h, int(it), C.byref(r)))
        return {k: getattr(r, k) for k, _ in _abi.Residuals._fields_}

    def mnorm(self) -> float:
        v = C.c_double()
        self._check(self._fn("mnorm")(self.h, C.byref(v)))
        return v.value

    def get(self, name: str) -> np.ndarray:
        ln = self.m if name in _abi.VEC_IS_M else self.n
        out = np.zeros(ln)
        self._check(self._fn("get_vector")(self.h, _abi.VEC[name], _abi.ptr_f64(out)))
        return out

    def set(self, name: str, v):
        ln = self.m if name in _abi.VEC_IS_M else self.n
        a = _abi.f64(v)
        if a.shape != (ln,):
            raise ValueError(f"{name}: expected length {ln}")
        self._check(self._fn("set_vector")(self.h, _abi.VEC[name], _abi.ptr_f64(a)))

    def scalars(self) -> dict:
        s = _abi.Scalars()
        self._check(self._fn("get_scalars")(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _abi.Scalars._fields_}

    def set_scalars(self, sigma=-1.0, lambda_A=-1.0, lambda_Q=-1.0):
        self._check(self._fn("set_scalars")(self.h, float(sigma), float(lambda_A), float(lambda_Q)))

    def scaled_A(self) -> np.ndarray:
        out = np.zeros(self._buf.problem.A_nnz)
        self._check(self._fn("get_scaled_A")(self.h, _abi.ptr_f64(out)))
        return out

    def spmv(self, v, trans=False) -> np.ndarray:
        a = _abi.f64(v)
        out = np.zeros(self.n if trans else self.m)
        self._check(self._fn("spmv")(self.h, int(trans), _abi.ptr_f64(a), _abi.ptr_f64(out)))
        return out

    def _results(self, r) -> HPRSOCP_results:
        x, y, z, w = np.zeros(self.n), np.zeros(self.m), np.zeros(self.n), np.zeros(self.n)
        self._check(self._fn("get_solution")(self.h, _abi.ptr_f64(x), _abi.ptr_f64(y),
                                              _abi.ptr_f64(z), _abi.ptr_f64(w)))
        return HPRSOCP_results.from_c(r, x, y, z, w)

    def solve(self) -> HPRSOCP_results:
        r = _abi.Result()
        self._check(self._fn("solve")(self.h, C.byref(r)))
        return self._results(r)

    def iterate(self, k: int, reset: bool = False) -> HPRSOCP_results:
        """Runs the next ``k`` bodies of the solve loop (hprsocp_iterate);
        ``status`` is "CONTINUE" until the loop terminates."""
        r = _abi.Result()
        self._check(self._fn("iterate")(self.h, int(k), _abi.ITERATE_RESET if reset else 0, C.byref(r)))
        return self._results(r)

    def plan_info(self) -> dict:
        """The autotuner's launch-plan choice for each operator (hprsocp_plan_info)."""
        return _plan_info(self.lib.hprsocp_plan_info, self.h, self._check)

    # restart / sigma seams (check_restart, update_sigma!, do_restart!)
    def init_restart(self) -> _abi.RestartInfo:
        r = _abi.RestartInfo()
        self._fn("init_restart")(C.byref(r))
        return r

    def check_restart(self, r: _abi.RestartInfo, it: int, check_iter: int, sigma: float, kkt: float):
        self._fn("check_restart")(C.byref(r), int(it), int(check_iter), float(sigma), float(kkt))

    def update_sigma(self, r: _abi.RestartInfo, res: dict):
        rr = _abi.Residuals(**{k: float(res.get(k, 0.0)) for k, _ in _abi.Residuals._fields_})
        self._check(self._fn("update_sigma")(self.h, C.byref(r), C.byref(rr)))

    def restart(self, r: _abi.RestartInfo):
        self._check(self._fn("restart")(self.h, C.byref(r)))

    def probe_copy(self, nbytes: int) -> float:
        """Plain device copy of `nbytes` (counter calibration); returns ms."""
        ms = C.c_double()
        fn = self.lib.hprsocp_probe_copy
        fn.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]
        fn.restype = C.c_int
        self._check(fn(self.h, int(nbytes), C.byref(ms)))
        return ms.value

    def time_iterations(self, iters: int, per_kernel: bool = False):
        ms = C.c_double()
        kern = (C.c_double * 6)()
        self._check(self.lib.hprsocp_time_iterations(self.h, int(iters), C.byref(ms),
                                                      kern if per_kernel else None))
        return ms.value, list(kern) if per_kernel else None

    def time_spmv(self, trans: bool = False, reps: int = 20, fp32: bool = False) -> float:
        """Median ms of the iteration's A x (or A^T y) SpMV on resident operands
        (``fp32``: the FP32 operator of the FP32 precisions)."""
        ms = C.c_double()
        self._check(self.lib.hprsocp_time_spmv(self.h, int(bool(trans)) | (2 if fp32 else 0), int(reps),
                                                C.byref(ms)))
        return ms.value

    def probe_gather(self, nnz: int, width: int, slices: int = 8, sorted_chunks: bool = False,
                     fp32: bool = False) -> float:
        """Median ms of the synthetic gather-ceiling kernel (see include/hprsocp.h);
        ``fp32`` probes float entries and gathers (the FP32 operators)."""
        ms = C.c_double()
        fn = self.lib.hprsocp_probe_gather32 if fp32 else self.lib.hprsocp_probe_gather
        self._check(fn(self.h, int(nnz), int(width), int(slices), int(bool(sorted_chunks)), C.byref(ms)))
        return ms.value

    def probe_ctl(self, xs) -> "tuple[np.ndarray, np.ndarray]":
        """The control arithmetic of batch solves evaluated at xs on the device
        and on the host (see include/hprsocp.h): two (len(xs), 5) arrays."""
        x = np.ascontiguousarray(xs, dtype=np.float64)
        dev = np.zeros((x.size, 5))
        host = np.zeros((x.size, 5))
        ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        self._check(self.lib.hprsocp_probe_ctl(self.h, ptr(x), x.size, ptr(dev), ptr(host)))
        return dev, host


def _plan_info(fn, h, check) -> dict:
    import json
    n = C.c_int64()
    check(fn(h, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(fn(h, buf, n.value + 1, C.byref(n)))
    return json.loads(buf.value.decode())


def project_cones(inp: int, out: int, n: int, cone_idx: int, ncones: int, stream: int = 0,
                  lib: Optional[C.CDLL] = None, noalloc: bool = False) -> None:
    """hprsocp_project_cones on device pointers (ints, e.g. torch tensors'
    data_ptr()): rows [cone_idx[k], cone_idx[k+1]) of ``inp`` projected onto the
    second-order cone into ``out``; waits for ``stream``.  ``noalloc`` calls
    hprsocp_project_cones_noalloc instead (no allocation, no wait)."""
    lib = lib or load_library()
    fn = lib.hprsocp_project_cones_noalloc if noalloc else lib.hprsocp_project_cones
    rc = fn(C.c_void_p(inp), C.c_void_p(out), int(n), C.c_void_p(cone_idx), int(ncones), C.c_void_p(stream or None))
    if rc != 0:
        raise HPRSOCPError(rc, lib.hprsocp_project_cones_last_error().decode())


def optimize(model: QP_info, params: Optional[HPRSOCP_parameters] = None) -> HPRSOCP_results:
    """optimize(model, params) of the reference (algorithm.jl:2538): optional
    200-iteration warm-up, then setup + solve on the GPU; returns unscaled x, y, z, w."""
    lib = load_library()
    params = params or HPRSOCP_parameters()
    keep = []
    buf = model.buffers()
    cp = params.to_c(keep)
    m, n = model.A.shape
    r = _abi.Result()
    x, y, z, w = np.zeros(n), np.zeros(m), np.zeros(n), np.zeros(n)
    rc = lib.hprsocp_optimize(C.byref(buf.problem), C.byref(cp), C.byref(r), _abi.ptr_f64(x),
                              _abi.ptr_f64(y), _abi.ptr_f64(z), _abi.ptr_f64(w))
    if rc != 0:
        msg = lib.hprsocp_last_error(None)
        raise HPRSOCPError(rc, msg.decode() if msg else "optimize failed")
    return HPRSOCP_results.from_c(r, x, y, z, w)


def load_autosave(filename: str) -> dict:
    """Reads an auto-save checkpoint written by a solve with ``auto_save=True``:
    the current / best / parameters groups of the reference's
    save_state_to_hdf5! (src/algorithm.jl:306-423), keyed by the same dataset
    paths ("best/x_org", "current/iteration", ...).  Vectors come back as
    float64 arrays, scalars as Python float / int / bool / str.  The file is
    an .npz archive, read with allow_pickle=False."""
    out = {}
    with np.load(filename, allow_pickle=False) as d:
        for k in d.files:
            a = d[k]
            out[k] = a.item() if a.ndim == 0 else a
    return out


def resume_parameters(filename: str, params: Optional[HPRSOCP_parameters] = None, which: str = "best",
                      use_sigma: bool = True) -> HPRSOCP_parameters:
    """Warm-start parameters from an auto-save checkpoint: initial_x and
    initial_y from the checkpoint's ``which`` group ("best" or "current"),
    sigma from the same group when ``use_sigma``, every other field from
    ``params`` (the defaults when None).  optimize() applies the initial points
    as process_initial_points! does (reference src/algorithm.jl:2612-2677)."""
    if which not in ("best", "current"):
        raise ValueError("which must be 'best' or 'current'")
    d = load_autosave(filename)
    p = dataclasses.replace(params) if params is not None else HPRSOCP_parameters()
    p.initial_x = np.array(d[f"{which}/x_org"], dtype=np.float64)
    y = d.get(f"{which}/y_org")
    p.initial_y = None if y is None else np.array(y, dtype=np.float64)
    if use_sigma:
        p.sigma = float(d[f"{which}/sigma"])
    return p


class BatchSolver:
    """Many independent problems solved together on one GPU.

    The instances are packed into one block-diagonal layout and iterated in
    lockstep by the same kernels; each keeps its own scaling, sigma, restarts
    and termination (hprsocp_batch_* in include/hprsocp.h).
    """

    def __init__(self, models, params: Optional[HPRSOCP_parameters] = None, lib: Optional[C.CDLL] = None):
        self.lib = _abi.bind_batch(lib if lib is not None else load_library())
        self.models = list(models)
        self.params = params or HPRSOCP_parameters(verbose=False)
        self._keep = []
        self._bufs = [m.buffers() for m in self.models]
        arr = (C.POINTER(_abi.Problem) * len(self._bufs))(*[C.pointer(b.problem) for b in self._bufs])
        self._keep.append(arr)
        cp = self.params.to_c(self._keep)
        h = C.c_void_p()
        rc = self.lib.hprsocp_batch_create(C.cast(arr, C.c_void_p), len(self._bufs), C.byref(cp), C.byref(h))
        if rc != 0:
            msg = self.lib.hprsocp_batch_last_error(None)
            raise HPRSOCPError(rc, msg.decode() if msg else "batch create failed")
        self.h = h
        self.dims = [m.A.shape for m in self.models]

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.hprsocp_batch_last_error(self.h)
            raise HPRSOCPError(rc, msg.decode() if msg else "error")

    def close(self):
        if getattr(self, "h", None):
            self.lib.hprsocp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return len(self.models)

    def estimate_lambda(self):
        self._check(self.lib.hprsocp_batch_estimate_lambda(self.h))

    def set_initial(self, xs=None, ys=None):
        """xs / ys: per-instance initial points (None entries = cold start)."""
        k = len(self.models)
        keep = []

        def table(vs):
            if vs is None:
                return None
            t = (C.c_void_p * k)()
            for i, v in enumerate(vs):
                if v is not None:
                    a = _abi.f64(v)
                    keep.append(a)
                    t[i] = a.ctypes.data
            keep.append(t)
            return C.cast(t, C.c_void_p)

        self._check(self.lib.hprsocp_batch_set_initial(self.h, table(xs), table(ys)))

    def solve(self):
        """Solves all instances; returns one HPRSOCP_results per instance."""
        k = len(self.models)
        res = (_abi.Result * k)()
        self._check(self.lib.hprsocp_batch_solve(self.h, res))
        out = []
        for i, (m, n) in enumerate(self.dims):
            x, z, w = np.zeros(n), np.zeros(n), np.zeros(n)
            y = np.zeros(m)
            self._check(self.lib.hprsocp_batch_get_solution(
                self.h, i, _abi.ptr_f64(x), _abi.ptr_f64(y) if m else None, _abi.ptr_f64(z), _abi.ptr_f64(w)))
            out.append(HPRSOCP_results.from_c(res[i], x, y, z, w))
        return out

    def plan_info(self) -> dict:
        """The autotuner's launch-plan choice for the packed operators."""
        return _plan_info(self.lib.hprsocp_batch_plan_info, self.h, self._check)

    def time_iterations(self, iters: int, per_kernel: bool = False):
        """(ms per iteration, per-launch-group ms or None) of `iters` lockstep iterations."""
        ms = C.c_double()
        kern = np.zeros(6) if per_kernel else None
        self._check(self.lib.hprsocp_batch_time_iterations(self.h, int(iters), C.byref(ms),
                                                           _abi.ptr_f64(kern)))
        return ms.value, (list(kern) if per_kernel else None)


def build_from_cbf(filename: str, verbose: bool = False) -> QP_info:
    """Model from a .cbf file, build_from_cbf (reference src/utils/builders.jl:114-130)
    through the library's reader (hprsocp_cbf_read; host code only)."""
    lib = load_library()
    lib.hprsocp_cbf_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    lib.hprsocp_cbf_read.restype = C.c_int
    lib.hprsocp_cbf_last_error.restype = C.c_char_p
    lib.hprsocp_cbf_problem.argtypes = [C.c_void_p, C.POINTER(_abi.Problem), C.POINTER(C.c_int32)]
    lib.hprsocp_cbf_destroy.argtypes = [C.c_void_p]
    h = C.c_void_p()
    rc = lib.hprsocp_cbf_read(str(filename).encode(), C.byref(h))
    if rc != 0:
        raise HPRSOCPError(rc, lib.hprsocp_cbf_last_error().decode())
    try:
        p = _abi.Problem()
        neq = C.c_int32()
        lib.hprsocp_cbf_problem(h, C.byref(p), C.byref(neq))
        m, n, nnz = p.m, p.n, p.A_nnz

        def arr(ptr, k, dt):
            return np.ctypeslib.as_array(ptr, shape=(k,)).astype(dt, copy=True) if k > 0 else np.zeros(0, dt)

        rp = arr(p.A_rowptr, m + 1, np.int64)
        ci = arr(p.A_colidx, nnz, np.int64)
        v = arr(p.A_val, nnz, np.float64)
        A = sp.csr_matrix((v, ci, rp), shape=(m, n))
        m1 = p.number_linear_con
        model = build_qp(A, arr(p.c, n, np.float64), arr(p.AL, m, np.float64), arr(p.AU, m, np.float64),
                         arr(p.l, n, np.float64), arr(p.u, n, np.float64), arr(p.soc_rhs, m - m1, np.float64),
                         m1, p.number_lu_x, arr(p.soc_con_idx, p.n_soc_con + 1, np.int64),
                         arr(p.soc_var_idx, p.n_soc_var + 1, np.int64), obj_constant=p.obj_constant)
        model.number_eq = int(neq.value)
        model.number_ineq = int(m1 - neq.value)
        if verbose:
            print(f"READ CBF {filename}: m={m} n={n} nnz={nnz}")
        return model
    finally:
        lib.hprsocp_cbf_destroy(h)
