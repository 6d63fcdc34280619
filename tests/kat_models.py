"""Models of the reference's known-answer tests (tests/golden/reference_kats.json)."""
import json
import os

import numpy as np
import scipy.sparse as sp

import hprsocp
from hprsocp import build_from_SOCP_data, build_qp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_kats.json")


def kats():
    with open(GOLDEN) as f:
        return json.load(f)


def dense(spec, n=None):
    if isinstance(spec, str):
        if spec.startswith("identity"):
            return np.eye(int(spec[len("identity"):]))
        if spec.startswith("zeros"):
            k = int(spec[len("zeros"):])
            return np.zeros((k, k))
        raise ValueError(spec)
    return np.array(spec, dtype=float)


def model_from(spec):
    """QP_info of one KAT entry (canonical layout, 0-based cone offsets)."""
    Qd = spec.get("Q_diag")
    A = dense(spec["A"]) if "A" in spec else None
    if "Q" in spec:
        Qm = dense(spec["Q"])
        n = Qm.shape[0]
    elif Qd is not None:
        Qm = np.diag(Qd)
        n = len(Qd)
    elif A is not None:
        n = A.shape[1]
        Qm = np.zeros((n, n))
    else:
        n = spec["n"]
        Qm = np.zeros((n, n))
    if A is None:
        A = np.zeros((0, n))
    m = A.shape[0]
    rhs = np.array(spec.get("rhs", np.zeros(m)), float)
    neq = spec.get("number_eq", 0)
    nin = spec.get("number_ineq", 0)
    nl = spec.get("number_linear_con", neq + nin)
    ci = spec.get("soc_con_idx0", [nl])
    vi = spec.get("soc_var_idx0", [n])
    c = np.array(spec.get("c", -np.array(spec["a"]) if "a" in spec else np.zeros(n)), float)
    l = np.array(spec.get("l", np.full(n, -np.inf)), float)
    u = np.array(spec.get("u", np.full(n, np.inf)), float)
    if "number_linear_con" in spec and "number_eq" not in spec and "number_ineq" not in spec:
        nin = nl
    return build_from_SOCP_data(sp.csr_matrix(Qm), c, sp.csr_matrix(A), rhs,
                                [i + 1 for i in ci], neq, nin, l, u, [i + 1 for i in vi],
                                verbose=False)


def trust_region_optimum(Qd, c):
    """min 1/2 xi'D xi + c'xi s.t. |xi| <= 1 (x1 = 1 fixed): secular equation by bisection."""
    d = np.array(Qd[1:], float)
    g = np.array(c[1:], float)

    def xi(mu):
        return -g / (d + mu)

    if np.linalg.norm(xi(0.0)) <= 1.0:
        return np.concatenate([[1.0], xi(0.0)])
    lo, hi = 0.0, 1.0
    while np.linalg.norm(xi(hi)) > 1.0:
        hi *= 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if np.linalg.norm(xi(mid)) > 1.0:
            lo = mid
        else:
            hi = mid
    return np.concatenate([[1.0], xi(0.5 * (lo + hi))])


def params(**kw):
    p = hprsocp.HPRSOCP_parameters(verbose=False, print_frequency=-1, warm_up=False)
    for k, v in kw.items():
        setattr(p, k, v)
    return p
