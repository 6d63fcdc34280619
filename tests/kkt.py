"""Independent KKT certificate for returned (x, y, z), in original units.

Recomputes the reference's relative measures (src/algorithm.jl:77-175) from
the unscaled problem and the solver outputs only — no solver internals:
  err_Rp  = |Pi(Ax) - Ax|_inf / (1 + max(|b|_inf, |Ax|_inf))   (linear rows: box,
            cone rows: Pi_K(Ax - b) - (Ax - b))
  err_Rd  = |Qx + c - A'y - z|_inf / (1 + max(|c|_inf, |A'y|_inf, |Qx|_inf))
  rel_gap from the primal objective and a dual objective built from support
  functions of the bounds (so it does not reuse the solver's s vector)
plus the dual-cone violations of y (cone rows) and z (cone variables).
"""
import numpy as np
import scipy.sparse as sp


def soc_proj(v):
    t, xi = v[0], v[1:]
    ns = np.linalg.norm(xi)
    if ns <= -t:
        return np.zeros_like(v)
    if ns <= t:
        return v.copy()
    pt = (ns + t) / 2
    return np.concatenate([[pt], (pt / ns) * xi])


def q_apply(model, v):
    Q = model.Q
    if isinstance(Q, tuple):
        _, F, d = Q
        return F.T @ (F @ v) + (0 if d is None else d * v)
    return Q @ v


def certificate(model, x, y, z):
    A = sp.csr_matrix(model.A)
    m, n = A.shape
    m1 = model.number_linear_con
    ci = np.asarray(model.SOC_con_idx) - 1
    vi = np.asarray(model.SOC_var_idx) - 1
    ax = A @ x
    rp = np.zeros(m)
    rp[:m1] = np.clip(ax[:m1], model.AL[:m1], model.AU[:m1]) - ax[:m1]
    for k in range(len(ci) - 1):
        s, e = ci[k], ci[k + 1]
        v = ax[s:e] - model.soc_rhs_full[s:e]
        rp[s:e] = soc_proj(v) - v
    al = np.where(np.isinf(model.AL), 0, model.AL)
    au = np.where(np.isinf(model.AU), 0, model.AU)
    nb = np.max(np.maximum(np.abs(al), np.abs(au)), initial=0.0)
    nb = max(nb, np.max(np.abs(model.soc_rhs), initial=0.0))
    err_rp = np.max(np.abs(rp), initial=0.0) / (1 + max(nb, np.max(np.abs(ax), initial=0.0)))
    qx = q_apply(model, x)
    aty = A.T @ y
    rd = qx + model.c - aty - z
    err_rd = np.max(np.abs(rd), initial=0.0) / (1 + max(np.max(np.abs(model.c), initial=0.0),
                                                      np.max(np.abs(aty), initial=0.0),
                                                      np.max(np.abs(qx), initial=0.0)))
    pobj = model.c @ x + 0.5 * x @ qx + model.obj_constant
    # dual objective: -1/2 x'Qx + support terms
    d = -0.5 * x @ qx + model.obj_constant
    yl = y[:m1]
    d += np.sum(np.where(yl > 0, yl * np.where(np.isinf(model.AL[:m1]), 0, model.AL[:m1]),
                         yl * np.where(np.isinf(model.AU[:m1]), 0, model.AU[:m1])))
    d += y[m1:] @ model.soc_rhs_full[m1:]
    n1 = model.number_lu_x
    zl = z[:n1]
    d += np.sum(np.where(zl > 0, zl * np.where(np.isinf(model.l[:n1]), 0, model.l[:n1]),
                         zl * np.where(np.isinf(model.u[:n1]), 0, model.u[:n1])))
    gap = abs(pobj - d) / (1 + max(abs(pobj), abs(d)))
    yviol = 0.0
    for k in range(len(ci) - 1):
        s, e = ci[k], ci[k + 1]
        yviol = max(yviol, np.linalg.norm(y[s + 1:e]) - y[s])
    zviol = 0.0
    for k in range(len(vi) - 1):
        s, e = vi[k], vi[k + 1]
        zviol = max(zviol, np.linalg.norm(z[s + 1:e]) - z[s])
    return {"err_Rp": err_rp, "err_Rd": err_rd, "rel_gap": gap, "primal_obj": pobj,
            "dual_obj": d, "y_cone_violation": yviol, "z_cone_violation": zviol}
