"""Seeded synthetic SOCP instances with a known-feasible, bounded construction.

Construction (shared by every workload): draw an interior primal point x0
(box coordinates U(-1,1) inside l = x0 - U(.5,2), u = x0 + U(.5,2); cone
blocks with t = ||xi|| + U(.1,1)), set the linear rows' bounds around A1 x0
(+-U(0,1)), set the cone-row right-hand side b2 = A2 x0 - s with s strictly
inside the cones, and set c = A1'y1 + A2'y2 + z - Q x0 from a dual-feasible
(y, z) (cone parts of y2 and z strictly inside the self-dual cones).

Layout is the canonical one of include/hprsocp.h: rows [linear | cone rows],
columns [boxed | cone variables].  numpy Generator(PCG64(seed)).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import build_qp, QP_info


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _interior_cone_point(rng, sizes):
    """Concatenated cone blocks (t, xi) with t = ||xi|| + U(0.1, 1)."""
    out = np.empty(int(np.sum(sizes)) if len(sizes) else 0)
    p = 0
    for s in sizes:
        xi = rng.standard_normal(s - 1)
        out[p] = np.linalg.norm(xi) + rng.uniform(0.1, 1.0)
        out[p + 1:p + s] = xi
        p += s
    return out


def _interior_cone_point_fast(rng, sizes):
    """Vectorised variant for many cones."""
    sizes = np.asarray(sizes, dtype=np.int64)
    total = int(sizes.sum())
    if total == 0:
        return np.zeros(0)
    v = rng.standard_normal(total)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    sq = v * v
    sq[starts] = 0.0
    norms = np.sqrt(np.add.reduceat(sq, starts)) if total else np.zeros(0)
    v[starts] = norms + rng.uniform(0.1, 1.0, size=len(sizes))
    return v


def _random_rows(rng, nrows, ncols, lengths):
    """CSR with the given row lengths, uniform random columns, N(0,1) values."""
    lengths = np.minimum(np.asarray(lengths, dtype=np.int64), ncols)
    indptr = np.concatenate([[0], np.cumsum(lengths)])
    nnz = int(indptr[-1])
    cols = rng.integers(0, ncols, size=nnz, dtype=np.int64)
    vals = rng.standard_normal(nnz)
    M = sp.csr_matrix((vals, cols.astype(np.int32), indptr), shape=(nrows, ncols))
    M.sum_duplicates()
    M.sort_indices()
    return M


def assemble(rng, A1, A2, n1, var_sizes, con_sizes, Q=None, Qx_fn=None):
    """Feasibility construction around given A1 (linear rows) and A2 (cone rows)."""
    n = A1.shape[1] if A1 is not None else A2.shape[1]
    m1 = A1.shape[0]
    m2 = A2.shape[0]
    x0 = np.empty(n)
    x0[:n1] = rng.uniform(-1, 1, n1)
    x0[n1:] = _interior_cone_point_fast(rng, var_sizes)
    l = np.full(n, -np.inf)
    u = np.full(n, np.inf)
    l[:n1] = x0[:n1] - rng.uniform(0.5, 2.0, n1)
    u[:n1] = x0[:n1] + rng.uniform(0.5, 2.0, n1)
    a1x = A1 @ x0
    AL = np.concatenate([a1x - rng.uniform(0, 1, m1), np.full(m2, -np.inf)])
    AU = np.concatenate([a1x + rng.uniform(0, 1, m1), np.full(m2, np.inf)])
    s = _interior_cone_point_fast(rng, con_sizes)
    b2 = A2 @ x0 - s
    y1 = rng.standard_normal(m1)
    y2 = _interior_cone_point_fast(rng, con_sizes)
    z = np.empty(n)
    z[:n1] = rng.standard_normal(n1)
    z[n1:] = _interior_cone_point_fast(rng, var_sizes)
    c = A1.T @ y1 + A2.T @ y2 + z
    if Qx_fn is not None:
        c = c - Qx_fn(x0)
    A = sp.vstack([A1, A2], format="csr")
    A.sort_indices()
    con_idx = np.concatenate([[m1], m1 + np.cumsum(con_sizes)]).astype(np.int64)
    var_idx = np.concatenate([[n1], n1 + np.cumsum(var_sizes)]).astype(np.int64)
    model = build_qp(A, c, AL, AU, l, u, b2, m1, n1, con_idx, var_idx, Q=Q)
    model.x0 = x0
    return model


def make_c0(seed: int = 0) -> QP_info:
    """Plain FP64 small SOCP: n1=800 boxed + 4 variable cones of 50; A1 400x1000
    at 2% density, A2 210x1000 at 5% density as 10 cone row blocks of 21; Q=0."""
    rng = _rng(seed)
    n1, var_sizes = 800, [50] * 4
    n = n1 + sum(var_sizes)
    A1 = sp.random(400, n, density=0.02, format="csr", random_state=rng,
                   data_rvs=rng.standard_normal)
    A2 = sp.random(210, n, density=0.05, format="csr", random_state=rng,
                   data_rvs=rng.standard_normal)
    return assemble(rng, A1, A2, n1, var_sizes, [21] * 10)


def make_c1(scale: float = 1.0, seed: int = 1000, q_form: str = "factored") -> QP_info:
    """Large single QP-SOCP: n = 1e6 boxed, m_A = 5e5 linear rows plus 10,000
    constraint cones of sizes U[3,201]; log-normal row lengths (mean 20);
    Q = F'F + diag(U(0,1)), F 1e5 x 1e6 with 5 nnz per column.  `scale` shrinks
    every count (tests use scale << 1)."""
    rng = _rng(seed)
    n = max(int(1_000_000 * scale), 10)
    m1 = max(int(500_000 * scale), 2)
    ncones = max(int(10_000 * scale), 1)
    con_sizes = rng.integers(3, 202, size=ncones)
    m2 = int(con_sizes.sum())
    sig = 1.0
    mu = np.log(20.0) - 0.5 * sig * sig
    lens = np.maximum(1, np.rint(rng.lognormal(mu, sig, size=m1 + m2))).astype(np.int64)
    M = _random_rows(rng, m1 + m2, n, lens)
    A1, A2 = M[:m1], M[m1:]
    mf = max(int(100_000 * scale), 1)
    # F with 5 nnz per column: build F^T row-wise
    FT = _random_rows(rng, n, mf, np.full(n, 5))
    F = FT.T.tocsr()
    F.sort_indices()
    d = rng.uniform(0, 1, n)
    if q_form == "factored":
        Q = ("factored", F, d)
    else:
        Q = (F.T @ F + sp.diags(d)).tocsr()
    Qx = lambda v: F.T @ (F @ v) + d * v
    return assemble(rng, A1, A2, n, [], con_sizes, Q=Q, Qx_fn=Qx)


def make_c2(scale: float = 1.0, seed: int = 2000) -> QP_info:
    """Cone-projection stress QP-SOCP: 1e5 size-3 + 100 size-10,000 cones on
    each side (variables and rows), 1e5 boxed variables and 1e5 linear rows;
    A = 3x3 dense N(0,1) blocks + 4 random couplings per row; Q = diag U(0.01,1)."""
    rng = _rng(seed)
    k3 = max(int(100_000 * scale), 1)
    kb = max(int(100 * scale), 1)
    big = 10_000 if scale >= 0.01 else max(int(10_000 * scale * 10), 4)
    n1 = max(int(100_000 * scale), 3)
    n1 -= n1 % 3
    m1 = n1
    var_sizes = [3] * k3 + [big] * kb
    con_sizes = var_sizes
    n = n1 + sum(var_sizes)
    m = m1 + sum(con_sizes)
    # 3x3 dense blocks on a random block permutation (rows in triples)
    nbr = m // 3
    nbc = n // 3
    rb = np.arange(nbr)
    cb = rng.integers(0, nbc, size=nbr)
    r_idx = (3 * rb[:, None, None] + np.arange(3)[None, :, None]) * np.ones((1, 1, 3), np.int64)
    c_idx = (3 * cb[:, None, None] + np.arange(3)[None, None, :]) * np.ones((1, 3, 1), np.int64)
    rows_b = r_idx.ravel()
    cols_b = c_idx.ravel()
    vals_b = rng.standard_normal(rows_b.size)
    rows_r = np.repeat(np.arange(m), 4)
    cols_r = rng.integers(0, n, size=rows_r.size)
    vals_r = rng.standard_normal(rows_r.size)
    Afull = sp.csr_matrix((np.concatenate([vals_b, vals_r]),
                           (np.concatenate([rows_b, rows_r]), np.concatenate([cols_b, cols_r]))),
                          shape=(m, n))
    Afull.sum_duplicates()
    A1, A2 = Afull[:m1], Afull[m1:]
    d = rng.uniform(0.01, 1.0, n)
    Q = sp.diags(d).tocsr()
    return assemble(rng, A1, A2, n1, var_sizes, con_sizes, Q=Q, Qx_fn=lambda v: d * v)


def make_small_qp(seed: int = 7, n1=60, var_sizes=(5, 9), m1=30, con_sizes=(4, 7, 3),
                  density=0.15, q="diag") -> QP_info:
    """Small mixed instance for parity tests (q in none/diag/csr/factored)."""
    rng = _rng(seed)
    n = n1 + sum(var_sizes)
    m2 = sum(con_sizes)
    A1 = sp.random(m1, n, density=density, format="csr", random_state=rng, data_rvs=rng.standard_normal)
    A2 = sp.random(m2, n, density=density, format="csr", random_state=rng, data_rvs=rng.standard_normal)
    Q = None
    Qx = None
    if q == "diag":
        d = rng.uniform(0.01, 1.0, n)
        Q = sp.diags(d).tocsr()
        Qx = lambda v: d * v
    elif q == "csr":
        F = sp.random(n // 2, n, density=0.1, format="csr", random_state=rng, data_rvs=rng.standard_normal)
        Qm = (F.T @ F + sp.diags(rng.uniform(0, 1, n))).tocsr()
        Q = Qm
        Qx = lambda v: Qm @ v
    elif q == "factored":
        F = sp.random(n // 2, n, density=0.1, format="csr", random_state=rng, data_rvs=rng.standard_normal)
        d = rng.uniform(0, 1, n)
        Q = ("factored", F, d)
        Qx = lambda v: F.T @ (F @ v) + d * v
    return assemble(rng, A1, A2, n1, list(var_sizes), list(con_sizes), Q=Q, Qx_fn=Qx)


# ------------------------------------------------------------------ batches
def _simplex_projection(v):
    """Euclidean projection of v onto {x >= 0, sum x = 1} (sort-based)."""
    u = np.sort(v)[::-1]
    css = np.cumsum(u) - 1.0
    k = np.arange(1, v.size + 1)
    rho = np.nonzero(u - css / k > 0)[0][-1]
    theta = css[rho] / (rho + 1.0)
    return np.maximum(v - theta, 0.0)


def _min_risk_upper(F, steps=60):
    """Risk ||F^T x|| of an approximate minimum-risk portfolio (accelerated
    projected gradient on the simplex); an upper bound of the minimum."""
    n = F.shape[0]
    L = 2.0 * np.linalg.norm(F, 2) ** 2 + 1e-12
    x = np.full(n, 1.0 / n)
    yv, t = x.copy(), 1.0
    for _ in range(steps):
        g = 2.0 * (F @ (F.T @ yv))
        xn = _simplex_projection(yv - g / L)
        tn = 0.5 * (1.0 + np.sqrt(1.0 + 4.0 * t * t))
        yv = xn + ((t - 1.0) / tn) * (xn - x)
        x, t = xn, tn
    return float(np.linalg.norm(F.T @ x))


def make_portfolio(rng, n: int, k: int) -> QP_info:
    """min -mu'x  s.t.  sum x = 1, x >= 0, ||F'x||_2 <= gamma.

    F is n x k dense N(0, 1/k), mu ~ N(0.05, 0.02).  gamma = 1.5 x max(r_min,
    r_eq / 4) with r_min the risk of an approximate minimum-risk portfolio (an
    upper bound of the minimum) and r_eq = ||F' 1/n|| the equal-weight risk:
    for n > k the minimum risk is 0 (F' has a nonnegative null vector on the
    simplex), and gamma = 1.5 x 0 would leave only the cone's apex; the r_eq
    floor keeps the instance strictly feasible with a binding risk limit.
    Layout: one linear row (sum x = 1), then the cone
    block (gamma, F'x) in Q^{k+1} as an empty t-row with soc_rhs = -gamma and k
    rows F'; every variable boxed in [0, inf)."""
    F = rng.standard_normal((n, k)) / np.sqrt(k)
    mu = rng.normal(0.05, 0.02, n)
    r_eq = float(np.linalg.norm(F.T @ np.full(n, 1.0 / n)))
    gamma = 1.5 * max(_min_risk_upper(F), 0.25 * r_eq)
    ones = sp.csr_matrix(np.ones((1, n)))
    rows = sp.vstack([ones, sp.csr_matrix((1, n)), sp.csr_matrix(F.T)], format="csr")
    rows.sort_indices()
    m = k + 2
    AL = np.concatenate([[1.0], np.full(k + 1, -np.inf)])
    AU = np.concatenate([[1.0], np.full(k + 1, np.inf)])
    soc_rhs = np.zeros(k + 1)
    soc_rhs[0] = -gamma
    model = build_qp(rows, -mu, AL, AU, np.zeros(n), np.full(n, np.inf), soc_rhs, 1, n,
                     np.array([1, m]), np.array([n]))
    model.gamma = gamma
    return model


def make_c3(count: int = 512, seed: int = 3000, n_lo: int = 200, n_hi: int = 20000, k_lo: int = 5,
            k_hi: int = 50):
    """Batch of `count` portfolio SOCPs, n log-uniform in [n_lo, n_hi], k in U{k_lo..k_hi}."""
    out = []
    for i in range(count):
        rng = _rng(seed + i)
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        k = int(rng.integers(k_lo, k_hi + 1))
        out.append(make_portfolio(rng, n, k))
    return out


def make_robust_regression(rng, d: int, lam: float = 0.1):
    """min 1/2 ||r||^2 + lam 1'(x+ + x-)  s.t.  D(x+ - x-) - r = y,  ||W(x+ - x- - xbar)||_2 <= rho.

    D is (d/2) x d sparse with 10 N(0,1) entries per row, W = diag(U(0.5, 2)),
    x* has 10% N(0,1) nonzeros, y = D x* + 0.01 N(0,1), xbar = 0, rho = 2||x*||.
    Variables [x+ (d), x- (d), r (d/2)], x+- >= 0 and r free (all boxed);
    Q = diag(0, 0, 1) (diagonal); rows: d/2 equalities, then the cone block
    (rho, W(x+ - x-)) in Q^{d+1} as an empty t-row with soc_rhs = -rho.
    Returns (model, warm_x) with warm_x the point built from x* (x+- from its
    sign split, r = D x* - y)."""
    h = d // 2
    # 10 distinct columns per row: uniform draws, redrawn for the few rows with a repeat
    cols = rng.integers(0, d, size=(h, 10))
    while True:
        srt = np.sort(cols, axis=1)
        bad = np.nonzero(np.any(srt[:, 1:] == srt[:, :-1], axis=1))[0]
        if bad.size == 0:
            break
        cols[bad] = rng.integers(0, d, size=(bad.size, 10))
    cols = cols.ravel()
    rows = np.repeat(np.arange(h), 10)
    D = sp.csr_matrix((rng.standard_normal(h * 10), (rows, cols)), shape=(h, d))
    D.sum_duplicates()
    xs = np.where(rng.uniform(size=d) < 0.1, rng.standard_normal(d), 0.0)
    y = D @ xs + 0.01 * rng.standard_normal(h)
    W = rng.uniform(0.5, 2.0, d)
    rho = 2.0 * float(np.linalg.norm(xs))
    n = 2 * d + h
    A1 = sp.hstack([D, -D, -sp.eye(h)], format="csr")
    Wd = sp.diags(W)
    A2 = sp.vstack([sp.csr_matrix((1, n)), sp.hstack([Wd, -Wd, sp.csr_matrix((d, h))])], format="csr")
    A = sp.vstack([A1, A2], format="csr")
    A.sort_indices()
    m = h + d + 1
    AL = np.concatenate([y, np.full(d + 1, -np.inf)])
    AU = np.concatenate([y, np.full(d + 1, np.inf)])
    soc_rhs = np.zeros(d + 1)
    soc_rhs[0] = -rho
    c = np.concatenate([np.full(2 * d, lam), np.zeros(h)])
    l = np.concatenate([np.zeros(2 * d), np.full(h, -np.inf)])
    u = np.full(n, np.inf)
    Q = sp.diags(np.concatenate([np.zeros(2 * d), np.ones(h)])).tocsr()
    model = build_qp(A, c, AL, AU, l, u, soc_rhs, h, n, np.array([h, m]), np.array([n]), Q=Q)
    warm = np.concatenate([np.maximum(xs, 0), np.maximum(-xs, 0), D @ xs - y])
    return model, warm


def make_c4(count: int = 64, seed: int = 4000, d_lo: int = 10000, d_hi: int = 200000):
    """Batch of `count` robust-regression QP-SOCPs, d ~ U[d_lo, d_hi] (even);
    returns (models, warm_starts) with warm starts for the even-indexed half,
    perturbed as w * (1 + 0.01 N(0,1))."""
    models, warms = [], []
    for i in range(count):
        rng = _rng(seed + i)
        d = 2 * int(rng.integers(d_lo // 2, d_hi // 2 + 1))
        m, w = make_robust_regression(rng, d)
        models.append(m)
        warms.append(w * (1 + 0.01 * rng.standard_normal(w.size)) if i % 2 == 0 else None)
    return models, warms
