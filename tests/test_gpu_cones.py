"""hprsocp_project_cones (stand-alone SOC projection on device pointers) against
the oracle's projection (oracle_soc_project, kernels.jl:7585-7663 restated):
cones from 1 row to several 2048-row tiles, empty cones, rows outside every
cone, all three cases of the projection, in place and out of place; invalid
offsets are rejected with the output untouched.  Tolerances, relative to the
cone's scale: 1e-14 against the projection with an exactly rounded |xi|^2
(math.fsum); 1e-14 + 4e-16 len against the oracle, whose |xi|^2 is a
sequential sum."""
import ctypes as C
import math

import numpy as np
import pytest

import hprsocp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def oracle_project(oracle_lib, v, idx):
    fn = oracle_lib.oracle_soc_project
    fn.restype = None
    fn.argtypes = [C.POINTER(C.c_double), C.c_int32]
    out = v.copy()
    for k in range(len(idx) - 1):
        s, e = int(idx[k]), int(idx[k + 1])
        if e > s:
            seg = np.ascontiguousarray(out[s:e])
            fn(seg.ctypes.data_as(C.POINTER(C.c_double)), e - s)
            out[s:e] = seg
    return out


def fsum_project(seg):
    """Pi_Q of one cone with |xi| from an exactly rounded sum of squares."""
    t, xi = seg[0], seg[1:]
    ns = math.sqrt(math.fsum(xi * xi))
    if ns <= -t:
        return np.zeros_like(seg)
    if ns <= t:
        return seg.copy()
    pt = 0.5 * (ns + t)
    return np.concatenate([[pt], (pt / ns) * xi])


def make_case(seed, sizes, lead=3, tail=2):
    rng = np.random.default_rng(seed)
    idx = np.concatenate([[lead], lead + np.cumsum(sizes)]).astype(np.int32)
    n = int(idx[-1]) + tail
    v = rng.normal(size=n)
    # one cone inside, one in the polar cone, the rest mixed
    for k in range(len(sizes)):
        s, e = int(idx[k]), int(idx[k + 1])
        if e - s < 2:
            continue
        nx = np.linalg.norm(v[s + 1:e])
        if k % 3 == 0:
            v[s] = 1.5 * nx + 0.1
        elif k % 3 == 1:
            v[s] = -1.5 * nx - 0.1
    return v, idx


CASES = {
    "small": (1, [1, 2, 3, 5, 0, 4, 1, 3]),
    "tile_edges": (2, [2047, 2048, 2049, 2050, 4097, 1]),
    "huge": (3, [100001, 3, 250000, 0, 7]),
    "many_small": (4, [3] * 40000 + [5] * 1000),
}


@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_project_cones_vs_oracle(case, inplace, oracle_lib, gpu_lib):
    seed, sizes = CASES[case]
    v, idx = make_case(seed, sizes)
    print(f"[case] project_cones {case} n={v.size} ncones={len(idx) - 1} inplace={inplace}", flush=True)
    want = oracle_project(oracle_lib, v, idx)
    dv = torch.from_numpy(v).cuda()
    di = torch.from_numpy(idx).cuda()
    dout = dv if inplace else torch.full_like(dv, 7.0)
    st = torch.cuda.current_stream().cuda_stream
    hprsocp.project_cones(dv.data_ptr(), dout.data_ptr(), v.size, di.data_ptr(), len(idx) - 1, st)
    got = dout.cpu().numpy()
    inside = np.zeros(v.size, bool)
    scale = np.zeros(v.size)
    for k in range(len(idx) - 1):
        s, e = int(idx[k]), int(idx[k + 1])
        inside[s:e] = True
        if e > s:
            scale[s:e] = 1.0 + np.abs(v[s:e]).max()
    # against the oracle: its |xi|^2 is a sequential sum, accurate to ~len eps
    exact = v.copy()
    bound = np.zeros(v.size)
    for k in range(len(idx) - 1):
        s, e = int(idx[k]), int(idx[k + 1])
        if e > s:
            bound[s:e] = 1e-14 + 4e-16 * (e - s)
            exact[s:e] = fsum_project(v[s:e])
    err = np.abs(got[inside] - want[inside]) / scale[inside]
    assert np.all(err <= bound[inside]), err.max()
    # against the projection with an exactly rounded |xi|^2 (math.fsum)
    err = np.abs(got[inside] - exact[inside]) / scale[inside]
    assert err.max(initial=0.0) <= 1e-14, err.max()
    outside = ~inside
    np.testing.assert_array_equal(got[outside], v[outside] if inplace else 7.0)
    # all three cases of the projection occur
    modes = set()
    for k in range(len(idx) - 1):
        s, e = int(idx[k]), int(idx[k + 1])
        if e - s >= 2:
            modes.add("zero" if not np.any(want[s:e]) else
                      ("keep" if np.array_equal(want[s:e], v[s:e]) else "scale"))
    assert modes == {"zero", "keep", "scale"}


def test_project_cones_rejects_bad_offsets(gpu_lib):
    v = np.arange(10, dtype=np.float64)
    dv = torch.from_numpy(v).cuda()
    for bad in ([0, 4, 2, 10], [-1, 5], [0, 11], [3, 2]):
        di = torch.tensor(bad, dtype=torch.int32, device="cuda")
        dout = torch.full_like(dv, 3.0)
        with pytest.raises(hprsocp.HPRSOCPError) as ei:
            hprsocp.project_cones(dv.data_ptr(), dout.data_ptr(), 10, di.data_ptr(), len(bad) - 1)
        assert "cone_idx" in str(ei.value)
        assert torch.all(dout == 3.0)
    with pytest.raises(hprsocp.HPRSOCPError):
        hprsocp.project_cones(dv.data_ptr(), dv.data_ptr(), -1, 0, 1)
    hprsocp.project_cones(0, 0, 0, 0, 0)  # no cones: nothing to do
