"""hprsocp_project_cones_noalloc (no allocation, fixed device scratch) against
hprsocp_project_cones (scratch allocated per call), through the C ABI: two
calls in a row, the second with a smaller input, in place and out of place;
the outputs agree bit for bit (same kernels, same tiling, same summation
order).  Sizes beyond the scratch's limits are rejected before any work is
queued."""
import numpy as np
import pytest

import hprsocp

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _case(seed, sizes, lead=2, tail=3):
    rng = np.random.default_rng(seed)
    idx = np.concatenate([[lead], lead + np.cumsum(sizes)]).astype(np.int32)
    v = rng.normal(size=int(idx[-1]) + tail)
    for k in range(len(sizes)):
        s, e = int(idx[k]), int(idx[k + 1])
        if e - s >= 2 and k % 3:
            v[s] = (1.5 if k % 3 == 1 else -1.5) * np.linalg.norm(v[s + 1:e])
    return v, idx


def _run(v, idx, noalloc, inplace):
    st = torch.cuda.current_stream()
    dv = torch.from_numpy(v).cuda()
    di = torch.from_numpy(idx).cuda()
    dout = dv if inplace else torch.full_like(dv, -5.0)
    torch.cuda.synchronize()
    hprsocp.project_cones(dv.data_ptr(), dout.data_ptr(), v.size, di.data_ptr(), len(idx) - 1,
                          st.cuda_stream, noalloc=noalloc)
    torch.cuda.synchronize()
    return dout.cpu().numpy()


@pytest.mark.parametrize("inplace", [False, True])
def test_noalloc_matches_allocating(gpu_lib, inplace):
    calls = [
        ("large", _case(1, [300000, 5, 2049, 0, 4096, 7] + [3] * 20000)),
        ("small", _case(2, [9, 1, 2, 2050, 4])),
    ]
    for name, (v, idx) in calls:
        print(f"[case] no-alloc {name} n={v.size} ncones={len(idx) - 1} inplace={inplace}", flush=True)
        want = _run(v, idx, noalloc=False, inplace=inplace)
        got = _run(v, idx, noalloc=True, inplace=inplace)
        np.testing.assert_array_equal(got, want)
        assert not np.array_equal(got, v)  # the projection changed something


def test_noalloc_limits(gpu_lib):
    dv = torch.zeros(8, dtype=torch.float64, device="cuda")
    di = torch.zeros(65538, dtype=torch.int32, device="cuda")
    with pytest.raises(hprsocp.HPRSOCPError, match="ncones"):
        hprsocp.project_cones(dv.data_ptr(), dv.data_ptr(), 8, di.data_ptr(), 65537, noalloc=True)
    n_big = (131072 - 1) * 2048 + 1  # ceil(n / 2048) + 1 cone = 131073 tiles
    with pytest.raises(hprsocp.HPRSOCPError, match="exceeds the limit"):
        hprsocp.project_cones(dv.data_ptr(), dv.data_ptr(), n_big, di.data_ptr(), 1, noalloc=True)
    # at the limit the call is accepted (all cones empty: nothing to project)
    hprsocp.project_cones(dv.data_ptr(), dv.data_ptr(), 8, di.data_ptr(), 65536, noalloc=True)
    torch.cuda.synchronize()
    assert torch.all(dv == 0)
