"""Measurement hooks used by bench.py: the resident-operand SpMV timer and the
gather-ceiling probe (include/hprsocp.h).  They return positive times, reject
invalid arguments, and leave the handle usable (a solve after them matches a
solve on a fresh handle bit for bit)."""
import numpy as np
import pytest

import hprsocp
from kat_models import params
from test_gpu_parity import model

pytestmark = pytest.mark.gpu


def test_time_spmv_and_probe(gpu_lib):
    m = model("c1_tiny")
    with hprsocp.Solver(m, params(stoptol=1e-6)) as g:
        ta, tat = g.time_spmv(False, 5), g.time_spmv(True, 5)
        print(f"[result] c1_tiny SpMV A {ta:.4f} ms, A^T {tat:.4f} ms", flush=True)
        assert ta > 0 and tat > 0
        for srt in (False, True):
            t = g.probe_gather(200000, 4096, 8, srt)
            print(f"[result] probe sorted={srt}: {t:.4f} ms", flush=True)
            assert t > 0
        with pytest.raises(hprsocp.HPRSOCPError):
            g.probe_gather(200000, 1 << 22, 8, False)  # column offsets would not fit the packed word
        with pytest.raises(hprsocp.HPRSOCPError):
            g.probe_gather(100, 4096, 8, True)  # less than one chunk
        with pytest.raises(hprsocp.HPRSOCPError):
            g.time_spmv(False, 0)
        g.estimate_lambda()
        r1 = g.solve()
    with hprsocp.Solver(m, params(stoptol=1e-6)) as g2:
        g2.estimate_lambda()
        r2 = g2.solve()
    assert r1.iter == r2.iter and r1.primal_obj == r2.primal_obj
    np.testing.assert_array_equal(r1.x, r2.x)
