"""The backend's batched golden-section search (hpr-socp_amd/csrc/golden.h) visits
exactly the points of the reference's sequential loop (unified_golden_Q_diag,
unified_operations.jl:371-449) and returns the same sigma bit for bit, for every
lookahead depth; the answer is the minimiser of a dense scan.  The header is
shared with the device; here only its host side is compiled (hipcc
--offload-host-only) and run."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "hpr-socp_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "golden.h"
int main() {
    int n; double a, b;
    if (scanf("%d %lf %lf", &n, &a, &b) != 3) return 1;
    std::vector<double> c(n), d(n), q(n);
    for (int i = 0; i < n; ++i) if (scanf("%lf %lf %lf", &c[i], &d[i], &q[i]) != 3) return 1;
    auto g = [&](double x) { double s = 0; for (int i = 0; i < n; ++i) s += c[i] * (d[i] / (1.0 + x * q[i])); return s; };
    long evals = 0;
    auto ev = [&](const std::vector<double>& xs, std::vector<double>& gs) {
        gs.resize(xs.size()); evals++; for (size_t k = 0; k < xs.size(); ++k) gs[k] = g(xs[k]); };
    printf("%.17g\n", hpr::golden_sequential(a, b, 1e-12, 1e12, 1e-13, g));
    for (int L = 1; L <= 6; ++L) { evals = 0; double x = hpr::golden_speculative(a, b, 1e-12, 1e12, 1e-13, L, ev);
        printf("%.17g %ld\n", x, evals); }
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver():
    tmp = tempfile.mkdtemp()
    src = os.path.join(tmp, "g.cpp")
    exe = os.path.join(tmp, "g")
    with open(src, "w") as f:
        f.write(DRIVER)
    hipcc = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-I", CSRC, "-o", exe, src],
                   check=True)
    return exe


@pytest.mark.parametrize("seed", range(6))
def test_speculative_equals_sequential(driver, seed):
    rng = np.random.default_rng(seed)
    n = 50
    c = rng.standard_normal(n)
    q = rng.uniform(0, 2, n) * (rng.random(n) < 0.7)
    d = q * c  # QATdy = diag(Q) .* ATdy
    a, b = rng.uniform(0.01, 5, 2)
    inp = f"{n} {a:.17g} {b:.17g}\n" + "\n".join(f"{c[i]:.17g} {d[i]:.17g} {q[i]:.17g}" for i in range(n))
    out = subprocess.run([driver], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    seq = float(out[0])
    rounds = []
    for line in out[1:7]:
        x, ev = line.split()
        assert float(x) == seq  # bit for bit
        rounds.append(int(ev))
    assert rounds[4] < rounds[0]  # batching cuts device round trips
    f = lambda x: a * x + b / x + x * x * np.sum(c * (d / (1 + x * q)))
    grid = np.logspace(-4, 4, 20001)
    vals = np.array([f(x) for x in grid])
    xs = grid[np.argmin(vals)]
    assert f(seq) <= f(xs) * (1 + 1e-9) + 1e-12
