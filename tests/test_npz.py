"""The .npz writer of the auto-save checkpoint (hpr-socp_amd/csrc/npz.cpp),
built with g++ into a temporary driver and read back with numpy.load
(allow_pickle=False) and zipfile.  Host code only: no GPU involved.

Pins: the CRC-32 check value of the ZIP polynomial ("123456789" ->
0xCBF43926, and the published value 0x414FA339 of the pangram), zipfile's own
CRC verification of every member (testzip), the ZIP64 size fields read back by
zipfile, and the dtype/shape/value of every member type the checkpoint uses.
"""
import os
import shutil
import subprocess
import zipfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "hpr-socp_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("npz") / "npz_driver"
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-I", CSRC, "-o", str(out),
                    os.path.join(HERE, "npz_driver.cpp"), os.path.join(CSRC, "npz.cpp")], check=True)
    return str(out)


@pytest.mark.parametrize("n", [0, 1, 1000, 300001])
def test_npz_roundtrip(driver, tmp_path, n):
    path = str(tmp_path / "state.npz")
    res = subprocess.run([driver, path, str(n)], check=True, capture_output=True, text=True)
    lines = res.stdout.split()
    assert lines[0] == "cbf43926"
    assert lines[1] == lines[2] == "414fa339"
    assert not os.path.exists(path + ".tmp")
    with zipfile.ZipFile(path) as z:
        assert z.testzip() is None  # every member's CRC-32 verified
        names = sorted(z.namelist())
        info = z.getinfo("current/x_org.npy")
        assert info.file_size == info.compress_size
        assert info.compress_type == zipfile.ZIP_STORED
    assert "current/x_org.npy" in names and "best/y_org.npy" in names
    with np.load(path, allow_pickle=False) as d:
        x = d["current/x_org"]
        assert x.dtype == np.float64 and x.shape == (n,)
        np.testing.assert_array_equal(x, 0.5 * np.arange(n) - 3.0)
        assert d["best/y_org"].shape == (0,)
        s = d["current/sigma"]
        assert s.shape == () and s.dtype == np.float64 and float(s) == 0.125
        it = d["current/iteration"]
        assert it.dtype == np.int64 and int(it) == 123456789012
        dn = d["parameters/device_number"]
        assert dn.dtype == np.int32 and int(dn) == -7
        assert d["parameters/auto_save"].dtype == np.bool_ and bool(d["parameters/auto_save"]) is True
        assert bool(d["parameters/warm_up"]) is False
        assert str(d["parameters/bc_scaling_norm_type"]) == "l2"
        assert str(d["current/timestamp"]) == "2026-12-03T10:11:12.345"


def test_npz_rewrite_replaces(driver, tmp_path):
    path = str(tmp_path / "state.npz")
    subprocess.run([driver, path, "50"], check=True, capture_output=True)
    subprocess.run([driver, path, "7"], check=True, capture_output=True)
    with np.load(path) as d:
        assert d["current/x_org"].shape == (7,)


def test_npz_unwritable_path_fails(driver, tmp_path):
    path = str(tmp_path / "no_such_dir" / "state.npz")
    res = subprocess.run([driver, path, "5"], capture_output=True, text=True)
    assert res.returncode == 1
    assert "cannot open" in res.stderr
