"""The C-ABI library on the CPU: it loads, exports every function
include/inmsom_bt.h declares, and its host-side routines (mask
construction, Rayleigh coefficient) equal the oracle's bit for bit.  No
GPU call is made."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from _util import bitwise_equal
from inmsom_bt import _capi, gridcon, make_basin, randomize_state, rayleigh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    with open(os.path.join(ROOT, "include", "inmsom_bt.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(inmsom_bt_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = _capi.lib()
    names = header_functions()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_capi.SIGNATURES), "python binding and header disagree"


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gridcon_matches_oracle(seed):
    b = make_basin(50, 37, seed=seed, sigma=2.5, hland=80.0)
    for per in ((0, 0), (1, 0), (0, 1), (1, 1)):
        a, o = gridcon(b.lbasins, per), oracle.gridcon(b.lbasins, *per)
        for k in a:
            assert np.array_equal(a[k], o[k]), (per, k)


@pytest.mark.parametrize("metrics", ["uniform", "stretched"])
def test_rayleigh_matches_oracle(metrics):
    b = make_basin(64, 48, seed=4, sigma=3.0, metrics=metrics)
    assert bitwise_equal(rayleigh(b), oracle.OracleGrid(b).rayleigh())


def test_error_path_sets_message():
    L = _capi.lib()
    lu = np.zeros(4, dtype=np.float32)
    code = L.inmsom_bt_gridcon(2, 2, 0, 0, None, *[lu.ctypes.data_as(_capi._fp)] * 7)
    assert code == _capi.EINVAL
    assert b"at least 5" in L.inmsom_bt_last_error()
    code = L.inmsom_bt_rayleigh(None, 1e-4, None)
    assert code == _capi.EINVAL and b"null" in L.inmsom_bt_last_error()


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_more_than_one_gpu(n):
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and
    says so, before touching a device."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
