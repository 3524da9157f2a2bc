"""The Fortran host (ISO_C_BINDING module + driver in the reference's
shape) on the GPU: its final state equals the oracle's bit for bit."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from _util import STATE, assert_bitwise
from inmsom_bt.cases import make_basin, randomize_state, write_case

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "inmsom-elfs_amd", "bin", "inmsom_bt_host")
PARDIR = os.path.join(ROOT, "inmsom-elfs_amd", "fortran", "par")


def _run_host(b, rdis, nsteps, mode, tmp):
    case = os.path.join(tmp, "case.bin")
    out = os.path.join(tmp, "out.bin")
    write_case(case, b, rdis, nsteps, 2, 0, 0, [])
    p = subprocess.run([HOST, PARDIR, case, out, mode, repr(b.g)], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = np.fromfile(out, dtype="<f8").reshape(6, b.ny, b.nx)
    return dict(zip(("ssh", "sshp", "u", "up", "v", "vp"), raw)), p.stdout


@pytest.mark.parametrize("mode", ["literal", "fp64_exact"])
def test_fortran_host_bitwise(mode):
    assert os.path.exists(HOST), "build the host with make -C inmsom-elfs_amd/fortran"
    b = make_basin(96, 64, seed=51, sigma=5.0)
    rd = oracle.OracleGrid(b).rayleigh()
    steps = 12  # ocean_run.par: nstep_barotrop = 2
    want = oracle.run(b, steps, 2, out_steps=[steps], rdis=rd)[steps]
    with tempfile.TemporaryDirectory() as tmp:
        got, out = _run_host(b, rd, steps, mode, tmp)
    assert "Time barotropic" in out
    assert_bitwise(got, want, STATE, f"fortran host {mode}")


def test_fortran_host_range_check_stops_run():
    """A NaN in the initial state makes the host stop after the first step
    with the range-check status, as the reference stops (dyncall.f90:452-516)."""
    b = make_basin(40, 30, seed=52, sigma=4.0)
    b.ssh[10, 10] = np.nan
    b.sshp[10, 10] = np.nan
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        case = os.path.join(tmp, "case.bin")
        write_case(case, b, rd, 3, 2, 0, 0, [])
        p = subprocess.run([HOST, PARDIR, case, os.path.join(tmp, "o.bin"), "literal"],
                           capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "range check failed at step 1" in p.stdout
