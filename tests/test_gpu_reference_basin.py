"""The reference's own shipped basin on the GPU: its Black Sea land mask
(mask.txt, read as gridcon reads it) and bathymetry (topo.dat, read with
the DIOGIN reader into the box init_pars reads, Control/init_pars.f90:
151-152), 289 x 163 points as Inc/basinpar.fi sets them.  The
stage-by-stage path on spherical metrics and the fused paths on uniform
4 km metrics against the oracle."""
import os

import numpy as np
import pytest

import oracle
from _util import STATE, assert_bitwise, rel_l2
from inmsom_bt import BarotropicModel
from inmsom_bt.cases import reference_basin

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "blacksea")


def _basin(metrics):
    b = reference_basin(os.path.join(DATA, "mask.txt"), DATA, metrics=metrics)
    j, i = np.meshgrid(np.arange(b.ny), np.arange(b.nx), indexing="ij")
    bump = 0.5 * np.exp(-((i - 150.0) ** 2 + (j - 80.0) ** 2) / (2 * 12.0 ** 2)) * b.wet
    b.ssh, b.sshp = bump.copy(), bump.copy()
    return b


def test_reference_basin_stage_by_stage_bitwise():
    print("black sea, spherical metrics, stage-by-stage path", flush=True)
    b = _basin("spherical")
    steps, nstep = 40, 2
    want = oracle.run(b, steps, nstep, out_steps=[steps])[steps]
    with BarotropicModel(b) as m:
        m.model_steps(steps, nstep=nstep)
        got = m.get_state()
    assert_bitwise(got, want, STATE, "black sea stage-by-stage")
    assert np.abs(got["ssh"]).max() > 0.1


def test_reference_basin_fused():
    print("black sea, uniform metrics, fused paths", flush=True)
    b = _basin("uniform")
    steps, nstep = 100, 2
    want = oracle.run(b, steps, nstep, out_steps=[steps])[steps]
    with BarotropicModel(b) as m:
        m.run(steps, "fp64_exact", nstep=nstep)
        exact = m.get_state()
    assert_bitwise(exact, want, STATE, "black sea fused fp64_exact")
    with BarotropicModel(b) as m:
        m.run(steps, "fp64", nstep=nstep)
        f64 = m.get_state()
    with BarotropicModel(b) as m:
        m.run(steps, "fp32", nstep=nstep)
        f32 = m.get_state()
    for k in ("ssh", "u", "v"):
        assert rel_l2(f64[k], want[k]) <= 1e-12, (k, rel_l2(f64[k], want[k]))
        assert rel_l2(f32[k], want[k]) <= 1e-5, (k, rel_l2(f32[k], want[k]))
