"""The Smagorinsky viscosity and its depth integral on the GPU
(Control/dyncall.f90:40-43: Function/mixing.f90:171-219 then
Function/depth.f90:223-254) against the oracle, bit for bit; the oracle is
pinned to the compiled reference (harness mode 10) and by the closed forms
of the formula."""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise, periodic_pair
from inmsom_bt import BarotropicModel, make_basin

pytestmark = pytest.mark.gpu

FIELDS = ("amts", "amuv", "amuv2d")


def _inputs(b, nz, seed):
    rng = np.random.default_rng(seed)
    sh = (b.ny, b.nx)
    dz = 0.5 + rng.random(nz)
    return {"dz": dz / dz.sum(), "stress_t": 1e-5 * rng.standard_normal((nz,) + sh),
            "stress_s": 1e-5 * rng.standard_normal((nz,) + sh), "amts": rng.random((nz,) + sh),
            "amuv": rng.random((nz,) + sh), "amuv2d": rng.random(sh)}


def _channel(nxi, nyi, seed, all_wet=False):
    b = make_basin(nxi, nyi, seed=seed, sigma=3.0, metrics="stretched")
    b.lbasins[:, :] = 0 if all_wet else 1
    b.lbasins[3:-3, :] = 0
    return b


CASES = [("closed stretched nz=20", lambda: make_basin(285, 159, seed=4, sigma=5.0, metrics="stretched"),
          (0, 0), 20),
         ("channel px nz=7", lambda: _channel(90, 40, 6), (1, 0), 7),
         ("doubly periodic nz=1", lambda: _channel(40, 30, 7, all_wet=True), (1, 1), 1),
         ("ragged 67x5 nz=3", lambda: make_basin(67, 5, seed=9, sigma=2.0), (0, 0), 3),
         ("largest nz=1000", lambda: make_basin(24, 18, seed=8, sigma=3.0), (0, 0), 1000)]


@pytest.mark.parametrize("name,make,periodicity,nz", CASES, ids=[c[0] for c in CASES])
def test_smagorinsky_matches_oracle(name, make, periodicity, nz):
    print(f"smagorinsky {name}", flush=True)
    b = make()
    f = _inputs(b, nz, seed=nz)
    visc2, dift = 1.0e3, 5.0e2
    og = oracle.OracleGrid(b, periodicity=periodicity)
    want = oracle.smagorinsky(og, f, visc2, dift)
    with BarotropicModel(b, periodicity=periodicity) as m:
        got = m.smagorinsky(f["dz"], f["stress_t"], f["stress_s"], f["amts"], f["amuv"], f["amuv2d"], visc2, dift)
        assert m.smagorinsky_timed(nz, 2) > 0.0
    got = dict(zip(FIELDS, got))
    assert_bitwise(got, want, FIELDS, f"smagorinsky {name}")
    for k in FIELDS:                     # untouched cells keep their values
        assert np.any(got[k] == f[k]) and np.any(got[k] != f[k]), k


def test_smagorinsky_device_arrays_feed_the_step():
    """The _dev entry on device arrays, twice in a row with different
    coefficients (nothing carried over between calls), and amuv2d handed to
    a barotropic step as its lateral viscosity mu equals the oracle's step
    with the oracle's amuv2d."""
    from inmsom_bt.hipmem import DeviceArray
    b, _, _ = periodic_pair((1, 0))       # a zonal channel whose state is zero on land
    nz = 20
    f = _inputs(b, nz, seed=3)
    og = oracle.OracleGrid(b, periodicity=(1, 0))
    with BarotropicModel(b, periodicity=(1, 0)) as m:
        d = {k: DeviceArray.from_host(np.ascontiguousarray(f[k]), m.device) for k in
             ("stress_t", "stress_s", "amts", "amuv", "amuv2d")}
        try:
            for visc2, dift in ((40.0, 10.0), (20.0, 5.0)):   # step-stable viscosities
                want = oracle.smagorinsky(og, f, visc2, dift)
                m.smagorinsky_dev(f["dz"], d["stress_t"], d["stress_s"], d["amts"], d["amuv"], d["amuv2d"],
                                  visc2, dift)
                got = {k: d[k].to_host() for k in FIELDS}
                assert_bitwise(got, want, FIELDS, f"smagorinsky device arrays visc2={visc2}")
        finally:
            for a in d.values():
                a.free()
    steps, nstep = 4, 2
    rd = oracle.OracleGrid(b, periodicity=(1, 0)).rayleigh()
    before = oracle.run(b, steps, nstep, out_steps=[steps], rdis=rd, periodicity=(1, 0))[steps]
    b.mu = np.where(og.masks["lu"] > 0.5, want["amuv2d"], b.mu)
    ref = oracle.run(b, steps, nstep, out_steps=[steps], rdis=rd, periodicity=(1, 0))[steps]
    assert not np.array_equal(ref["u"], before["u"])      # the viscosity reaches the step
    with BarotropicModel(b, rdis=rd, periodicity=(1, 0)) as m:
        m.model_steps(steps, nstep=nstep)
        st = m.get_state()
    assert_bitwise(st, ref, ("ssh", "u", "v"), "step with the Smagorinsky mu")


@pytest.mark.parametrize("name,make,periodicity,nz", CASES, ids=[c[0] for c in CASES])
def test_stress_components_matches_oracle(name, make, periodicity, nz):
    """stress_components of 3-D velocities (Function/mixing.f90:115-168) on
    the GPU against the oracle's level-by-level restatement, bit for bit;
    cells it does not write keep the caller's values."""
    print(f"3-D stress {name}", flush=True)
    b = make()
    rng = np.random.default_rng(nz + 100)
    s3 = (nz, b.ny, b.nx)
    u, v = 0.1 * rng.standard_normal(s3), 0.1 * rng.standard_normal(s3)
    st0, ss0 = rng.random(s3), rng.random(s3)
    og = oracle.OracleGrid(b, periodicity=periodicity)
    want_t, want_s = oracle.stress_components_3d(og, u, v, st0, ss0)
    with BarotropicModel(b, periodicity=periodicity) as m:
        got_t, got_s = m.stress_components(u, v, st0, ss0)
        assert m.stress_components_timed(nz, 2) > 0.0
    assert_bitwise({"str_t": got_t, "str_s": got_s}, {"str_t": want_t, "str_s": want_s}, ("str_t", "str_s"),
                   f"3-D stress {name}")
    assert np.any(got_t == st0) and np.any(got_t != st0)


def test_viscosity_chain_on_device():
    """Control/dyncall.f90:36-43 on HBM-resident arrays: the 3-D stress of
    the level n-1 velocities, then the Smagorinsky viscosity from it, equal
    the oracle's chain bit for bit (a stretched basin, 20 levels)."""
    from inmsom_bt.hipmem import DeviceArray
    b = make_basin(150, 110, seed=12, sigma=5.0, metrics="stretched")
    nz = 20
    rng = np.random.default_rng(12)
    s3, s2 = (nz, b.ny, b.nx), (b.ny, b.nx)
    dz = 0.5 + rng.random(nz)
    dz = dz / dz.sum()
    f = {"uup": 0.1 * rng.standard_normal(s3), "vvp": 0.1 * rng.standard_normal(s3), "stress_t": np.zeros(s3),
         "stress_s": np.zeros(s3), "amts": np.zeros(s3), "amuv": np.zeros(s3), "amuv2d": np.zeros(s2)}
    og = oracle.OracleGrid(b)
    st, ss = oracle.stress_components_3d(og, f["uup"], f["vvp"], f["stress_t"], f["stress_s"])
    want = oracle.smagorinsky(og, dict(f, dz=dz, stress_t=st, stress_s=ss), 1.0e3, 5.0e2)
    with BarotropicModel(b) as m:
        d = {k: DeviceArray.from_host(np.ascontiguousarray(f[k]), m.device) for k in f}
        try:
            m.stress_components_dev(nz, d["uup"], d["vvp"], d["stress_t"], d["stress_s"])
            m.smagorinsky_dev(dz, d["stress_t"], d["stress_s"], d["amts"], d["amuv"], d["amuv2d"], 1.0e3, 5.0e2)
            got = {k: d[k].to_host() for k in FIELDS}
        finally:
            for a in d.values():
                a.free()
    assert_bitwise(got, want, FIELDS, "stress -> Smagorinsky chain")
