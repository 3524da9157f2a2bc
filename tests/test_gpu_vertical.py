"""vertical_velocity on the GPU (Function/vel_ssh.f90:1030-1076) against
the oracle, bit for bit; the oracle is pinned to the compiled reference
(harness mode 5) and by the continuity balance."""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise
from inmsom_bt import BarotropicModel, make_basin, randomize_state

pytestmark = pytest.mark.gpu


def _inputs(b, nz, seed):
    rng = np.random.default_rng(seed)
    sh = (b.ny, b.nx)
    dz = 0.5 + rng.random(nz)
    return {"dz": dz / dz.sum(), "uu": 0.1 * rng.standard_normal((nz,) + sh),
            "vv": 0.1 * rng.standard_normal((nz,) + sh), "ww": 3.0 + rng.standard_normal((nz + 1,) + sh),
            "hu": 40.0 + rng.random(sh), "hv": 40.0 + rng.random(sh), "wflux": 1e-5 * rng.standard_normal(sh)}


def _channel(nxi, nyi, seed, all_wet=False):
    b = make_basin(nxi, nyi, seed=seed, sigma=3.0, metrics="stretched")
    b.lbasins[:, :] = 0 if all_wet else 1
    b.lbasins[3:-3, :] = 0
    return b


CASES = [("closed stretched nz=20", lambda: randomize_state(make_basin(120, 88, seed=4, sigma=5.0,
                                                                       metrics="stretched"), 5), (0, 0), 20),
         ("channel px nz=7", lambda: _channel(90, 40, 6), (1, 0), 7),
         ("doubly periodic nz=1", lambda: _channel(40, 30, 7, all_wet=True), (1, 1), 1),
         ("largest nz=1000", lambda: make_basin(24, 18, seed=8, sigma=3.0), (0, 0), 1000)]


@pytest.mark.parametrize("name,make,periodicity,nz", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("ffs", [1, 0])
def test_vertical_velocity_matches_oracle(name, make, periodicity, nz, ffs):
    print(f"vertical velocity {name} ffs={ffs}", flush=True)
    b = make()
    f = _inputs(b, nz, seed=nz + ffs)
    og = oracle.OracleGrid(b, periodicity=periodicity, full_free_surface=ffs)
    want = oracle.vertical_velocity(og, f)
    with BarotropicModel(b, periodicity=periodicity, full_free_surface=ffs) as m:
        got = m.vertical_velocity(f["dz"], f["uu"], f["vv"], f["ww"], f["hu"], f["hv"], f["wflux"])
        assert m.vertical_velocity_timed(nz, 2) > 0.0
    assert_bitwise({"ww": got}, {"ww": want}, ("ww",), f"vertical velocity {name}")
    assert np.any(got == f["ww"]) and np.any(got != f["ww"])     # untouched cells keep their values


def test_vertical_velocity_device_arrays():
    from inmsom_bt.hipmem import DeviceArray
    b = _channel(70, 34, 8)
    nz = 20
    f = _inputs(b, nz, seed=3)
    og = oracle.OracleGrid(b, periodicity=(1, 0))
    want = oracle.vertical_velocity(og, f)
    with BarotropicModel(b, periodicity=(1, 0)) as m:
        d = {k: DeviceArray.from_host(np.ascontiguousarray(f[k]), m.device) for k in
             ("uu", "vv", "ww", "hu", "hv", "wflux")}
        try:
            m.vertical_velocity_dev(f["dz"], d["uu"], d["vv"], d["ww"], d["hu"], d["hv"], d["wflux"])
            got = d["ww"].to_host()
        finally:
            for a in d.values():
                a.free()
    assert_bitwise({"ww": got}, {"ww": want}, ("ww",), "vertical velocity device arrays")
