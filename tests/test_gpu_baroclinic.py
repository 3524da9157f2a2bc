"""baroclinic_dynamics on the GPU (Function/vel_ssh.f90:794-1027 with
factor8) against the oracle, bit for bit; the oracle is pinned to the
compiled reference (harness mode 4) and by invariants."""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise
from inmsom_bt import BarotropicModel, make_basin, randomize_state

pytestmark = pytest.mark.gpu

OUT = ("uu", "uup", "vv", "vvp")


def _fields(b, nz, seed):
    rng = np.random.default_rng(seed)
    sh = (b.ny, b.nx)
    f = {"uu": 0.1 * rng.standard_normal((nz,) + sh), "uup": 0.1 * rng.standard_normal((nz,) + sh),
         "vv": 0.1 * rng.standard_normal((nz,) + sh), "vvp": 0.1 * rng.standard_normal((nz,) + sh),
         "ww": 1e-4 * rng.standard_normal((nz + 1,) + sh), "nu": 1e-3 + 1e-2 * rng.random((nz + 1,) + sh)}
    for k in ("rhsx", "rhsy", "rhsx_tran", "rhsy_tran"):
        f[k] = 1e3 * rng.standard_normal((nz,) + sh)
    for k in ("pgrx", "pgry"):
        f[k] = 1e3 * rng.standard_normal(sh)
    f["rdis"] = 1e-5 * rng.random(sh)
    for k in ("tx_surf", "ty_surf", "tx_bot", "ty_bot"):
        f[k] = 1e-4 * rng.standard_normal(sh)
    for k in ("hhu", "hhup", "hhun", "hhv", "hhvp", "hhvn", "hhh"):
        f[k] = 40.0 + rng.random(sh)
    dz = 0.5 + rng.random(nz)
    f["dz"] = dz / dz.sum()
    f["hzt"] = (0.5 + rng.random(nz + 1)) / nz
    return f


def _channel(nxi, nyi, seed, all_wet=False):
    """A zonal channel wet across the x boundary (land rows north and south),
    or an ocean wet everywhere (for periodicity in both directions)."""
    b = make_basin(nxi, nyi, seed=seed, sigma=3.0, metrics="stretched")
    b.lbasins[:, :] = 0 if all_wet else 1
    b.lbasins[3:-3, :] = 0
    b.hhq_rest = np.where(b.lbasins == 0, 60.0 + 10.0 * np.sin(np.arange(b.nx) / 5.0)[None, :], 0.0)
    return b


CASES = [("closed nz=20", lambda: randomize_state(make_basin(150, 90, seed=3, sigma=5.0), 4), (0, 0), 20),
         ("stretched nz=7", lambda: make_basin(77, 61, seed=5, sigma=4.0, metrics="stretched"), (0, 0), 7),
         ("channel px nz=20", lambda: _channel(96, 40, 6), (1, 0), 20),
         ("doubly periodic nz=3", lambda: _channel(40, 30, 7, all_wet=True), (1, 1), 3),
         ("largest nz=128", lambda: make_basin(30, 24, seed=8, sigma=3.0), (0, 0), 128)]


@pytest.mark.parametrize("name,make,periodicity,nz", CASES, ids=[c[0] for c in CASES])
def test_baroclinic_matches_oracle(name, make, periodicity, nz):
    print(f"baroclinic {name}", flush=True)
    b = make()
    f = _fields(b, nz, seed=nz)
    tau = 40.0
    og = oracle.OracleGrid(b, periodicity=periodicity)
    want = oracle.baroclinic(og, tau, f)
    with BarotropicModel(b, periodicity=periodicity) as m:
        got = m.baroclinic_dynamics(tau, f)
        # a second call with fewer levels reuses the staging
        f2 = _fields(b, 2, seed=99)
        want2 = oracle.baroclinic(og, tau, f2)
        got2 = m.baroclinic_dynamics(tau, f2)
        assert m.baroclinic_timed(tau, 2, 2) > 0.0
    assert_bitwise(got, want, OUT, f"baroclinic {name}")
    assert_bitwise(got2, want2, OUT, f"baroclinic {name}, second call nz=2")
    assert not np.array_equal(got["uu"], f["uu"])


def test_baroclinic_device_arrays():
    from inmsom_bt.hipmem import DeviceArray
    print("baroclinic device arrays", flush=True)
    b = _channel(70, 34, 8)
    nz, tau = 20, 40.0
    f = _fields(b, nz, seed=1)
    og = oracle.OracleGrid(b, periodicity=(1, 0))
    want = oracle.baroclinic(og, tau, f)
    with BarotropicModel(b, periodicity=(1, 0)) as m:
        dev = {k: DeviceArray.from_host(np.ascontiguousarray(v, dtype=np.float64), m.device)
               for k, v in f.items() if k not in ("dz", "hzt")}
        try:
            m.baroclinic_dynamics_dev(tau, dict(dev, dz=f["dz"], hzt=f["hzt"]))
            got = {k: dev[k].to_host() for k in OUT}
        finally:
            for a in dev.values():
                a.free()
    assert_bitwise(got, want, OUT, "baroclinic device arrays")


def test_baroclinic_argument_errors():
    from inmsom_bt._capi import BtError, EINVAL, ESTATE
    b = make_basin(20, 16, seed=2)
    f = _fields(b, 4, seed=2)
    with BarotropicModel(b) as m:
        with pytest.raises(BtError) as e:
            m.baroclinic_timed(40.0, 4, 1)
        assert e.value.code == ESTATE
        with pytest.raises(BtError) as e:
            m.baroclinic_dynamics(-1.0, f)
        assert e.value.code == EINVAL
        g = _fields(b, 1, seed=3)
        with pytest.raises(BtError) as e:
            m.baroclinic_dynamics(40.0, g)              # nz = 1: the reference's column needs two levels
        assert e.value.code == EINVAL
