"""pressure_gradients on the GPU (Function/vel_ssh.f90:320-376) against the
oracle, bit for bit; the oracle is pinned to the compiled reference (harness
mode 6) and by two hydrostatic closed forms.  The last test chains it into
the forcing of the barotropic step as the producer of pgx, pgy."""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise
from inmsom_bt import BarotropicModel, make_basin, randomize_state

pytestmark = pytest.mark.gpu


def _inputs(b, nz, seed):
    rng = np.random.default_rng(seed)
    sh = (b.ny, b.nx)
    z = np.sort(rng.random(nz))
    return {"z": z, "den": 1025.0 + 0.3 * np.arange(1, nz + 1)[:, None, None] + rng.standard_normal((nz,) + sh),
            "rhsx": rng.standard_normal((nz,) + sh), "rhsy": rng.standard_normal((nz,) + sh),
            "hhq": 30.0 + 20.0 * rng.random(sh), "hhu": 30.0 + rng.random(sh), "hhv": 30.0 + rng.random(sh)}


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
def test_pressure_gradients_matches_oracle(name, make, periodicity, nz):
    print(f"pressure gradients {name}", flush=True)
    b = make()
    f = _inputs(b, nz, seed=nz)
    og = oracle.OracleGrid(b, periodicity=periodicity)
    want = oracle.pressure_gradients(og, f)
    with BarotropicModel(b, periodicity=periodicity) as m:
        rx, ry = m.pressure_gradients(f["z"], f["den"], f["rhsx"], f["rhsy"], f["hhq"], f["hhu"], f["hhv"])
        # a second call with fewer levels reuses the staging
        f1 = _inputs(b, 2, seed=nz + 40)
        want1 = oracle.pressure_gradients(og, f1)
        rx1, ry1 = m.pressure_gradients(f1["z"], f1["den"], f1["rhsx"], f1["rhsy"], f1["hhq"], f1["hhu"], f1["hhv"])
        assert m.pressure_gradients_timed(2, 2) > 0.0
    assert_bitwise({"rhsx": rx, "rhsy": ry}, want, ("rhsx", "rhsy"), f"pressure gradients {name}")
    assert_bitwise({"rhsx": rx1, "rhsy": ry1}, want1, ("rhsx", "rhsy"), "pressure gradients, second call nz=2")
    for got, k in ((rx, "rhsx"), (ry, "rhsy")):
        assert np.any(got == f[k]) and np.any(got != f[k])     # cells off lcu / lcv keep their values


def test_pressure_gradients_device_arrays():
    from inmsom_bt.hipmem import DeviceArray
    b = _channel(70, 34, 8)
    nz = 20
    f = _inputs(b, nz, seed=3)
    og = oracle.OracleGrid(b, periodicity=(1, 0))
    want = oracle.pressure_gradients(og, f)
    with BarotropicModel(b, periodicity=(1, 0)) as m:
        d = {k: DeviceArray.from_host(np.ascontiguousarray(f[k]), m.device) for k in
             ("den", "rhsx", "rhsy", "hhq", "hhu", "hhv")}
        try:
            m.pressure_gradients_dev(f["z"], d["den"], d["rhsx"], d["rhsy"], d["hhq"], d["hhu"], d["hhv"])
            got = {"rhsx": d["rhsx"].to_host(), "rhsy": d["rhsy"].to_host()}
        finally:
            for a in d.values():
                a.free()
    assert_bitwise(got, want, ("rhsx", "rhsy"), "pressure gradients device arrays")


def test_pressure_gradients_feed_the_forcing():
    """pressure_gradients on zeroed right-hand sides gives the pgx, pgy of
    the forcing; the GPU chain equals the oracle chain bit for bit."""
    from test_gpu_forcing3d import _fields
    nz = 9
    b = randomize_state(make_basin(70, 54, seed=71, sigma=4.0, metrics="stretched"), 3)
    f = _fields(b, nz, seed=5)
    p = _inputs(b, nz, seed=6)
    zero = np.zeros((nz, b.ny, b.nx))
    og = oracle.OracleGrid(b)
    pg = oracle.pressure_gradients(og, dict(p, rhsx=zero, rhsy=zero, hhq=f["hhq"], hhu=f["hhu"], hhv=f["hhv"]))
    want = oracle.forcing_3d(og, dict(f, pgx=pg["rhsx"], pgy=pg["rhsy"]), b.cb_nl, b.ebottom)
    with BarotropicModel(b) as m:
        pgx, pgy = m.pressure_gradients(p["z"], p["den"], zero, zero, f["hhq"], f["hhu"], f["hhv"])
        got = m.forcing_3d(dict(f, taux=b.taux, tauy=b.tauy), f["dz"], f["rhsx2d"], f["rhsy2d"], pgx, pgy)
    assert_bitwise({"pgx": pgx, "pgy": pgy}, {"pgx": pg["rhsx"], "pgy": pg["rhsy"]}, ("pgx", "pgy"), "pgx, pgy")
    assert_bitwise(got, want, ("rhsx2d", "rhsy2d"), "forcing fed by pressure_gradients")
    assert np.abs(pgx).max() > 0.0


def test_pressure_gradients_argument_errors():
    from inmsom_bt._capi import BtError, ESTATE
    b = make_basin(30, 20, seed=1)
    f = _inputs(b, 3, seed=1)
    with BarotropicModel(b) as m:
        with pytest.raises(ValueError):
            m.pressure_gradients(f["z"][:2], f["den"], f["rhsx"], f["rhsy"], f["hhq"], f["hhu"], f["hhv"])
        with pytest.raises(ValueError):
            m.pressure_gradients(f["z"], f["den"], f["rhsx"][:2], f["rhsy"], f["hhq"], f["hhu"], f["hhv"])
        with pytest.raises(BtError) as e:
            m.pressure_gradients_timed(4, 1)
        assert e.value.code == ESTATE
