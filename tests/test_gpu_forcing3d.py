"""The forcing of the barotropic step from the 3-D model on the GPU
(Control/dyncall.f90:37-130) against the oracle, bit for bit; the oracle
is pinned to the compiled reference harness (mode 3) and by a closed form."""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise
from inmsom_bt import BarotropicModel, make_basin, randomize_state

pytestmark = pytest.mark.gpu


def _fields(b, nz, seed):
    rng = np.random.default_rng(seed)
    sh = (b.ny, b.nx)
    og = oracle.OracleGrid(b)
    lcu = (og.masks["lcu"] > 0.5).astype(np.float64)
    lcv = (og.masks["lcv"] > 0.5).astype(np.float64)
    f = {"uu": rng.standard_normal((nz,) + sh) * 0.1 * lcu, "vv": rng.standard_normal((nz,) + sh) * 0.1 * lcv,
         "uup": rng.standard_normal((nz,) + sh) * 0.1 * lcu, "vvp": rng.standard_normal((nz,) + sh) * 0.1 * lcv,
         "amuv": 10.0 + np.abs(rng.standard_normal((nz,) + sh)),
         "uu2d": rng.standard_normal(sh) * 0.1 * lcu, "vv2d": rng.standard_normal(sh) * 0.1 * lcv,
         "uup2d": rng.standard_normal(sh) * 0.1 * lcu, "vvp2d": rng.standard_normal(sh) * 0.1 * lcv,
         "amuv2d": 10.0 + np.abs(rng.standard_normal(sh)),
         "hhq": 50.0 + np.abs(rng.standard_normal(sh)), "hhu": 50.0 + np.abs(rng.standard_normal(sh)),
         "hhv": 50.0 + np.abs(rng.standard_normal(sh)), "hhh": 50.0 + np.abs(rng.standard_normal(sh)),
         "slpr": 1.0e5 + 100.0 * rng.standard_normal(sh),
         "pgx": rng.standard_normal((nz,) + sh) * 1e-3, "pgy": rng.standard_normal((nz,) + sh) * 1e-3,
         "rhsx2d": np.full(sh, 7.0), "rhsy2d": np.full(sh, -3.0)}
    w = np.arange(1, nz + 1) + 0.5
    f["dz"] = w / w.sum()
    return f


@pytest.mark.parametrize("nz,periodicity,metrics,pg", [(1, (0, 0), "uniform", True), (20, (0, 0), "stretched", True),
                                                       (6, (1, 0), "uniform", False), (5, (1, 1), "stretched", True),
                                                       (1000, (0, 0), "uniform", True)])
def test_forcing_3d_matches_oracle(nz, periodicity, metrics, pg):
    print(f"forcing 3-D nz={nz} periodicity={periodicity} {metrics} pg={pg}", flush=True)
    b = randomize_state(make_basin(70, 54, seed=60 + nz, sigma=4.0, metrics=metrics), 3)
    f = _fields(b, nz, seed=nz)
    og = oracle.OracleGrid(b, periodicity=periodicity)
    want = oracle.forcing_3d(og, f, b.cb_nl, b.ebottom, pressure=pg)
    with BarotropicModel(b, periodicity=periodicity) as m:
        fld = dict(f, taux=b.taux, tauy=b.tauy)
        got = m.forcing_3d(fld, f["dz"], f["rhsx2d"], f["rhsy2d"], *((f["pgx"], f["pgy"]) if pg else (None, None)))
        # a second call with fewer levels reuses the staging
        f1 = _fields(b, 1, seed=nz + 50)
        want1 = oracle.forcing_3d(og, f1, b.cb_nl, b.ebottom)
        got1 = m.forcing_3d(dict(f1, taux=b.taux, tauy=b.tauy), f1["dz"], f1["rhsx2d"], f1["rhsy2d"], f1["pgx"],
                            f1["pgy"])
        assert m.forcing_3d_timed(1, 2) > 0.0
    assert_bitwise(got, want, ("rhsx2d", "rhsy2d"), f"forcing 3-D nz={nz} {periodicity} {metrics}")
    assert_bitwise(got1, want1, ("rhsx2d", "rhsy2d"), "forcing 3-D, second call nz=1")
    assert np.abs(got["rhsx2d"] - 7.0).max() > 0.0 and (got["rhsx2d"] == 7.0).any()


@pytest.mark.parametrize("periodicity,pg", [((0, 0), True), ((1, 1), False)])
def test_forcing_3d_device_arrays(periodicity, pg):
    """The device-address entry on HBM arrays: rhsx2d, rhsy2d updated in place,
    bit for bit the oracle's."""
    from inmsom_bt.hipmem import DeviceArray
    print(f"forcing 3-D device arrays periodicity={periodicity} pg={pg}", flush=True)
    nz = 7
    b = randomize_state(make_basin(66, 58, seed=5, sigma=4.0, metrics="stretched"), 4)
    f = _fields(b, nz, seed=3)
    og = oracle.OracleGrid(b, periodicity=periodicity)
    want = oracle.forcing_3d(og, f, b.cb_nl, b.ebottom, pressure=pg)
    with BarotropicModel(b, periodicity=periodicity) as m:
        host = dict(f, taux=b.taux, tauy=b.tauy)
        keys = m.FORCING_3D + m.FORCING_2D
        dev = {k: DeviceArray.from_host(np.ascontiguousarray(host[k], dtype=np.float64), m.device) for k in keys}
        rx = DeviceArray.from_host(f["rhsx2d"], m.device)
        ry = DeviceArray.from_host(f["rhsy2d"], m.device)
        pgd = [DeviceArray.from_host(f[k], m.device) for k in ("pgx", "pgy")] if pg else [None, None]
        try:
            m.forcing_3d_dev(dev, f["dz"], rx, ry, *pgd)
            got = {"rhsx2d": rx.to_host(), "rhsy2d": ry.to_host()}
        finally:
            for a in list(dev.values()) + [rx, ry] + [q for q in pgd if q is not None]:
                a.free()
    assert_bitwise(got, want, ("rhsx2d", "rhsy2d"), f"device-array forcing {periodicity}")


def test_forcing_3d_argument_errors():
    from inmsom_bt._capi import BtError, ESTATE, EINVAL
    b = make_basin(30, 24, seed=2)
    f = _fields(b, 3, seed=1)
    with BarotropicModel(b) as m:
        with pytest.raises(BtError) as e:
            m.forcing_3d_timed(3, 1)            # nothing staged yet
        assert e.value.code == ESTATE
        fld = dict(f, taux=b.taux, tauy=b.tauy)
        with pytest.raises(BtError) as e:
            m.forcing_3d(fld, f["dz"], f["rhsx2d"], f["rhsy2d"], f["pgx"], None)
        assert e.value.code == EINVAL
        m.forcing_3d(fld, f["dz"], f["rhsx2d"], f["rhsy2d"])
        with pytest.raises(BtError) as e:
            m.forcing_3d_timed(4, 1)            # staged with 3 levels
        assert e.value.code == ESTATE
        assert m.forcing_3d_timed(3, 1) > 0.0
