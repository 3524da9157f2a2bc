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
                                                       (6, (1, 0), "uniform", False), (5, (1, 1), "stretched", True)])
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
