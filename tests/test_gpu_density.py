"""The density definition on the GPU (Control/dyncall.f90:437-449 with the
equation of state of Service/oc_algorithms.f90:138-161) against the oracle,
bit for bit; the oracle is pinned to the reference's own density function
(harness mode 7) and by special cases of the polynomial.  The last tests
chain it into pressure_gradients, whose den it produces."""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise
from inmsom_bt import BarotropicModel, make_basin, randomize_state

pytestmark = pytest.mark.gpu


def _inputs(b, nz, seed):
    rng = np.random.default_rng(seed)
    sh = (b.ny, b.nx)
    s3 = (nz,) + sh
    return {"z": np.sort(rng.random(nz)), "tt": 2.0 + 25.0 * rng.random(s3), "ss": 5.0 + 30.0 * rng.random(s3),
            "hhq": 10.0 + 90.0 * rng.random(sh), "den": np.full(s3, 3.0), "den_pot": np.full(s3, -3.0)}


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
def test_density_matches_oracle(name, make, periodicity, nz):
    print(f"density {name}", flush=True)
    b = make()
    f = _inputs(b, nz, seed=nz)
    og = oracle.OracleGrid(b, periodicity=periodicity)
    want = oracle.density(og, f)
    with BarotropicModel(b, periodicity=periodicity) as m:
        den, dpot = m.density(f["z"], f["tt"], f["ss"], f["hhq"], f["den"], f["den_pot"])
        f1 = _inputs(b, 2, seed=nz + 40)
        want1 = oracle.density(og, f1)
        den1, dpot1 = m.density(f1["z"], f1["tt"], f1["ss"], f1["hhq"], f1["den"], f1["den_pot"])
        assert m.density_timed(2, 2) > 0.0
    assert_bitwise({"den": den, "den_pot": dpot}, want, ("den", "den_pot"), f"density {name}")
    assert_bitwise({"den": den1, "den_pot": dpot1}, want1, ("den", "den_pot"), "density, second call nz=2")
    assert np.any(den == 3.0) and np.any(den != 3.0)        # cells the reference does not write keep theirs


def test_density_device_arrays_feed_pressure_gradients():
    """density_dev then pressure_gradients_dev on HBM arrays, against the
    oracle chain, bit for bit."""
    from inmsom_bt.hipmem import DeviceArray
    b = randomize_state(make_basin(70, 54, seed=9, sigma=4.0, metrics="stretched"), 2)
    nz = 20
    f = _inputs(b, nz, seed=3)
    rng = np.random.default_rng(4)
    sh = (b.ny, b.nx)
    hu, hv = 30.0 + rng.random(sh), 30.0 + rng.random(sh)
    rx, ry = rng.standard_normal((nz,) + sh), rng.standard_normal((nz,) + sh)
    og = oracle.OracleGrid(b)
    eos = oracle.density(og, f)
    want = oracle.pressure_gradients(og, {"z": f["z"], "den": eos["den"], "rhsx": rx, "rhsy": ry, "hhq": f["hhq"],
                                          "hhu": hu, "hhv": hv})
    with BarotropicModel(b) as m:
        host = {"tt": f["tt"], "ss": f["ss"], "hhq": f["hhq"], "den": f["den"], "den_pot": f["den_pot"],
                "hhu": hu, "hhv": hv, "rhsx": rx, "rhsy": ry}
        d = {k: DeviceArray.from_host(np.ascontiguousarray(v), m.device) for k, v in host.items()}
        try:
            m.density_dev(f["z"], d["tt"], d["ss"], d["hhq"], d["den"], d["den_pot"])
            m.pressure_gradients_dev(f["z"], d["den"], d["rhsx"], d["rhsy"], d["hhq"], d["hhu"], d["hhv"])
            got = {k: d[k].to_host() for k in ("den", "den_pot", "rhsx", "rhsy")}
        finally:
            for a in d.values():
                a.free()
    assert_bitwise(got, eos, ("den", "den_pot"), "density device arrays")
    assert_bitwise(got, want, ("rhsx", "rhsy"), "pressure gradients of the device density")


def test_density_argument_errors():
    from inmsom_bt._capi import BtError, ESTATE
    b = make_basin(30, 20, seed=1)
    f = _inputs(b, 3, seed=1)
    with BarotropicModel(b) as m:
        with pytest.raises(ValueError):
            m.density(f["z"][:2], f["tt"], f["ss"], f["hhq"], f["den"], f["den_pot"])
        with pytest.raises(ValueError):
            m.density(f["z"], f["tt"], f["ss"][:2], f["hhq"], f["den"], f["den_pot"])
        with pytest.raises(BtError) as e:
            m.density_timed(4, 1)
        assert e.value.code == ESTATE
