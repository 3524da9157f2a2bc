"""The 2-D -> 3-D coupling on the GPU (Control/dyncall.f90:169-231 with
ssh_internal, Function/vel_ssh.f90:682-751) against the oracle, bit for
bit; the oracle itself is pinned to the compiled reference harness."""
import numpy as np
import pytest

import oracle
from _util import assert_bitwise
from inmsom_bt import BarotropicModel, make_basin, randomize_state

pytestmark = pytest.mark.gpu

OUT = ("uu", "vv", "uu2d", "vv2d", "pgrx", "pgry", "ssh_i", "sshp_i", "hhqn", "hhun", "hhvn", "hhhn")
ACC = ("ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")


def _inputs(b, nz, seed):
    rng = np.random.default_rng(seed)
    shape = (b.ny, b.nx)
    wet = b.wet.astype(np.float64)

    def f(scale):
        return rng.standard_normal(shape) * scale
    return {"ubrtr_i": f(5.0), "vbrtr_i": f(5.0), "hhu": 50.0 + np.abs(f(10.0)), "hhv": 50.0 + np.abs(f(10.0)),
            "slpr": 1.0e5 + f(100.0), "wflux": f(1e-6) * wet, "uu2d": f(0.1), "vv2d": f(0.1), "pgrx": f(1.0),
            "pgry": f(1.0), "ssh_i": f(0.2) * wet, "sshp_i": f(0.2) * wet,
            "uu": rng.standard_normal((nz,) + shape), "vv": rng.standard_normal((nz,) + shape)}


def _device(m, tau, inp):
    uu, vv = inp["uu"].copy(), inp["vv"].copy()
    m.remove_barotropic(uu, vv, inp["uu2d"], inp["vv2d"])
    z = np.zeros_like(inp["ssh_i"])
    got = m.internal_update(tau, dict(inp, hhqn=z, hhun=z, hhvn=z, hhhn=z), uu, vv)
    got["uu"], got["vv"] = uu, vv
    return got


@pytest.mark.parametrize("nz,periodicity,metrics,ffs", [(1, (0, 0), "uniform", 1), (20, (0, 0), "stretched", 1),
                                                         (5, (1, 0), "uniform", 1), (3, (1, 1), "stretched", 1),
                                                         (4, (0, 0), "uniform", 0), (400, (1, 0), "stretched", 1)])
def test_coupling_matches_oracle(nz, periodicity, metrics, ffs):
    print(f"coupling nz={nz} periodicity={periodicity} {metrics} ffs={ffs}", flush=True)
    b = randomize_state(make_basin(70, 50, seed=31 + nz, sigma=4.0, metrics=metrics), 5)
    inp = _inputs(b, nz, seed=nz)
    want = oracle.coupling(oracle.OracleGrid(b, periodicity=periodicity, full_free_surface=ffs), b.dt, inp)
    with BarotropicModel(b, periodicity=periodicity, full_free_surface=ffs) as m:
        got = _device(m, b.dt, inp)
        # a second call with fewer levels reuses the staging buffer
        inp2 = _inputs(b, 1, seed=nz + 100)
        want2 = oracle.coupling(oracle.OracleGrid(b, periodicity=periodicity, full_free_surface=ffs), b.dt, inp2)
        got2 = _device(m, b.dt, inp2)
    assert_bitwise(got, want, OUT, f"coupling nz={nz} {periodicity} {metrics} ffs={ffs}")
    assert_bitwise(got2, want2, OUT, "coupling, second call nz=1")
    if ffs:
        assert np.abs(got["hhhn"]).max() > 0.0
    else:
        assert not np.abs(got["hhhn"]).any()


def test_coupling_after_dropin_step():
    """The reference's order: barotropic_dynamics (the drop-in) produces the
    time-mean transports and the barotropic pressure gradient, then the
    coupling consumes them; both on the device, both against the oracle."""
    b = randomize_state(make_basin(64, 48, seed=41, sigma=4.0), 6)
    og = oracle.OracleGrid(b)
    rd = og.rayleigh()
    rng = np.random.default_rng(3)
    rhsx = rng.standard_normal((b.ny, b.nx)) * 1e3
    rhsy = rng.standard_normal((b.ny, b.nx)) * 1e3
    st_o = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
    st_g = {k: a.copy() for k, a in st_o.items()}
    acc_o = {k: np.zeros((b.ny, b.nx)) for k in ACC}
    acc_g = {k: np.zeros((b.ny, b.nx)) for k in ACC}
    oracle.barotropic_dynamics(og, b.dt, 2, 0, st_o, acc_o, rhsx, rhsy, b.wflux, b.mu, b.mu4, rd)
    inp = _inputs(b, 4, seed=9)
    inp.update(ubrtr_i=acc_o["ubrtr_i"], vbrtr_i=acc_o["vbrtr_i"], pgrx=acc_o["ssh4gradx"],
               pgry=acc_o["ssh4grady"])
    want = oracle.coupling(og, b.dt, inp)
    with BarotropicModel(b, rdis=rd) as m:
        m.barotropic_dynamics(b.dt, 2, 0, st_g["u"], st_g["up"], st_g["v"], st_g["vp"], st_g["ssh"],
                              st_g["sshp"], acc_g["ubrtr_i"], acc_g["vbrtr_i"], acc_g["ssh4gradx"],
                              acc_g["ssh4grady"], rhsx, rhsy, np.ascontiguousarray(b.wflux),
                              np.ascontiguousarray(b.mu), np.ascontiguousarray(b.mu4), rd)
        assert_bitwise(acc_g, acc_o, ACC, "drop-in time means")
        ginp = dict(inp, ubrtr_i=acc_g["ubrtr_i"], vbrtr_i=acc_g["vbrtr_i"], pgrx=acc_g["ssh4gradx"],
                    pgry=acc_g["ssh4grady"])
        got = _device(m, b.dt, ginp)
    assert_bitwise(got, want, OUT, "coupling after the drop-in step")


def test_coupling_timed_needs_staged_fields():
    from inmsom_bt._capi import BtError
    b = make_basin(40, 30, seed=44, sigma=4.0)
    with BarotropicModel(b) as m:
        with pytest.raises(BtError, match="inmsom_bt_internal_update first"):
            m.coupling_timed(b.dt, 3, 2)
        inp = _inputs(b, 3, seed=1)
        got = _device(m, b.dt, inp)
        assert np.isfinite(got["ssh_i"]).all()
        assert m.coupling_timed(b.dt, 3, 2) > 0.0
        assert m.coupling_timed(b.dt, 2, 1) > 0.0     # fewer levels fit the staging
        with pytest.raises(BtError, match="staged"):
            m.coupling_timed(b.dt, 4, 1)


@pytest.mark.parametrize("periodicity", [(0, 0), (1, 1)])
def test_coupling_device_pointers_match_oracle(periodicity):
    """The device-address entry points on arrays already in HBM give the
    oracle's result bit for bit, like the host-array ones."""
    from inmsom_bt.hipmem import DeviceArray
    nz = 6
    print(f"coupling on device arrays, periodicity={periodicity}", flush=True)
    b = randomize_state(make_basin(66, 52, seed=45, sigma=4.0, metrics="stretched"), 8)
    inp = _inputs(b, nz, seed=12)
    want = oracle.coupling(oracle.OracleGrid(b, periodicity=periodicity), b.dt, inp)
    z = np.zeros_like(inp["ssh_i"])
    host = dict(inp, hhqn=z, hhun=z, hhvn=z, hhhn=z)
    with BarotropicModel(b, periodicity=periodicity) as m:
        dev = {k: DeviceArray.from_host(a) for k, a in host.items()}
        m.remove_barotropic_dev(dev["uu"], dev["vv"], dev["uu2d"], dev["vv2d"])
        m.internal_update_dev(b.dt, dev, dev["uu"], dev["vv"])
        with pytest.raises(ValueError):
            m.remove_barotropic_dev(dev["uu"], dev["vv"], dev["uu"], dev["vv2d"])
        got = {k: dev[k].to_host() for k in OUT}
        for a in dev.values():
            a.free()
    assert_bitwise(got, want, OUT, f"coupling on device arrays {periodicity}")
