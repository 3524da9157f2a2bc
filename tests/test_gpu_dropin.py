"""The fused barotropic_dynamics (one kernel launch per substep, look-ahead
substeps and averages included) behind the reference's own boundary, bit
for bit against the oracle and against the stage-by-stage path: uniform,
stretched (every metric a different 2-D field) and longitude-latitude
metrics, random states, water flux and caller RHS, nstep 1..3, chained
calls, the biharmonic viscosity (ksw4 = 1), the shipped run's
nstep_barotrop = 50 (ocean_run.par:9) with ksw4 = 1 on the Black Sea grid,
the device-address entry, the full 2048 x 1536 basin, and the rule that
sends other calls (periodic, open halo) down the stage-by-stage path."""
import os

import numpy as np
import pytest

import oracle
from _util import STATE, assert_bitwise, periodic_pair
from inmsom_bt import BarotropicModel, make_basin, randomize_state
from inmsom_bt._capi import BtError
from inmsom_bt.cases import reference_basin
from inmsom_bt.hipmem import DeviceArray
from inmsom_bt.workloads import single_basin

pytestmark = pytest.mark.gpu

ACC = ("ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")
DEPTHS = ("hhq_e", "hhqp_e", "hhqn_e", "hhu_e", "hhup_e", "hhun_e",
          "hhv_e", "hhvp_e", "hhvn_e", "hhh_e", "hhhp_e", "hhhn_e")
BLACKSEA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "blacksea")


def _basin(kind):
    if kind == "uniform":
        return make_basin(150, 110, seed=21, sigma=6.0)
    if kind == "stretched":
        return randomize_state(make_basin(130, 97, seed=22, sigma=5.0, metrics="stretched"), 7, wflux_amp=1e-5)
    if kind == "spherical":
        b = reference_basin(os.path.join(BLACKSEA, "mask.txt"), BLACKSEA)
        return randomize_state(b, 3, amp_ssh=0.2, amp_uv=0.1, wflux_amp=1e-6)
    if kind == "ragged":
        return randomize_state(make_basin(7, 5, seed=23, sigma=1.0, hland=60.0), 8)
    if kind == "wide":
        return randomize_state(make_basin(301, 64, seed=24, sigma=5.0), 9)
    raise KeyError(kind)


def _inputs(b, seed):
    rng = np.random.default_rng(seed)
    return {"RHSx": rng.standard_normal((b.ny, b.nx)) * 1e3, "RHSy": rng.standard_normal((b.ny, b.nx)) * 1e3,
            "wflux": np.ascontiguousarray(b.wflux), "mu": np.ascontiguousarray(b.mu),
            "mu4": np.ascontiguousarray(b.mu4)}


def _call_model(m, b, nstep, st, acc, inp, rdis, ksw4=0):
    m.barotropic_dynamics(b.dt, nstep, ksw4, st["u"], st["up"], st["v"], st["vp"], st["ssh"], st["sshp"],
                          acc["ubrtr_i"], acc["vbrtr_i"], acc["ssh4gradx"], acc["ssh4grady"], inp["RHSx"],
                          inp["RHSy"], inp["wflux"], inp["mu"], inp["mu4"], rdis)


@pytest.mark.parametrize("kind,nstep,calls,ksw4", [("uniform", 2, 3, 0), ("stretched", 3, 2, 0), ("spherical", 2, 2, 0),
                                                   ("ragged", 1, 3, 0), ("wide", 2, 1, 0), ("uniform", 2, 2, 1),
                                                   ("stretched", 3, 2, 1), ("spherical", 2, 2, 1),
                                                   ("ragged", 1, 2, 1), ("spherical", 50, 1, 1)])
def test_fused_dropin_matches_oracle(kind, nstep, calls, ksw4):
    b = _basin(kind)
    og = oracle.OracleGrid(b)
    rdis = og.rayleigh()
    st_o = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
    st_g = {k: v.copy() for k, v in st_o.items()}
    with BarotropicModel(b, rdis=rdis) as m:
        assert m.dropin_path(ksw4) == "fused", kind
        for call in range(calls):
            print(f"fused drop-in {kind} {b.nx}x{b.ny} nstep={nstep} ksw4={ksw4} call {call}", flush=True)
            inp = _inputs(b, 100 + call)
            acc_o = {k: np.zeros((b.ny, b.nx)) for k in ACC}
            acc_g = {k: np.full((b.ny, b.nx), np.nan) for k in ACC}  # outputs only
            dep_o = oracle.barotropic_dynamics(og, b.dt, nstep, ksw4, st_o, acc_o, inp["RHSx"], inp["RHSy"],
                                               inp["wflux"], inp["mu"], inp["mu4"], rdis)
            _call_model(m, b, nstep, st_g, acc_g, inp, rdis, ksw4)
            got = {**st_g, **acc_g, **m.depths()}
            want = {**st_o, **acc_o, **dep_o}
            assert_bitwise(got, want, STATE + ACC + DEPTHS, f"{kind} ksw4={ksw4} call {call}")


@pytest.mark.parametrize("nstep,ksw4", [(1, 0), (2, 0), (3, 0), (4, 0), (1, 1), (3, 1)])
def test_fused_equals_stage_by_stage(nstep, ksw4):
    """Five outer steps of the 2-D model (wind stress and drag formed from
    each step's state, then barotropic_dynamics) through either path."""
    b = randomize_state(single_basin("cfg0"), 4, wflux_amp=1e-6)
    out = {}
    for path in ("literal", "fused"):
        print(f"256x256 model steps, nstep={nstep}, ksw4={ksw4}, {path}", flush=True)
        with BarotropicModel(b) as m:
            m.set_dropin_path(path)
            assert m.dropin_path(ksw4) == path
            m.model_steps(5, nstep=nstep, ksw4=ksw4)
            out[path] = {**m.get_state(), **m.means(), **m.depths()}
    assert_bitwise(out["fused"], out["literal"], STATE + ACC + DEPTHS, f"nstep={nstep} ksw4={ksw4}")


@pytest.mark.parametrize("path", ["fused", "literal"])
def test_device_address_entry(path):
    """inmsom_bt_barotropic_dynamics_dev on device arrays equals the host-array
    entry bit for bit, and leaves the context's own state and averages
    alone."""
    b = _basin("stretched")
    rdis = oracle.OracleGrid(b).rayleigh()
    inp = _inputs(b, 7)
    st_h = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
    acc_h = {k: np.zeros((b.ny, b.nx)) for k in ACC}
    print(f"device-address drop-in, {path}", flush=True)
    with BarotropicModel(b, rdis=rdis) as m:
        m.set_dropin_path(path)
        m.model_steps(1, nstep=2)  # non-zero averages of the context's own
        before = {**m.get_state(), **m.means()}
        d = {k: DeviceArray.from_host(v) for k, v in st_h.items()}
        da = {k: DeviceArray.from_host(np.full((b.ny, b.nx), 7.0)) for k in ACC}
        di = {k: DeviceArray.from_host(v) for k, v in inp.items()}
        dr = DeviceArray.from_host(rdis)
        m.barotropic_dynamics_dev(b.dt, 3, 0, d["u"], d["up"], d["v"], d["vp"], d["ssh"], d["sshp"],
                                  da["ubrtr_i"], da["vbrtr_i"], da["ssh4gradx"], da["ssh4grady"], di["RHSx"],
                                  di["RHSy"], di["wflux"], di["mu"], di["mu4"], dr)
        got = {**{k: v.to_host() for k, v in d.items()}, **{k: v.to_host() for k, v in da.items()}}
        assert_bitwise({**m.get_state(), **m.means()}, before, STATE + ACC, "context state untouched")
        _call_model(m, b, 3, st_h, acc_h, inp, rdis)
    assert_bitwise(got, {**st_h, **acc_h}, STATE + ACC, f"device vs host entry ({path})")


@pytest.mark.parametrize("ksw4", [0, 1])
def test_full_size_call_matches_oracle(ksw4):
    b = single_basin("cfg1")
    og = oracle.OracleGrid(b)
    rdis = og.rayleigh()
    inp = _inputs(b, 11)
    st_o = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
    st_g = {k: v.copy() for k, v in st_o.items()}
    acc_o = {k: np.zeros((b.ny, b.nx)) for k in ACC}
    acc_g = {k: np.zeros((b.ny, b.nx)) for k in ACC}
    print(f"2048x1536 drop-in call, nstep=2, ksw4={ksw4}", flush=True)
    oracle.barotropic_dynamics(og, b.dt, 2, ksw4, st_o, acc_o, inp["RHSx"], inp["RHSy"], inp["wflux"], inp["mu"],
                               inp["mu4"], rdis)
    with BarotropicModel(b, rdis=rdis) as m:
        assert m.dropin_path(ksw4) == "fused"
        _call_model(m, b, 2, st_g, acc_g, inp, rdis, ksw4)
    assert_bitwise({**st_g, **acc_g}, {**st_o, **acc_o}, STATE + ACC, f"2048x1536 ksw4={ksw4}")


def test_routing():
    """Calls the fused kernel does not cover take the stage-by-stage path,
    and asking for the fused path then fails with a clear error."""
    b = make_basin(40, 30, seed=5, sigma=3.0)
    with BarotropicModel(b) as m:
        assert m.dropin_path(0) == "fused"
        assert m.dropin_path(1) == "fused"  # biharmonic viscosity: pre-pass per substep
        m.set_dropin_path("literal")
        assert m.dropin_path(0) == "literal"
    pb, _, _ = periodic_pair((1, 0))
    with BarotropicModel(pb, periodicity=(1, 0)) as m:
        assert m.dropin_path(0) == "literal"
        assert m.dropin_path(1) == "literal"
        m.set_dropin_path("fused")
        with pytest.raises(BtError, match="periodic"):
            m.model_steps(1, nstep=2)
    ob = make_basin(40, 30, seed=6, sigma=3.0)
    ob.lbasins[:, :] = 0  # ocean on the halo: the loop bounds differ from the masks
    with BarotropicModel(ob) as m:
        assert m.dropin_path(0) == "literal"
