"""The barotropic_dynamics entry that allocates no device memory
(inmsom_bt_barotropic_dynamics_noalloc, scratch in a fixed array of the
library) against the device-address entry it varies, bit for bit, on two
calls in a row on two contexts, the second with a smaller grid, each call
after the entry has already been used (the scratch keeps the previous
call's contents); a grid beyond the scratch is refused with an error."""
import numpy as np
import pytest

import oracle
from _util import STATE, assert_bitwise
from inmsom_bt import BarotropicModel, make_basin, randomize_state
from inmsom_bt._capi import BtError
from inmsom_bt.hipmem import DeviceArray

pytestmark = pytest.mark.gpu

ACC = ("ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")


def _arrays(b, seed):
    rng = np.random.default_rng(seed)
    st = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
    inp = {"RHSx": rng.standard_normal((b.ny, b.nx)) * 1e3, "RHSy": rng.standard_normal((b.ny, b.nx)) * 1e3,
           "wflux": np.ascontiguousarray(b.wflux), "mu": np.ascontiguousarray(b.mu)}
    return st, inp


def _run(m, b, st, inp, rdis, nstep, entry):
    d = {k: DeviceArray.from_host(v) for k, v in st.items()}
    da = {k: DeviceArray.from_host(np.full((b.ny, b.nx), -3.0)) for k in ACC}
    di = {k: DeviceArray.from_host(v) for k, v in inp.items()}
    dr = DeviceArray.from_host(rdis)
    args = (b.dt, nstep, d["u"], d["up"], d["v"], d["vp"], d["ssh"], d["sshp"], da["ubrtr_i"], da["vbrtr_i"],
            da["ssh4gradx"], da["ssh4grady"], di["RHSx"], di["RHSy"], di["wflux"], di["mu"])
    if entry == "noalloc":
        m.barotropic_dynamics_noalloc(*args, dr)
        m.synchronize()
    else:
        mu4 = DeviceArray.from_host(np.zeros((b.ny, b.nx)))
        m.barotropic_dynamics_dev(args[0], args[1], 0, *args[2:], mu4, dr)
    return {**{k: v.to_host() for k, v in d.items()}, **{k: v.to_host() for k, v in da.items()}}


def test_noalloc_entry_equals_device_entry():
    cases = [(randomize_state(make_basin(700, 420, seed=61, sigma=6.0, metrics="stretched"), 3, wflux_amp=1e-6), 3),
             (randomize_state(make_basin(150, 97, seed=62, sigma=4.0), 4, wflux_amp=1e-6), 2)]
    for i, (b, nstep) in enumerate(cases):
        print(f"no-alloc call {i}: {b.nx}x{b.ny}, nstep={nstep}", flush=True)
        rdis = oracle.OracleGrid(b).rayleigh()
        st, inp = _arrays(b, 70 + i)
        with BarotropicModel(b, rdis=rdis) as m:
            assert m.dropin_path(0) == "fused"
            want = _run(m, b, st, inp, rdis, nstep, "dev")
            got = _run(m, b, st, inp, rdis, nstep, "noalloc")
        assert_bitwise(got, want, STATE + ACC, f"call {i} ({b.nx}x{b.ny})")


def test_noalloc_entry_refuses_grids_beyond_its_scratch():
    b = make_basin(1900, 1900, seed=63, sigma=8.0)  # 1904 x 1904 > 3,595,117 cells
    print(f"no-alloc refusal: {b.nx}x{b.ny}", flush=True)
    rdis = np.zeros((b.ny, b.nx))
    st, inp = _arrays(b, 79)
    with BarotropicModel(b, rdis=rdis) as m:
        with pytest.raises(BtError, match="too large"):
            _run(m, b, st, inp, rdis, 1, "noalloc")
