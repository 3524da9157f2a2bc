"""The barotropic step on periodic grids (Service/oc_algorithms.f90:50-135
applied as a decomposed run applies it) against the oracle, bit for bit:
the stage-by-stage path over several outer steps (state, accumulators
ubrtr_i / vbrtr_i / ssh4grad*, the 12 depth arrays) and the drop-in
barotropic_dynamics with caller-supplied forcing, on a zonal channel and on
a doubly periodic ocean whose bathymetry, state and water flux vary along
x.  The oracle's wrap is pinned by the doubled-domain test of
test_oracle_invariants.py."""
import numpy as np
import pytest

import oracle
from _util import STATE, assert_bitwise, periodic_pair
from inmsom_bt import BarotropicModel

pytestmark = pytest.mark.gpu

ACC = ("ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")
DEPTHS = ("hhq_e", "hhqp_e", "hhqn_e", "hhu_e", "hhup_e", "hhun_e",
          "hhv_e", "hhvp_e", "hhvn_e", "hhh_e", "hhhp_e", "hhhn_e")


@pytest.mark.parametrize("periodicity,nstep", [((1, 0), 2), ((1, 1), 3)])
def test_periodic_model_steps_bitwise(periodicity, nstep):
    b, _, _ = periodic_pair(periodicity)
    steps = 15
    print(f"periodic {periodicity}: {b.nx}x{b.ny} nstep={nstep}", flush=True)
    og = oracle.OracleGrid(b, periodicity)
    rd = og.rayleigh()
    want = oracle.run(b, steps, nstep, out_steps=[steps], rdis=rd, periodicity=periodicity)[steps]
    with BarotropicModel(b, rdis=rd, periodicity=periodicity) as m:
        m.model_steps(steps, nstep=nstep)
        got = m.get_state()
        got.update(m.means())
        got.update(m.depths())
    assert_bitwise(got, want, STATE + ACC + DEPTHS, f"periodic {periodicity} model_steps")


@pytest.mark.parametrize("periodicity", [(1, 0), (1, 1)])
def test_periodic_dropin_bitwise(periodicity):
    b, _, _ = periodic_pair(periodicity, nxi=57, nyi=33, seed=14)
    nstep = 2
    print(f"periodic drop-in {periodicity}: {b.nx}x{b.ny}", flush=True)
    og = oracle.OracleGrid(b, periodicity)
    rdis = og.rayleigh()
    rng = np.random.default_rng(8)
    rhsx = rng.standard_normal((b.ny, b.nx)) * 1e3
    rhsy = rng.standard_normal((b.ny, b.nx)) * 1e3
    st_o = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
    st_g = {k: v.copy() for k, v in st_o.items()}
    acc_o = {k: np.zeros((b.ny, b.nx)) for k in ACC}
    acc_g = {k: np.zeros((b.ny, b.nx)) for k in ACC}
    dep_o = oracle.barotropic_dynamics(og, b.dt, nstep, 0, st_o, acc_o, rhsx, rhsy, b.wflux, b.mu, b.mu4, rdis)
    with BarotropicModel(b, rdis=rdis, periodicity=periodicity) as m:
        m.barotropic_dynamics(b.dt, nstep, 0, st_g["u"], st_g["up"], st_g["v"], st_g["vp"], st_g["ssh"],
                              st_g["sshp"], acc_g["ubrtr_i"], acc_g["vbrtr_i"], acc_g["ssh4gradx"],
                              acc_g["ssh4grady"], rhsx, rhsy, np.ascontiguousarray(b.wflux),
                              np.ascontiguousarray(b.mu), np.ascontiguousarray(b.mu4), rdis)
        dep_g = m.depths()
    got = {**st_g, **acc_g, **dep_g}
    want = {**st_o, **acc_o, **dep_o}
    assert_bitwise(got, want, STATE + ACC + DEPTHS, f"periodic {periodicity} drop-in")
