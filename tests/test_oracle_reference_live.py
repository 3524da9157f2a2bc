"""Live comparison of the oracle with the compiled reference harness on
further random cases (runs where oracle/_ref has been built, i.e. where the
reference sources are mounted; skipped elsewhere)."""
import os
import tempfile

import numpy as np
import pytest

import oracle
from _util import assert_bitwise
from inmsom_bt.cases import make_basin, randomize_state, write_case

pytestmark = pytest.mark.skipif(not oracle.have_reference(),
                                reason="oracle/_ref not built (reference sources not mounted)")

FIELDS = ("ssh", "sshp", "u", "up", "v", "vp", "ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")


@pytest.mark.parametrize("seed,nstep,ksw4,metrics", [(201, 2, 0, "uniform"), (202, 4, 1, "stretched"),
                                                    (203, 1, 1, "uniform")])
def test_random_case_bitwise(seed, nstep, ksw4, metrics):
    b = make_basin(285, 159, seed=seed, sigma=4.0, g=9.8, hmax=120.0, metrics=metrics,
                   lvisc_4=1e-3 if ksw4 else 0.0)
    b = randomize_state(b, seed + 1, amp_uv=0.05, wflux_amp=1e-6)
    rd = oracle.OracleGrid(b).rayleigh()
    steps = [1, 5]
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, rd, 5, nstep, ksw4, 0, steps)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 0, steps)
    got = oracle.run(b, 5, nstep, ksw4, out_steps=steps, rdis=rd)
    for s in steps:
        assert_bitwise(got[s], ref[s], FIELDS + oracle.DEPTHS, f"seed {seed} step {s}")


@pytest.mark.parametrize("seed,metrics", [(211, "uniform"), (212, "stretched")])
def test_coupling_bitwise(seed, metrics):
    """2-D -> 3-D coupling (Control/dyncall.f90:169-231 with the reference's
    own ssh_internal) against the oracle on the inputs the harness wrote."""
    b = make_basin(285, 159, seed=seed, sigma=4.0, g=9.8, hmax=120.0, metrics=metrics)
    b = randomize_state(b, seed + 1, amp_uv=0.05, wflux_amp=1e-6)
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, rd, 1, 2, 0, 2)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 2)
    inp, want = ref["coupling_in"], ref["coupling_out"]
    assert inp["uu"].shape[0] == 20 and np.abs(inp["slpr"]).max() > 0
    got = oracle.coupling(oracle.OracleGrid(b), b.dt, inp)
    assert_bitwise(got, want, ("uu", "vv") + oracle.COUPLING_OUT, f"coupling seed {seed}")


@pytest.mark.parametrize("seed,metrics", [(7, "uniform"), (8, "stretched")])
def test_forcing_3d_bitwise(seed, metrics):
    """The forcing of the barotropic step (Control/dyncall.f90:37-130: the
    reference's stress_components, surface and bottom fluxes, uv_trans,
    uv_diff2 at nz levels and depth_ave, then the assembly) against the
    oracle on the inputs the harness wrote."""
    b = make_basin(285, 159, seed=seed, sigma=4.0, g=9.8, hmax=120.0, metrics=metrics)
    b = randomize_state(b, seed + 1, amp_uv=0.05)
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, rd, 1, 2, 0, 3)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 3)
    inp, want = ref["forcing_in"], ref["forcing_out"]
    assert inp["uu"].shape[0] == 20 and abs(inp["dz"].sum() - 1.0) < 1e-12
    got = oracle.forcing_3d(oracle.OracleGrid(b), inp, b.cb_nl, b.ebottom)
    assert_bitwise(got, want, ("rhsx2d", "rhsy2d"), f"forcing 3-D seed {seed} {metrics}")
    # cells the reference leaves alone keep the caller's values
    untouched = (want["rhsx2d"] == 7.0)
    assert untouched.any() and (got["rhsx2d"][untouched] == 7.0).all()


@pytest.mark.parametrize("seed,metrics", [(9, "uniform"), (10, "stretched")])
def test_baroclinic_bitwise(seed, metrics):
    """baroclinic_dynamics (Function/vel_ssh.f90:794-1027 with factor8,
    Service/oc_algorithms.f90:494-524) called by the harness on its
    synthetic 3-D fields, against the oracle on the inputs it wrote."""
    b = make_basin(285, 159, seed=seed, sigma=4.0, g=9.8, hmax=120.0, metrics=metrics)
    b = randomize_state(b, seed + 1, amp_uv=0.05)
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, rd, 1, 2, 0, 4)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 4)
    inp, want = ref["baroclinic_in"], ref["baroclinic_out"]
    assert inp["uu"].shape[0] == 20 and inp["ww"].shape[0] == 21
    got = oracle.baroclinic(oracle.OracleGrid(b), b.dt, inp)
    assert_bitwise(got, want, oracle.BAROCLINIC_OUT, f"baroclinic seed {seed} {metrics}")
    assert not np.array_equal(got["uu"], inp["uu"])


@pytest.mark.parametrize("seed,metrics", [(11, "uniform"), (12, "stretched")])
def test_vertical_velocity_bitwise(seed, metrics):
    """vertical_velocity (Function/vel_ssh.f90:1030-1076) called by the
    harness, against the oracle on the inputs it wrote; cells the routine
    does not compute keep their initial values."""
    b = make_basin(285, 159, seed=seed, sigma=4.0, g=9.8, hmax=120.0, metrics=metrics)
    b = randomize_state(b, seed + 1, amp_uv=0.05, wflux_amp=1e-5)
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, rd, 1, 2, 0, 5)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 5)
    inp, want = ref["vv_in"], ref["ww_out"]
    assert inp["uu"].shape[0] == 20 and want.shape[0] == 21
    got = oracle.vertical_velocity(oracle.OracleGrid(b), inp)
    assert_bitwise({"ww": got}, {"ww": want}, ("ww",), f"vertical velocity seed {seed} {metrics}")
    assert np.any(want != inp["ww"]) and np.any(want == inp["ww"])


@pytest.mark.parametrize("seed,metrics", [(41, "uniform"), (42, "stretched")])
def test_smagorinsky_bitwise(seed, metrics):
    """smagorinsky_coeff (Function/mixing.f90:171-219) and depth_ave of its
    viscosity (Function/depth.f90:223-254), as Control/dyncall.f90:40-43
    runs them, called by the harness on the reference's own stress
    components, against the oracle on the inputs it wrote; cells the
    routines do not compute keep their initial values."""
    b = make_basin(285, 159, seed=seed, sigma=4.0, g=9.8, hmax=120.0, metrics=metrics)
    b = randomize_state(b, seed + 1, amp_uv=0.05)
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, rd, 1, 2, 0, 10)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 10)
    inp, want = ref["smag_in"], ref["smag_out"]
    assert inp["amuv"].shape[0] == 20
    got = oracle.smagorinsky(oracle.OracleGrid(b), inp, 1.0e3, 5.0e2)
    assert_bitwise(got, want, ("amts", "amuv", "amuv2d"), f"smagorinsky seed {seed} {metrics}")
    for k in ("amts", "amuv", "amuv2d"):
        assert np.any(want[k] != inp[k]) and np.any(want[k] == inp[k]), k


@pytest.mark.parametrize("seed,metrics", [(41, "uniform"), (42, "stretched")])
def test_stress_components_3d_bitwise(seed, metrics):
    """stress_components (Function/mixing.f90:115-168) of 3-D velocities at
    the compiled nz, as Control/dyncall.f90:36 calls it before the
    Smagorinsky viscosity, called by the harness (mode 10) on zeroed
    stress arrays, against the oracle's level-by-level restatement."""
    b = make_basin(285, 159, seed=seed, sigma=4.0, g=9.8, hmax=120.0, metrics=metrics)
    b = randomize_state(b, seed + 1, amp_uv=0.05)
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, rd, 1, 2, 0, 10)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 10)
    inp = ref["smag_in"]
    z = np.zeros_like(inp["uup"])
    st, ss = oracle.stress_components_3d(oracle.OracleGrid(b), inp["uup"], inp["vvp"], z, z)
    assert_bitwise({"stress_t": st, "stress_s": ss}, inp, ("stress_t", "stress_s"),
                   f"3-D stress seed {seed} {metrics}")
    assert np.any(st != 0) and np.any(ss != 0)


@pytest.mark.parametrize("seed,metrics", [(21, "uniform"), (22, "stretched")])
def test_pressure_gradients_bitwise(seed, metrics):
    """pressure_gradients (Function/vel_ssh.f90:320-376) called by the
    harness, against the oracle on the inputs it wrote; right-hand sides
    off lcu (lcv) keep their values."""
    b = make_basin(285, 159, seed=seed, sigma=4.0, g=9.8, hmax=120.0, metrics=metrics)
    b = randomize_state(b, seed + 1, amp_uv=0.05)
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, rd, 1, 2, 0, 6)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 6)
    inp, want = ref["pg_in"], ref["pg_out"]
    assert inp["den"].shape[0] == 20
    got = oracle.pressure_gradients(oracle.OracleGrid(b), inp)
    assert_bitwise(got, want, ("rhsx", "rhsy"), f"pressure gradients seed {seed} {metrics}")
    for k in ("rhsx", "rhsy"):
        assert np.any(want[k] != inp[k]) and np.any(want[k] == inp[k])


@pytest.mark.parametrize("seed,metrics", [(31, "uniform"), (32, "stretched")])
def test_density_bitwise(seed, metrics):
    """The density definition (Control/dyncall.f90:437-449, restated in the
    harness) with the reference's own density function
    (Service/oc_algorithms.f90:138-161), against the oracle on the inputs it
    wrote; cells off lu and outside m, n = 2..n-1 keep their values."""
    b = make_basin(285, 159, seed=seed, sigma=4.0, g=9.8, hmax=120.0, metrics=metrics)
    b = randomize_state(b, seed + 1, amp_uv=0.05)
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, rd, 1, 2, 0, 7)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 7)
    inp, want = ref["eos_in"], ref["eos_out"]
    assert inp["tt"].shape[0] == 20
    got = oracle.density(oracle.OracleGrid(b), inp)
    assert_bitwise(got, want, ("den", "den_pot"), f"density seed {seed} {metrics}")
    assert np.any(want["den"] == 7.0) and np.any(want["den"] != 7.0)
    assert np.any(want["den"] != want["den_pot"])


def test_reference_grid_bitwise():
    """inmsom_bt.cases.reference_grid against the reference's own basinpar
    (Service/basin_parameters.f90, harness mode 8) for the shipped
    configuration: every metric and rlh_s bit for bit."""
    from inmsom_bt.cases import REF_NX, REF_NY, reference_grid

    b = make_basin(REF_NX - 4, REF_NY - 4, seed=3, sigma=4.0)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "case.bin")
        write_case(path, b, np.zeros((b.ny, b.nx)), 1, 1, 0, 8)
        ref = oracle.run_reference(path, path + ".out", b.nx, b.ny, 8)
    assert_bitwise(reference_grid(), ref, tuple(ref), "basinpar")


def test_reference_timing_mode():
    """Mode 9 of the harness (the reference's time loop under its own
    start_timer / end_timer) reports the steps and threads it ran and a
    positive time, and the recipe that times it beside the C restatement
    runs end to end on a few steps."""
    import time_reference

    b = time_reference.basin()
    rdis = oracle.OracleGrid(b).rayleigh()
    sec = time_reference.fortran_seconds(b, rdis, 3, 2)
    assert 0.0 < sec < 60.0
    r = time_reference.measure(2, [1])
    t = r["threads"]["1"]
    assert t["fortran_rate"] > 0 and t["c_rate"] > 0
    assert r["substeps_per_outer_step"] == 2 * time_reference.NSTEP
