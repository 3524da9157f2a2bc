"""The Fortran host (ISO_C_BINDING module + driver in the reference's
shape) on the GPU: its final state equals the oracle's bit for bit."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from _util import STATE, assert_bitwise
from inmsom_bt.cases import make_basin, randomize_state, write_case

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "inmsom-elfs_amd", "bin", "inmsom_bt_host")
PARDIR = os.path.join(ROOT, "inmsom-elfs_amd", "fortran", "par")


def _run_host(b, rdis, nsteps, mode, tmp):
    case = os.path.join(tmp, "case.bin")
    out = os.path.join(tmp, "out.bin")
    write_case(case, b, rdis, nsteps, 2, 0, 0, [])
    p = subprocess.run([HOST, PARDIR, case, out, mode, repr(b.g)], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = np.fromfile(out, dtype="<f8").reshape(6, b.ny, b.nx)
    return dict(zip(("ssh", "sshp", "u", "up", "v", "vp"), raw)), p.stdout


@pytest.mark.parametrize("mode", ["literal", "fp64_exact"])
def test_fortran_host_bitwise(mode):
    assert os.path.exists(HOST), "build the host with make -C inmsom-elfs_amd/fortran"
    b = make_basin(96, 64, seed=51, sigma=5.0)
    rd = oracle.OracleGrid(b).rayleigh()
    steps = 12  # ocean_run.par: nstep_barotrop = 2
    want = oracle.run(b, steps, 2, out_steps=[steps], rdis=rd)[steps]
    with tempfile.TemporaryDirectory() as tmp:
        got, out = _run_host(b, rd, steps, mode, tmp)
    assert "Time barotropic" in out
    assert_bitwise(got, want, STATE, f"fortran host {mode}")


def test_fortran_host_range_check_stops_run():
    """A NaN in the initial state makes the host stop after the first step
    with the range-check status, as the reference stops (dyncall.f90:452-516)."""
    b = make_basin(40, 30, seed=52, sigma=4.0)
    b.ssh[10, 10] = np.nan
    b.sshp[10, 10] = np.nan
    rd = oracle.OracleGrid(b).rayleigh()
    with tempfile.TemporaryDirectory() as tmp:
        case = os.path.join(tmp, "case.bin")
        write_case(case, b, rd, 3, 2, 0, 0, [])
        p = subprocess.run([HOST, PARDIR, case, os.path.join(tmp, "o.bin"), "literal"],
                           capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "range check failed at step 1" in p.stdout


def test_fortran_host_coupling_bitwise():
    """The host's coupling step (remove_barotropic_gpu, internal_update_gpu
    through the ISO_C_BINDING module) after the last barotropic step, fed
    with that step's time means, against the oracle's coupling on the
    oracle's means (Control/dyncall.f90:137-231 in order)."""
    b = randomize_state(make_basin(80, 56, seed=53, sigma=5.0), 4, wflux_amp=1e-5)
    rd = oracle.OracleGrid(b).rayleigh()
    steps, nz = 6, 7
    want_run = oracle.run(b, steps, 2, out_steps=[steps], rdis=rd)[steps]
    rng = np.random.default_rng(17)
    shape = (b.ny, b.nx)
    wet = b.wet.astype(np.float64)
    side = {"hhu": 50.0 + np.abs(rng.standard_normal(shape)), "hhv": 50.0 + np.abs(rng.standard_normal(shape)),
            "slpr": 1.0e5 + 100.0 * rng.standard_normal(shape), "uu2d": 0.1 * rng.standard_normal(shape),
            "vv2d": 0.1 * rng.standard_normal(shape), "ssh_i": 0.2 * rng.standard_normal(shape) * wet,
            "sshp_i": 0.2 * rng.standard_normal(shape) * wet}
    uu = rng.standard_normal((nz,) + shape)
    vv = rng.standard_normal((nz,) + shape)
    inp = dict(side, uu=uu, vv=vv, ubrtr_i=want_run["ubrtr_i"], vbrtr_i=want_run["vbrtr_i"],
               pgrx=want_run["ssh4gradx"], pgry=want_run["ssh4grady"], wflux=np.ascontiguousarray(b.wflux))
    want = oracle.coupling(oracle.OracleGrid(b), b.dt, inp)
    with tempfile.TemporaryDirectory() as tmp:
        case = os.path.join(tmp, "case.bin")
        cpl = os.path.join(tmp, "coupling.bin")
        out = os.path.join(tmp, "out.bin")
        write_case(case, b, rd, steps, 2, 0, 0, [])
        with open(cpl, "wb") as fh:
            np.array([nz], dtype="<i4").tofile(fh)
            for k in ("hhu", "hhv", "slpr", "uu2d", "vv2d", "ssh_i", "sshp_i"):
                np.ascontiguousarray(side[k], dtype="<f8").tofile(fh)
            uu.astype("<f8").tofile(fh)
            vv.astype("<f8").tofile(fh)
        p = subprocess.run([HOST, PARDIR, case, out, "literal", repr(b.g), cpl], capture_output=True, text=True,
                           timeout=120)
        print(p.stdout, p.stderr)
        assert p.returncode == 0, p.stdout + p.stderr
        raw = np.fromfile(out, dtype="<f8")
    per = b.nx * b.ny
    st = dict(zip(("ssh", "sshp", "u", "up", "v", "vp"), raw[:6 * per].reshape(6, b.ny, b.nx)))
    assert_bitwise(st, want_run, STATE, "fortran host state before the coupling")
    rest = raw[6 * per:]
    got = {"uu": rest[:nz * per].reshape(nz, b.ny, b.nx), "vv": rest[nz * per:2 * nz * per].reshape(nz, b.ny, b.nx)}
    planes = rest[2 * nz * per:].reshape(10, b.ny, b.nx)
    got.update(zip(("uu2d", "vv2d", "pgrx", "pgry", "ssh_i", "sshp_i", "hhqn", "hhun", "hhvn", "hhhn"), planes))
    assert "coupling step done (7 levels)" in p.stdout
    assert_bitwise(got, want, tuple(got), "fortran host coupling")


def test_fortran_host_forcing_bitwise():
    """The host's forcing step (forcing_3d_gpu through the ISO_C_BINDING
    module, Control/dyncall.f90:37-130) with the case's wind stress and
    drag, against the oracle, bit for bit."""
    b = randomize_state(make_basin(72, 50, seed=31, sigma=5.0, metrics="stretched"), 4)
    rd = oracle.OracleGrid(b).rayleigh()
    nz = 6
    rng = np.random.default_rng(23)
    shape = (b.ny, b.nx)
    og = oracle.OracleGrid(b)
    lcu = (og.masks["lcu"] > 0.5).astype(np.float64)
    lcv = (og.masks["lcv"] > 0.5).astype(np.float64)
    inp = {"dz": (np.arange(nz) + 1.0) / (nz * (nz + 1) / 2.0)}
    for k, msk in (("uu", lcu), ("vv", lcv), ("uup", lcu), ("vvp", lcv)):
        inp[k] = 0.1 * rng.standard_normal((nz,) + shape) * msk
    inp["amuv"] = 10.0 + np.abs(rng.standard_normal((nz,) + shape))
    inp["pgx"] = 1e-3 * rng.standard_normal((nz,) + shape)
    inp["pgy"] = 1e-3 * rng.standard_normal((nz,) + shape)
    for k, msk in (("uu2d", lcu), ("vv2d", lcv), ("uup2d", lcu), ("vvp2d", lcv)):
        inp[k] = 0.1 * rng.standard_normal(shape) * msk
    inp["amuv2d"] = 10.0 + np.abs(rng.standard_normal(shape))
    for k in ("hhq", "hhu", "hhv", "hhh"):
        inp[k] = 50.0 + np.abs(rng.standard_normal(shape))
    inp["slpr"] = 1.0e5 + 100.0 * rng.standard_normal(shape)
    inp["rhsx2d"], inp["rhsy2d"] = np.full(shape, 2.0), np.full(shape, -1.0)
    want = oracle.forcing_3d(og, inp, b.cb_nl, b.ebottom)
    with tempfile.TemporaryDirectory() as tmp:
        case = os.path.join(tmp, "case.bin")
        frc = os.path.join(tmp, "forcing.bin")
        out = os.path.join(tmp, "out.bin")
        write_case(case, b, rd, 1, 2, 0, 0, [])
        with open(frc, "wb") as fh:
            np.array([nz], dtype="<i4").tofile(fh)
            inp["dz"].astype("<f8").tofile(fh)
            for k in ("uu", "vv", "uup", "vvp", "amuv", "pgx", "pgy"):
                np.ascontiguousarray(inp[k], dtype="<f8").tofile(fh)
            for k in ("uu2d", "vv2d", "uup2d", "vvp2d", "amuv2d", "hhq", "hhu", "hhv", "hhh", "slpr", "rhsx2d",
                      "rhsy2d"):
                np.ascontiguousarray(inp[k], dtype="<f8").tofile(fh)
        p = subprocess.run([HOST, PARDIR, case, out, "literal", repr(b.g), "-", frc], capture_output=True,
                           text=True, timeout=120)
        print(p.stdout, p.stderr)
        assert p.returncode == 0, p.stdout + p.stderr
        raw = np.fromfile(out, dtype="<f8")
    per = b.nx * b.ny
    got = dict(zip(("rhsx2d", "rhsy2d"), raw[6 * per:].reshape(2, b.ny, b.nx)))
    assert "forcing step done (6 levels)" in p.stdout
    assert_bitwise(got, want, ("rhsx2d", "rhsy2d"), "fortran host forcing")


def test_fortran_host_baroclinic_bitwise():
    """The host's baroclinic step (baroclinic_dynamics_gpu through the
    ISO_C_BINDING module, Function/vel_ssh.f90:794-1027) against the oracle,
    bit for bit."""
    b = randomize_state(make_basin(66, 48, seed=37, sigma=5.0, metrics="stretched"), 4)
    rd = oracle.OracleGrid(b).rayleigh()
    nz, tau = 20, 40.0
    rng = np.random.default_rng(29)
    sh = (b.ny, b.nx)
    f = {k: 0.1 * rng.standard_normal((nz,) + sh) for k in ("uu", "uup", "vv", "vvp")}
    for k in ("rhsx", "rhsy", "rhsx_tran", "rhsy_tran"):
        f[k] = 1e3 * rng.standard_normal((nz,) + sh)
    f["ww"] = 1e-4 * rng.standard_normal((nz + 1,) + sh)
    f["nu"] = 1e-3 + 1e-2 * rng.random((nz + 1,) + sh)
    two = ("pgrx", "pgry", "rdis", "tx_surf", "ty_surf", "tx_bot", "ty_bot", "hhu", "hhup", "hhun", "hhv", "hhvp",
           "hhvn", "hhh")
    for k in two[:7]:
        f[k] = 1e-3 * rng.standard_normal(sh)
    f["rdis"] = 1e-5 * rng.random(sh)
    for k in two[7:]:
        f[k] = 40.0 + rng.random(sh)
    dz = 0.5 + rng.random(nz)
    f["dz"] = dz / dz.sum()
    f["hzt"] = (0.5 + rng.random(nz + 1)) / nz
    want = oracle.baroclinic(oracle.OracleGrid(b), tau, f)
    with tempfile.TemporaryDirectory() as tmp:
        case = os.path.join(tmp, "case.bin")
        bcl = os.path.join(tmp, "baroclinic.bin")
        out = os.path.join(tmp, "out.bin")
        write_case(case, b, rd, 1, 2, 0, 0, [])
        with open(bcl, "wb") as fh:
            np.array([nz], dtype="<i4").tofile(fh)
            np.array([tau], dtype="<f8").tofile(fh)
            f["dz"].astype("<f8").tofile(fh)
            f["hzt"].astype("<f8").tofile(fh)
            for k in ("uu", "uup", "vv", "vvp", "rhsx", "rhsy", "rhsx_tran", "rhsy_tran", "ww", "nu") + two:
                np.ascontiguousarray(f[k], dtype="<f8").tofile(fh)
        p = subprocess.run([HOST, PARDIR, case, out, "literal", repr(b.g), "-", "-", bcl], capture_output=True,
                           text=True, timeout=120)
        print(p.stdout, p.stderr)
        assert p.returncode == 0, p.stdout + p.stderr
        raw = np.fromfile(out, dtype="<f8")
    per = b.nx * b.ny
    got = dict(zip(("uu", "uup", "vv", "vvp"), raw[6 * per:].reshape(4, nz, b.ny, b.nx)))
    assert "baroclinic step done (20 levels)" in p.stdout
    assert_bitwise(got, want, ("uu", "uup", "vv", "vvp"), "fortran host baroclinic")


def test_fortran_host_pressure_gradients_bitwise():
    """The host's pressure step (pressure_gradients_gpu through the
    ISO_C_BINDING module, Function/vel_ssh.f90:320-376) against the oracle,
    bit for bit; RHSx, RHSy are updated in place as the reference does."""
    b = randomize_state(make_basin(66, 48, seed=41, sigma=5.0, metrics="stretched"), 4)
    rd = oracle.OracleGrid(b).rayleigh()
    nz = 20
    rng = np.random.default_rng(31)
    sh = (b.ny, b.nx)
    f = {"z": np.sort(rng.random(nz)),
         "den": 1025.0 + 0.3 * np.arange(1, nz + 1)[:, None, None] + rng.standard_normal((nz,) + sh),
         "rhsx": rng.standard_normal((nz,) + sh), "rhsy": rng.standard_normal((nz,) + sh),
         "hhq": 30.0 + 20.0 * rng.random(sh), "hhu": 30.0 + rng.random(sh), "hhv": 30.0 + rng.random(sh)}
    want = oracle.pressure_gradients(oracle.OracleGrid(b), f)
    with tempfile.TemporaryDirectory() as tmp:
        case = os.path.join(tmp, "case.bin")
        prs = os.path.join(tmp, "pressure.bin")
        out = os.path.join(tmp, "out.bin")
        write_case(case, b, rd, 1, 2, 0, 0, [])
        with open(prs, "wb") as fh:
            np.array([nz], dtype="<i4").tofile(fh)
            f["z"].astype("<f8").tofile(fh)
            for k in ("den", "rhsx", "rhsy", "hhq", "hhu", "hhv"):
                np.ascontiguousarray(f[k], dtype="<f8").tofile(fh)
        p = subprocess.run([HOST, PARDIR, case, out, "literal", repr(b.g), "-", "-", "-", prs], capture_output=True,
                           text=True, timeout=120)
        print(p.stdout, p.stderr)
        assert p.returncode == 0, p.stdout + p.stderr
        raw = np.fromfile(out, dtype="<f8")
    per = b.nx * b.ny
    got = dict(zip(("rhsx", "rhsy"), raw[6 * per:].reshape(2, nz, b.ny, b.nx)))
    assert "pressure gradients done (20 levels)" in p.stdout
    assert_bitwise(got, want, ("rhsx", "rhsy"), "fortran host pressure gradients")


def test_fortran_host_density_bitwise():
    """The host's density step (density_gpu through the ISO_C_BINDING
    module, Control/dyncall.f90:437-449 with Service/oc_algorithms.f90:
    138-161) against the oracle, bit for bit."""
    b = randomize_state(make_basin(66, 48, seed=43, sigma=5.0, metrics="stretched"), 4)
    rd = oracle.OracleGrid(b).rayleigh()
    nz = 20
    rng = np.random.default_rng(33)
    sh = (b.ny, b.nx)
    s3 = (nz,) + sh
    f = {"z": np.sort(rng.random(nz)), "tt": 2.0 + 25.0 * rng.random(s3), "ss": 5.0 + 30.0 * rng.random(s3),
         "den": np.full(s3, 3.0), "den_pot": np.full(s3, -3.0), "hhq": 10.0 + 90.0 * rng.random(sh)}
    want = oracle.density(oracle.OracleGrid(b), f)
    with tempfile.TemporaryDirectory() as tmp:
        case = os.path.join(tmp, "case.bin")
        eos = os.path.join(tmp, "density.bin")
        out = os.path.join(tmp, "out.bin")
        write_case(case, b, rd, 1, 2, 0, 0, [])
        with open(eos, "wb") as fh:
            np.array([nz], dtype="<i4").tofile(fh)
            f["z"].astype("<f8").tofile(fh)
            for k in ("tt", "ss", "den", "den_pot", "hhq"):
                np.ascontiguousarray(f[k], dtype="<f8").tofile(fh)
        p = subprocess.run([HOST, PARDIR, case, out, "literal", repr(b.g), "-", "-", "-", "-", eos],
                           capture_output=True, text=True, timeout=120)
        print(p.stdout, p.stderr)
        assert p.returncode == 0, p.stdout + p.stderr
        raw = np.fromfile(out, dtype="<f8")
    per = b.nx * b.ny
    got = dict(zip(("den", "den_pot"), raw[6 * per:].reshape(2, nz, b.ny, b.nx)))
    assert "density done (20 levels)" in p.stdout
    assert_bitwise(got, want, ("den", "den_pot"), "fortran host density")


@pytest.mark.parametrize("metrics,ksw4", [("uniform", 0), ("stretched", 0), ("uniform", 1), ("stretched", 1)])
def test_fortran_host_dropin_bitwise(metrics, ksw4):
    """The host's 'dropin' mode calls barotropic_dynamics_gpu, the
    reference's own argument list (Function/vel_ssh.f90:379-439) through the
    ISO_C_BINDING module, every step on host arrays with a fixed RHS (the
    wind stress as a face force over the reference density): the state after
    the steps and the last call's averages equal the oracle's calls bit for
    bit.  ksw4 is the parameter file's ksw_lat4 (phys_proc.par line 6), which
    the host passes on as the reference's caller does; 1 adds the biharmonic
    viscosity."""
    b = randomize_state(make_basin(96, 64, seed=55, sigma=5.0, metrics=metrics), 6, wflux_amp=1e-6)
    rd = oracle.OracleGrid(b).rayleigh()
    steps, nstep = 7, 2  # ocean_run.par: nstep_barotrop = 2
    og = oracle.OracleGrid(b)
    rhsx = b.taux * b.dxt * b.dyh / 1025.0
    rhsy = b.tauy * b.dxh * b.dyt / 1025.0
    st = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
    for _ in range(steps):
        acc = {k: np.zeros((b.ny, b.nx)) for k in ("ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")}
        oracle.barotropic_dynamics(og, b.dt, nstep, ksw4, st, acc, rhsx, rhsy, b.wflux, b.mu, b.mu4, rd)
    with tempfile.TemporaryDirectory() as tmp:
        case = os.path.join(tmp, "case.bin")
        out = os.path.join(tmp, "out.bin")
        write_case(case, b, rd, steps, nstep, 0, 0, [])
        pardir = os.path.join(tmp, "par")
        shutil.copytree(PARDIR, pardir)
        pp = os.path.join(pardir, "phys_proc.par")
        lines = open(pp).read().split("\n")
        lines[5] = f"           {ksw4}  : Lateral 4th order momentum mixing: 0 - no, 1 - yes"
        with open(pp, "w") as fh:
            fh.write("\n".join(lines))
        p = subprocess.run([HOST, pardir, case, out, "dropin", repr(b.g)], capture_output=True, text=True,
                           timeout=120)
        print(p.stdout, p.stderr)
        assert p.returncode == 0, p.stdout + p.stderr
        raw = np.fromfile(out, dtype="<f8").reshape(10, b.ny, b.nx)
    got = dict(zip(("ssh", "sshp", "u", "up", "v", "vp", "ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady"), raw))
    assert_bitwise(got, {**st, **acc}, STATE + tuple(acc), f"fortran host drop-in ({metrics}, ksw4={ksw4})")
