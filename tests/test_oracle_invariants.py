"""Checks that pin the oracle by what the mathematics fixes, independent of
the compiled reference: rest state, volume conservation, closed forms of
the depth averages, stress tensor and Rayleigh coefficient, hand-derived
continuity increments, the C-grid mask definitions, gravity-wave speed,
mirror symmetry and (for the periodic wrap, which the shipped reference
build does not exercise) translation invariance."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from _util import periodic_pair
from inmsom_bt.cases import make_basin, randomize_state


def _flat(nxi, nyi, h=50.0, bumps=(), **kw):
    """All-wet rectangle of constant depth h; surface bumps (amplitude,
    x-centre, y-centre [cells], radius [m]) on the wet cells."""
    b = make_basin(nxi, nyi, seed=0, sigma=2.0, bumps=[], **kw)
    wet = np.zeros((b.ny, b.nx), dtype=bool)
    wet[2:-2, 2:-2] = True
    b.lbasins = np.where(wet, 0, 1).astype(np.int32)
    b.hhq_rest = np.where(wet, h, 0.0)
    b.mu = b.mu * 0 + 10.0 * wet
    xc = (np.arange(b.nx) - 2 + 0.5)[None, :]
    yc = (np.arange(b.ny) - 2 + 0.5)[:, None]
    ssh = np.zeros((b.ny, b.nx))
    for amp, x0, y0, rad in bumps:
        ssh += amp * np.exp(-(((xc - x0) * 1000.0) ** 2 + ((yc - y0) * 1000.0) ** 2) / rad ** 2)
    b.ssh = ssh * wet
    b.sshp = b.ssh.copy()
    return b


def test_rest_state_stays_exactly_at_rest():
    b = make_basin(60, 40, seed=5, sigma=4.0, tau0=0.0, bumps=[])
    out = oracle.run(b, 25, 2, out_steps=[25])[25]
    for k in ("ssh", "sshp", "u", "up", "v", "vp", "ubrtr_i", "vbrtr_i"):
        assert not np.any(out[k]), k


def test_volume_conserved_in_closed_basin():
    b = make_basin(80, 60, seed=6, sigma=5.0)
    og = oracle.OracleGrid(b)
    v0 = og.volume(b.ssh)
    outs = oracle.run(b, 60, 2, out_steps=[20, 60], grid=og)
    for s, o in outs.items():
        v = og.volume(o["ssh"])
        assert abs(v - v0) <= 1e-12 * abs(v0), (s, v, v0)


def test_hh_init_closed_form_uniform_metrics():
    b = randomize_state(make_basin(40, 30, seed=7, sigma=3.0), 3)
    og = oracle.OracleGrid(b)
    ds, da = og.new_depths()
    import ctypes as C
    oracle.lib().bto_hh_init(og.ref, C.byref(ds), oracle._d(np.ascontiguousarray(b.ssh)),
                             oracle._d(np.ascontiguousarray(b.sshp)))
    hq = b.hhq_rest + b.ssh
    lu = og.masks["lu"].astype(np.float64)
    llu = og.masks["llu"] > 0.5
    num = hq[:, :-1] * lu[:, :-1] + hq[:, 1:] * lu[:, 1:]
    den = lu[:, :-1] + lu[:, 1:]
    with np.errstate(invalid="ignore", divide="ignore"):
        want = np.where(llu[:, :-1], num / den, 0.0)
    got = da[3][:, :-1]
    assert np.allclose(got, want, rtol=1e-15, atol=0), np.max(np.abs(got - want))
    # the H-point depth is the wet-weighted mean of the 4 surrounding T depths
    luh = og.masks["luh"] > 0.5
    num4 = hq[:-1, :-1] * lu[:-1, :-1] + hq[:-1, 1:] * lu[:-1, 1:] + hq[1:, :-1] * lu[1:, :-1] + \
        hq[1:, 1:] * lu[1:, 1:]
    den4 = lu[:-1, :-1] + lu[:-1, 1:] + lu[1:, :-1] + lu[1:, 1:]
    with np.errstate(invalid="ignore", divide="ignore"):
        want4 = np.where(luh[:-1, :-1], num4 / den4, 0.0)
    assert np.allclose(da[9][:-1, :-1], want4, rtol=1e-15, atol=0)


def test_stress_of_linear_flow_and_zero_viscous_rhs():
    """u = a*x, v = b*y on a flat, all-wet basin: tension = a - b exactly in the
    interior, shear = 0, so the harmonic viscous force vanishes wherever
    the stencil sees only interior wet points."""
    b = _flat(30, 24)
    a, c = 2e-5, -3e-5
    x_u = (np.arange(b.nx) + 0.5) * 1000.0
    y_v = (np.arange(b.ny) + 0.5) * 1000.0
    og = oracle.OracleGrid(b)
    lcu, lcv = og.masks["lcu"] > 0.5, og.masks["lcv"] > 0.5
    b.up = np.where(lcu, a * x_u[None, :], 0.0)
    b.vp = np.where(lcv, c * y_v[:, None], 0.0)
    L, d = oracle.lib(), oracle._d
    z = lambda: np.zeros((b.ny, b.nx))
    strt, strs, rx, ry = z(), z(), z(), z()
    L.bto_stress_components(og.ref, d(b.up), d(b.vp), d(strt), d(strs))
    hq = np.ascontiguousarray(b.hhq_rest)  # flat depth at every point type
    mu = np.ascontiguousarray(b.mu)
    L.bto_uv_diff2(og.ref, d(mu), d(strt), d(strs), d(hq), d(hq), d(hq), d(hq), d(rx), d(ry))
    assert np.allclose(strt[4:-4, 4:-4], a - c, rtol=1e-12, atol=0)
    assert np.max(np.abs(strs[4:-4, 4:-4])) < 1e-18
    scale = 10.0 * 50.0 * abs(a - c) * 1000.0  # mu * H * |stress| * dy
    assert np.max(np.abs(rx[5:-5, 5:-5])) < 1e-9 * scale
    assert np.max(np.abs(ry[5:-5, 5:-5])) < 1e-9 * scale


def test_rayleigh_closed_form_linear_slope():
    b = _flat(20, 16)
    gx, gy = 3e-3, -4e-3
    xx = np.arange(b.nx)[None, :] * 1000.0
    yy = np.arange(b.ny)[:, None] * 1000.0
    wet = b.lbasins == 0
    b.hhq_rest = np.where(wet, 100.0 + gx * xx + gy * yy, 0.0)
    r = oracle.OracleGrid(b).rayleigh(1e-4)
    want = 1e-4 * math.hypot(gx, gy)
    assert np.allclose(r[3:-3, 3:-3], want, rtol=1e-9)
    # at a coast the one-sided slope is used: west column sees only its east face
    assert np.isclose(r[5, 2], 1e-4 * math.hypot(gx, gy), rtol=1e-9)


def test_continuity_single_transport_by_hand():
    """One nonzero u on a flat basin at rest: after one substep the two T
    cells beside that U face change by -/+ 2 tau u h dy / (dx dy)."""
    b = _flat(12, 10, h=40.0, tau0=0.0, bumps=[])
    b.cb_nl = 0.0
    m0, n0 = 7, 6  # 0-based U point between T(7,6) and T(8,6)
    b.u[n0, m0] = b.up[n0, m0] = 0.3
    rd = np.zeros((b.ny, b.nx))
    out = oracle.run(b, 1, 1, out_steps=[1], rdis=rd)[1]
    tau = b.dt
    d = 2 * tau * 0.3 * 40.0 * 1000.0 / (1000.0 * 1000.0)
    # Asselin: sshp = ssh + gamma/2 (sshn - 2 ssh + sshp) with ssh = sshp = 0
    assert math.isclose(out["ssh"][n0, m0], -d, rel_tol=1e-14)
    assert math.isclose(out["ssh"][n0, m0 + 1], d, rel_tol=1e-14)
    assert math.isclose(out["sshp"][n0, m0 + 1], 0.025 * d, rel_tol=1e-13)
    others = out["ssh"].copy()
    others[n0, m0] = others[n0, m0 + 1] = 0.0
    assert not np.any(others)


def test_gridcon_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(5):
        lb = (rng.random((9, 11)) < 0.35).astype(np.int32)
        m = oracle.gridcon(lb)
        lu = (lb == 0).astype(np.float32)
        ny, nx = lb.shape
        for n in range(ny):
            for mm in range(nx):
                inr = 1 <= mm <= nx - 3 and 1 <= n <= ny - 3
                four = [lu[n, mm], lu[n, min(mm + 1, nx - 1)], lu[min(n + 1, ny - 1), mm],
                        lu[min(n + 1, ny - 1), min(mm + 1, nx - 1)]]
                assert m["lu"][n, mm] == lu[n, mm]
                assert m["luh"][n, mm] == (1.0 if inr and sum(four) > 0 else 0.0)
                assert m["luu"][n, mm] == (1.0 if inr and min(four) > 0 else 0.0)
                assert m["llu"][n, mm] == (1.0 if inr and four[0] + four[1] > 0 else 0.0)
                assert m["lcu"][n, mm] == (1.0 if inr and four[0] * four[1] > 0 else 0.0)
                assert m["llv"][n, mm] == (1.0 if inr and four[0] + four[2] > 0 else 0.0)
                assert m["lcv"][n, mm] == (1.0 if inr and four[0] * four[2] > 0 else 0.0)


def test_gravity_wave_speed():
    """A 1-D ridge of surface height in a long flat channel (f = 0, no
    friction, no viscosity) splits into two fronts moving at sqrt(g H)."""
    h = 40.0
    b = _flat(400, 6, h=h, tau0=0.0, f0=0.0, beta=0.0, nu=0.0, bumps=[])
    b.mu[:] = 0.0
    b.cb_nl = 0.0
    x = (np.arange(b.nx) - 2 + 0.5) * 1000.0
    eta = 0.01 * np.exp(-((x - 200e3) / 8e3) ** 2)
    b.ssh = np.where(b.lbasins == 0, eta[None, :], 0.0)
    b.sshp = b.ssh.copy()
    steps, nstep = 150, 2
    rd = np.zeros((b.ny, b.nx))
    out = oracle.run(b, steps, nstep, out_steps=[steps], rdis=rd)[steps]
    row = out["ssh"][4, 2:-2]
    right = np.argmax(row[200:]) + 200
    dist = (right + 0.5) * 1000.0 - 200e3
    c = dist / (steps * b.dt)
    assert abs(c - math.sqrt(b.g * h)) / math.sqrt(b.g * h) < 0.03, c


def test_mirror_symmetry_without_rotation():
    b = _flat(40, 30, h=30.0, tau0=0.0, f0=0.0, beta=0.0,
              bumps=[(0.3, 20.0, 15.0, 4e3)])
    steps = 20
    out = oracle.run(b, steps, 2, out_steps=[steps], rdis=np.zeros((b.ny, b.nx)))[steps]
    s = out["ssh"][2:-2, 2:-2]
    assert np.allclose(s, s[:, ::-1], rtol=0, atol=1e-13 * np.max(np.abs(s)))
    assert np.allclose(s, s[::-1, :], rtol=0, atol=1e-13 * np.max(np.abs(s)))


@pytest.mark.parametrize("shift", [3, 11])
def test_periodic_x_translation_invariance(shift):
    """Periodic wrap (Service/oc_algorithms.f90:50-95; parity with the
    compiled reference is not available, its build has periodicity 0):
    in a zonally periodic channel with forcing uniform along x, shifting the
    initial state by k columns shifts the solution by k columns."""
    b = make_basin(48, 20, seed=9, sigma=3.0, bumps=[(0.2, 10.0, 10.0, 5e3)])
    wet = np.zeros((b.ny, b.nx), dtype=bool)
    wet[2:-2, :] = True
    b.lbasins = np.where(wet, 0, 1).astype(np.int32)
    b.hhq_rest = np.where(wet, 60.0, 0.0)
    b.mu = 10.0 * wet
    x = np.arange(b.nx) - 2
    b.ssh = np.where(wet, 0.2 * np.exp(-(((x % 48) - 10.0) ** 2) / 9.0)[None, :], 0.0)
    b.sshp = b.ssh.copy()
    rd = np.zeros((b.ny, b.nx))

    def shifted(a):
        inner = np.roll(a[:, 2:-2], shift, axis=1)
        out = a.copy()
        out[:, 2:-2] = inner
        out[:, 1] = inner[:, -1]
        out[:, -2] = inner[:, 0]
        return out

    b2 = b.copy()
    for k in ("ssh", "sshp", "u", "up", "v", "vp"):
        setattr(b2, k, shifted(getattr(b, k)))
    o1 = oracle.run(b, 12, 2, out_steps=[12], rdis=rd, periodicity=(1, 0))[12]
    o2 = oracle.run(b2, 12, 2, out_steps=[12], rdis=rd, periodicity=(1, 0))[12]
    for k in ("ssh", "u", "v"):
        want = np.roll(o1[k][:, 2:-2], shift, axis=1)
        assert np.allclose(o2[k][:, 2:-2], want, rtol=0, atol=1e-12 * max(1.0, np.max(np.abs(want))))


def _coupling_inputs(b, nz=3, seed=0, wflux=True):
    rng = np.random.default_rng(seed)
    shape = (b.ny, b.nx)

    def f(scale):
        return rng.standard_normal(shape) * scale
    wet = b.wet.astype(np.float64)
    inp = {"ubrtr_i": f(5.0), "vbrtr_i": f(5.0), "hhu": 50.0 + np.abs(f(10.0)), "hhv": 50.0 + np.abs(f(10.0)),
           "slpr": 1.0e5 + f(100.0), "wflux": f(1e-6) * wet if wflux else np.zeros(shape),
           "uu2d": f(0.1), "vv2d": f(0.1), "pgrx": f(1.0), "pgry": f(1.0), "ssh_i": f(0.2) * wet,
           "sshp_i": f(0.2) * wet}
    inp["uu"] = rng.standard_normal((nz,) + shape)
    inp["vv"] = rng.standard_normal((nz,) + shape)
    return inp


def test_coupling_removal_closed_form():
    b = _flat(40, 30)
    og = oracle.OracleGrid(b)
    inp = _coupling_inputs(b, nz=4)
    uu, vv = inp["uu"].copy(), inp["vv"].copy()
    oracle.lib().bto_remove_barotropic(C.byref(og.g), 4, oracle._d(uu), oracle._d(vv),
                                       oracle._d(np.ascontiguousarray(inp["uu2d"])),
                                       oracle._d(np.ascontiguousarray(inp["vv2d"])))
    lcu = og.masks["lcu"] > 0.5
    lcv = og.masks["lcv"] > 0.5
    for k in range(4):
        np.testing.assert_array_equal(uu[k], np.where(lcu, inp["uu"][k] - inp["uu2d"], inp["uu"][k]))
        np.testing.assert_array_equal(vv[k], np.where(lcv, inp["vv"][k] - inp["vv2d"], inp["vv"][k]))


def test_ssh_internal_conserves_volume():
    """With no surface water flux the continuity step with the time-mean
    transports telescopes in a closed basin: sum(ssh_new dx dy) over wet
    cells equals sum(sshp_i dx dy); ssh_i takes ssh_new."""
    b = _flat(48, 36)
    og = oracle.OracleGrid(b)
    inp = _coupling_inputs(b, nz=2, seed=3, wflux=False)
    lcu = og.masks["lcu"] > 0.5
    lcv = og.masks["lcv"] > 0.5
    inp["ubrtr_i"] = inp["ubrtr_i"] * lcu          # transports live on open faces
    inp["vbrtr_i"] = inp["vbrtr_i"] * lcv
    out = oracle.coupling(og, b.dt, inp)
    v0 = og.volume(inp["sshp_i"])
    v1 = og.volume(out["ssh_i"])
    assert abs(v1 - v0) <= 1e-12 * np.abs(inp["sshp_i"]).sum() * 1e6
    # and the surface moved: the transports are not divergence-free
    assert np.abs(out["ssh_i"] - inp["sshp_i"]).max() > 1e-6
    # uu, vv: removal then re-add of the new 2-D velocity ubrtr_i / hhu
    k = 1
    want = np.where(lcu, inp["uu"][k] - inp["uu2d"] + inp["ubrtr_i"] / inp["hhu"], inp["uu"][k])
    np.testing.assert_allclose(out["uu"][k], want, rtol=1e-15, atol=1e-15)


def test_coupling_velocity_and_pressure_closed_form():
    """A sea-level pressure linear in x and y changes pgrx by the constant
    -a hhu dyh / RefDen and pgry by -c hhv dxh / RefDen on open faces of
    the wide range, and sets uu2d = ubrtr_i / hhu there; the re-added 3-D
    velocity is uu - uu2d_old + uu2d_new on those faces; everything else
    keeps its value."""
    b = make_basin(30, 24, seed=4, sigma=3.0, metrics="stretched")
    og = oracle.OracleGrid(b)
    inp = _coupling_inputs(b, nz=2, seed=5)
    a, c = 3.0, -2.0
    mm = np.arange(1, b.nx + 1)[None, :]
    nn = np.arange(1, b.ny + 1)[:, None]
    inp["slpr"] = 1.0e5 + a * mm + c * nn + 0.0 * inp["slpr"]
    out = oracle.coupling(og, b.dt, inp)
    wide = np.zeros((b.ny, b.nx), dtype=bool)
    wide[1:-1, 1:-1] = True
    lcu = (og.masks["lcu"] > 0.5) & wide
    lcv = (og.masks["lcv"] > 0.5) & wide
    rho = b.rho0
    np.testing.assert_allclose(out["pgrx"], np.where(lcu, inp["pgrx"] - a * inp["hhu"] * b.dyh / rho, inp["pgrx"]),
                               rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(out["pgry"], np.where(lcv, inp["pgry"] - c * inp["hhv"] * b.dxh / rho, inp["pgry"]),
                               rtol=1e-13, atol=1e-15)
    np.testing.assert_array_equal(out["uu2d"], np.where(lcu, inp["ubrtr_i"] / inp["hhu"], inp["uu2d"]))
    np.testing.assert_array_equal(out["vv2d"], np.where(lcv, inp["vbrtr_i"] / inp["hhv"], inp["vv2d"]))
    for k in range(2):
        np.testing.assert_allclose(out["uu"][k], np.where(lcu, inp["uu"][k] - inp["uu2d"] + out["uu2d"],
                                                          inp["uu"][k]), rtol=1e-14, atol=1e-14)
        np.testing.assert_allclose(out["vv"][k], np.where(lcv, inp["vv"][k] - inp["vv2d"] + out["vv2d"],
                                                          inp["vv"][k]), rtol=1e-14, atol=1e-14)


def test_forcing_3d_closed_form_at_rest():
    """With every velocity zero, RHSx2d on open interior U faces is the depth
    mean of the pressure-gradient terms plus the wind stress times the face
    area minus the sea-level-pressure gradient (Control/dyncall.f90:
    92-126): the advective, viscous and bottom terms vanish."""
    b = make_basin(28, 22, seed=6, sigma=3.0, metrics="stretched")
    og = oracle.OracleGrid(b)
    nz = 4
    sh = (b.ny, b.nx)
    rng = np.random.default_rng(2)
    z3, z2 = np.zeros((nz,) + sh), np.zeros(sh)
    a = np.array([1.0, -2.0, 0.5, 3.0]) * 1e-3
    dz = np.array([0.1, 0.2, 0.3, 0.4])
    inp = {"uu": z3, "vv": z3, "uup": z3, "vvp": z3, "amuv": 10.0 + z3,
           "pgx": a[:, None, None] + z3, "pgy": -a[:, None, None] + z3,
           "uu2d": z2, "vv2d": z2, "uup2d": z2, "vvp2d": z2, "amuv2d": 10.0 + z2,
           "hhq": 40.0 + np.abs(rng.standard_normal(sh)), "hhu": 40.0 + np.abs(rng.standard_normal(sh)),
           "hhv": 40.0 + np.abs(rng.standard_normal(sh)), "hhh": 40.0 + z2,
           "slpr": 1.0e5 + 3.0 * np.arange(b.nx)[None, :] + 0.0 * z2, "rhsx2d": z2 + 5.0, "rhsy2d": z2 + 5.0, "dz": dz}
    out = oracle.forcing_3d(og, inp, b.cb_nl, b.ebottom)
    inner = np.zeros(sh, dtype=bool)
    inner[2:-2, 2:-2] = True
    lcu = (og.masks["lcu"] > 0.5) & inner
    taux = b.taux
    wind = (taux * b.dx * b.dy + np.roll(taux * b.dx * b.dy, -1, axis=1)) / b.rho0 / 2.0
    want = (a * dz).sum() + wind - 3.0 * inp["hhu"] * b.dyh / b.rho0
    np.testing.assert_allclose(out["rhsx2d"][lcu], want[lcu], rtol=1e-12, atol=1e-14)
    # on llu cells that are not open interior faces only the depth mean is written
    llu_only = (og.masks["llu"] > 0.5) & ~lcu
    llu_only[[0, -1], :] = False
    llu_only[:, [0, -1]] = False
    assert np.allclose(out["rhsx2d"][llu_only], (a * dz).sum(), rtol=1e-13)


def _baroclinic_pure_viscosity(seed, uniform_profile):
    """Inputs of baroclinic_dynamics with only the implicit vertical
    viscosity left: no vertical velocity, Coriolis, drag, RHS, pressure or
    stress terms, equal depths at all time levels."""
    b = make_basin(40, 30, seed=seed, sigma=3.0, metrics="stretched")
    b.rlh_s = np.zeros_like(b.rlh_s)
    og = oracle.OracleGrid(b)
    rng = np.random.default_rng(seed)
    nz, sh = 8, (b.ny, b.nx)
    z3, z2 = np.zeros((nz,) + sh), np.zeros(sh)
    h = 40.0 + np.abs(rng.standard_normal(sh))
    if uniform_profile:
        col = rng.standard_normal(sh)
        uup = np.broadcast_to(col, (nz,) + sh).copy()
        vvp = np.broadcast_to(-col, (nz,) + sh).copy()
    else:
        uup, vvp = rng.standard_normal((nz,) + sh), rng.standard_normal((nz,) + sh)
    inp = {"dz": 0.5 + rng.random(nz), "hzt": 0.5 + rng.random(nz + 1), "uu": uup.copy(), "uup": uup,
           "vv": vvp.copy(), "vvp": vvp, "ww": np.zeros((nz + 1,) + sh),
           "nu": 1e-2 * (1.0 + rng.random((nz + 1,) + sh)), "rhsx": z3, "rhsy": z3, "rhsx_tran": z3,
           "rhsy_tran": z3}
    inp["dz"] /= inp["dz"].sum()
    for k in ("pgrx", "pgry", "rdis", "tx_surf", "ty_surf", "tx_bot", "ty_bot"):
        inp[k] = z2
    for k in ("hhu", "hhup", "hhun", "hhv", "hhvp", "hhvn", "hhh"):
        inp[k] = h
    return b, og, inp


def test_baroclinic_viscosity_keeps_uniform_profile():
    """Vertical diffusion leaves a depth-uniform velocity unchanged: every
    row of the implicit system sums to bp (a + b + c = bp), so un = uup."""
    b, og, inp = _baroclinic_pure_viscosity(3, True)
    out = oracle.baroclinic(og, 40.0, inp)
    cu = oracle.OracleGrid(b).masks["lcu"] > 0.5
    cu[:2, :] = cu[-2:, :] = False
    cu[:, :2] = cu[:, -2:] = False
    for name, src in (("uu", "uup"), ("vv", "vvp")):
        err = np.abs(out[name] - inp[src])[:, cu].max()
        assert err <= 1e-12 * np.abs(inp[src]).max(), (name, err)


def test_baroclinic_viscosity_conserves_depth_integral_and_matches_dense_solve():
    """The implicit viscosity only moves momentum between levels: the
    dz-weighted sum of un equals that of uup (bp = bp0, no stress).  Each
    column also equals a dense solve of the tridiagonal system built from
    the reference's coefficients (factor8 is the Thomas algorithm)."""
    b, og, inp = _baroclinic_pure_viscosity(4, False)
    tau = 40.0
    out = oracle.baroclinic(og, tau, inp)
    dz, hzt, nu = inp["dz"], inp["hzt"], inp["nu"]
    masks = og.masks
    nz = dz.size
    checked = 0
    for n in range(2, b.ny - 2):
        for m in range(2, b.nx - 2):
            if masks["lcu"][n, m] <= 0.5:
                continue
            s_in, s_out = (dz * inp["uup"][:, n, m]).sum(), (dz * out["uu"][:, n, m]).sum()
            assert abs(s_out - s_in) <= 1e-12 * max(1.0, np.abs(inp["uup"][:, n, m]).max())
            if checked < 25:
                area = b.dxt[n, m] * b.dyh[n, m]
                bp = inp["hhun"][n, m] * area / tau / 2.0
                d = [(nu[k + 1, n, m] + nu[k + 1, n, m + 1]) / 2.0 / inp["hhun"][n, m] / hzt[k + 1] * area
                     for k in range(nz - 1)]
                A = np.zeros((nz, nz))
                for k in range(nz):
                    A[k, k] = bp
                    if k > 0:
                        A[k, k] += d[k - 1] / dz[k]
                        A[k, k - 1] = -d[k - 1] / dz[k]
                    if k < nz - 1:
                        A[k, k] += d[k] / dz[k]
                        A[k, k + 1] = -d[k] / dz[k]
                want = np.linalg.solve(A, bp * inp["uup"][:, n, m])
                assert np.allclose(out["uu"][:, n, m], want, rtol=1e-12, atol=1e-14)
                checked += 1
    assert checked == 25
    # the Robert-Asselin step of the filtered level (vel_ssh.f90:1006-1010)
    ts = b.time_smooth
    want_p = inp["uu"] + ts * (out["uu"] - 2.0 * inp["uu"] + inp["uup"]) / 2.0
    cu = masks["lcu"] > 0.5
    assert np.allclose(out["uup"][:, cu], want_p[:, cu], rtol=1e-13, atol=1e-15)


def test_vertical_velocity_closed_basin_balance():
    """Continuity: in a closed basin the transport divergence sums to zero
    over the wet cells, so sum(w(k=1) dx dy) over lu interior cells equals
    sum(wflux / RefDen) * sum(dz) (full free surface); with every velocity
    zero each column's w grows linearly from the bottom."""
    b = make_basin(50, 36, seed=8, sigma=3.0, metrics="stretched")
    og = oracle.OracleGrid(b)
    rng = np.random.default_rng(8)
    nz, sh = 6, (b.ny, b.nx)
    lcu, lcv, lu = (og.masks[k] > 0.5 for k in ("lcu", "lcv", "lu"))
    dz = 0.5 + rng.random(nz)
    wflux = 1e-5 * rng.standard_normal(sh) * lu
    inp = {"dz": dz, "uu": rng.standard_normal((nz,) + sh) * lcu, "vv": rng.standard_normal((nz,) + sh) * lcv,
           "ww": np.zeros((nz + 1,) + sh), "hu": 30.0 + rng.random(sh), "hv": 30.0 + rng.random(sh),
           "wflux": wflux}
    w = oracle.vertical_velocity(og, inp)
    inner = lu.copy()
    inner[:2, :] = inner[-2:, :] = False
    inner[:, :2] = inner[:, -2:] = False
    total = (w[0] * b.dx * b.dy)[inner].sum()
    want = (wflux[inner] / b.rho0).sum() * dz.sum()
    scale = np.abs(inp["uu"]).max() * 30.0 * b.dyh.max() * dz.sum() * inner.sum()
    assert abs(total - want) <= 1e-12 * scale
    inp0 = dict(inp, uu=np.zeros_like(inp["uu"]), vv=np.zeros_like(inp["vv"]))
    w0 = oracle.vertical_velocity(og, inp0)
    cum = np.concatenate([np.cumsum(dz[::-1])[::-1], [0.0]])
    rate = wflux / b.rho0 / (b.dx * b.dy)
    for k in range(nz + 1):
        assert np.allclose(w0[k][inner], (cum[k] * rate * b.dx * b.dy / (b.dx * b.dy))[inner], rtol=1e-12,
                           atol=1e-30)


def test_pressure_gradients_closed_forms():
    """Hydrostatics on sigma levels: a density that is the same in every
    cell leaves the right-hand sides unchanged whatever the depths (the
    sigma-coordinate pressure-gradient error of a uniform fluid is zero);
    with equal depths h and a density that differs only between the two
    columns, rhs(k) drops by g/RefDen * h * z(k) * (den1 - den0) * hhu * dyh
    (hhv * dxh for y), the hydrostatic pressure difference at depth h z(k)."""
    b = make_basin(40, 30, seed=9, sigma=3.0, metrics="stretched")
    og = oracle.OracleGrid(b)
    rng = np.random.default_rng(9)
    nz, sh = 7, (b.ny, b.nx)
    lcu, lcv = (og.masks[k] > 0.5 for k in ("lcu", "lcv"))
    z = np.sort(rng.random(nz))
    rhsx, rhsy = rng.standard_normal((nz,) + sh), rng.standard_normal((nz,) + sh)
    hhu, hhv = 20.0 + rng.random(sh), 20.0 + rng.random(sh)
    uniform = {"z": z, "den": np.full((nz,) + sh, 1027.5), "rhsx": rhsx, "rhsy": rhsy,
               "hhq": 10.0 + 40.0 * rng.random(sh), "hhu": hhu, "hhv": hhv}
    out = oracle.pressure_gradients(og, uniform)
    scale = 1027.5 * 50.0 * b.g / b.rho0 * 21.0 * max(b.dxh.max(), b.dyh.max())
    for k in ("rhsx", "rhsy"):
        assert np.abs(out[k] - uniform[k]).max() <= 1e-13 * scale
    h = 37.0
    den_col = 1025.0 + 3.0 * rng.standard_normal(sh)
    tilt = {"z": z, "den": np.broadcast_to(den_col, (nz,) + sh).copy(), "rhsx": rhsx, "rhsy": rhsy,
            "hhq": np.full(sh, h), "hhu": hhu, "hhv": hhv}
    out = oracle.pressure_gradients(og, tilt)
    dx_den = np.zeros(sh)
    dx_den[:, :-1] = den_col[:, 1:] - den_col[:, :-1]
    dy_den = np.zeros(sh)
    dy_den[:-1, :] = den_col[1:, :] - den_col[:-1, :]
    for k in range(nz):
        want_x = rhsx[k] - b.g / b.rho0 * h * z[k] * dx_den * hhu * b.dyh
        want_y = rhsy[k] - b.g / b.rho0 * h * z[k] * dy_den * hhv * b.dxh
        assert np.allclose(out["rhsx"][k][lcu], want_x[lcu], rtol=1e-11, atol=1e-11 * scale)
        assert np.allclose(out["rhsy"][k][lcv], want_y[lcv], rtol=1e-11, atol=1e-11 * scale)
        assert np.array_equal(out["rhsx"][k][~lcu], rhsx[k][~lcu])
        assert np.array_equal(out["rhsy"][k][~lcv], rhsy[k][~lcv])


def test_density_special_cases():
    """The equation of state where its form fixes the answer: at t = 0 and
    s = 0 every pressure term vanishes (the pure-pressure term is left out),
    so den = den_pot = 1000 x -2.00920601e-2 for any depth; at t = 0 it is
    affine in s at any pressure (equal steps in s give equal steps in den);
    den_pot is den at zero depth; and with t, s fixed den depends on the
    depth only through the pressure g RefDen hhq z, so two columns with the
    same hhq z agree.  Only lu points of m, n = 2..n-1 are written."""
    b = make_basin(30, 24, seed=5, sigma=3.0)
    og = oracle.OracleGrid(b)
    rng = np.random.default_rng(5)
    nz, sh = 4, (b.ny, b.nx)
    lu = og.masks["lu"] > 0.5
    inner = lu.copy()
    inner[0, :] = inner[-1, :] = False
    inner[:, 0] = inner[:, -1] = False
    z = np.array([0.1, 0.35, 0.6, 0.9])
    hhq = 20.0 + 80.0 * rng.random(sh)
    zero = np.zeros((nz,) + sh)
    out = oracle.density(og, {"z": z, "tt": zero, "ss": zero, "hhq": hhq, "den": np.full((nz,) + sh, 5.0),
                              "den_pot": np.full((nz,) + sh, 5.0)})
    c0 = -2.00920601e-02 * 1000.0
    for k in ("den", "den_pot"):
        assert np.all(out[k][:, inner] == c0)
        assert np.all(out[k][:, ~inner] == 5.0)
    s1 = 10.0 + 20.0 * rng.random((nz,) + sh)
    res = [oracle.density(og, {"z": z, "tt": zero, "ss": a * s1, "hhq": hhq, "den": zero, "den_pot": zero})
           for a in (0.0, 1.0, 2.0)]
    step1 = res[1]["den"] - res[0]["den"]
    step2 = res[2]["den"] - res[1]["den"]
    assert np.allclose(step1[:, inner], step2[:, inner], rtol=1e-12, atol=1e-12)
    tt = 25.0 * rng.random((nz,) + sh)
    ss = 35.0 * rng.random((nz,) + sh)
    surf = oracle.density(og, {"z": np.zeros(nz), "tt": tt, "ss": ss, "hhq": hhq, "den": zero, "den_pot": zero})
    assert np.array_equal(surf["den"][:, inner], surf["den_pot"][:, inner])
    full = oracle.density(og, {"z": z, "tt": tt, "ss": ss, "hhq": hhq, "den": zero, "den_pot": zero})
    assert np.array_equal(full["den_pot"], surf["den_pot"])
    assert np.all(np.abs(full["den"] - full["den_pot"])[:, inner] > 0.0)
    # same pressure in two columns: hhq z equal when hhq doubles and z halves
    col = {"tt": np.broadcast_to(tt[:, 5:6, 5:6], (nz,) + sh).copy(),
           "ss": np.broadcast_to(ss[:, 5:6, 5:6], (nz,) + sh).copy()}
    a = oracle.density(og, dict(col, z=z, hhq=np.full(sh, 40.0), den=zero, den_pot=zero))
    b2 = oracle.density(og, dict(col, z=z * 0.5, hhq=np.full(sh, 80.0), den=zero, den_pot=zero))
    assert np.allclose(a["den"][:, inner], b2["den"][:, inner], rtol=0, atol=1e-12)


@pytest.mark.parametrize("periodicity", [(0, 0), (1, 1)])
def test_smagorinsky_closed_forms(periodicity):
    """Smagorinsky viscosity (Function/mixing.f90:171-219) and its depth
    integral (Function/depth.f90:223-254) from the formula: with no strain
    amuv = visc2 and amts = dift exactly; a single shear value s at the H
    corner (m0, n0) raises exactly the four T cells m0..m0+1 x n0..n0+1 by
    mu = (0.4/pi)^2 dx dy |s| / 2 (a stencil that looks at the wrong corners
    raises other cells), amts by mu / 10; a single tension value t raises its
    own cell by (0.4/pi)^2 dx dy |t|; amuv2d = sum_k amuv dz on the lu cells
    of 2..nx-1 x 2..ny-1; cells outside those regions keep their values and
    the periodic wrap copies the interior's opposite edge."""
    b = make_basin(40, 30, seed=10, sigma=3.0, metrics="stretched")
    if periodicity != (0, 0):
        b.lbasins[:, :] = 0
    og = oracle.OracleGrid(b, periodicity=periodicity)
    rng = np.random.default_rng(10)
    nz, sh = 5, (b.ny, b.nx)
    lu = og.masks["lu"] > 0.5
    inner = np.zeros(sh, dtype=bool)
    inner[2:-2, 2:-2] = True
    dz = 0.5 + rng.random(nz)
    visc2, dift = 1.0e3, 5.0e2
    coef = (0.4 / math.pi) ** 2
    base = {"dz": dz, "stress_t": np.zeros((nz,) + sh), "stress_s": np.zeros((nz,) + sh),
            "amts": rng.random((nz,) + sh), "amuv": rng.random((nz,) + sh), "amuv2d": rng.random(sh)}
    calm = oracle.smagorinsky(og, base, visc2, dift)
    comp = inner & lu
    assert np.all(calm["amuv"][:, comp] == visc2) and np.all(calm["amts"][:, comp] == dift)
    m0, n0, k0, s, t = 17, 12, 3, 0.37, -0.21           # 1-based (m, n) of the corner / cell
    spike = dict(base, stress_s=base["stress_s"].copy(), stress_t=base["stress_t"].copy())
    spike["stress_s"][k0, n0 - 1, m0 - 1] = s
    spike["stress_t"][k0 - 2, 5, 8] = t
    out = oracle.smagorinsky(og, spike, visc2, dift)
    want_uv = calm["amuv"].copy()
    want_ts = calm["amts"].copy()
    for m, n in ((m0, n0), (m0 + 1, n0), (m0, n0 + 1), (m0 + 1, n0 + 1)):
        if comp[n - 1, m - 1]:
            mu = coef * b.dx[n - 1, m - 1] * b.dy[n - 1, m - 1] * abs(s) / 2
            want_uv[k0, n - 1, m - 1] = visc2 + mu
            want_ts[k0, n - 1, m - 1] = dift + mu / 10
    assert comp[5, 8]
    want_uv[k0 - 2, 5, 8] = visc2 + coef * b.dx[5, 8] * b.dy[5, 8] * abs(t)
    want_ts[k0 - 2, 5, 8] = dift + coef * b.dx[5, 8] * b.dy[5, 8] * abs(t) / 10
    assert np.allclose(out["amuv"], want_uv, rtol=1e-14, atol=0)
    assert np.allclose(out["amts"], want_ts, rtol=1e-14, atol=0)
    assert np.count_nonzero(out["amuv"] != calm["amuv"]) == 1 + sum(
        comp[n - 1, m - 1] for m, n in ((m0, n0), (m0 + 1, n0), (m0, n0 + 1), (m0 + 1, n0 + 1)))
    ring = np.zeros(sh, dtype=bool)
    ring[1:-1, 1:-1] = True
    ring &= lu
    for o in (calm, out):
        assert np.allclose(o["amuv2d"][ring], np.tensordot(dz, o["amuv"], axes=1)[ring], rtol=1e-14, atol=0)
        assert np.array_equal(o["amuv2d"][~ring], base["amuv2d"][~ring])
    if periodicity == (0, 0):
        keep = ~comp
        assert np.array_equal(out["amuv"][:, keep], base["amuv"][:, keep])
        assert np.array_equal(out["amts"][:, keep], base["amts"][:, keep])
    else:
        for f in ("amuv", "amts"):
            a = out[f]
            assert np.array_equal(a[:, :, 1], a[:, :, -3]) and np.array_equal(a[:, :, -2], a[:, :, 2])
            assert np.array_equal(a[:, 1, :], a[:, -3, :]) and np.array_equal(a[:, -2, :], a[:, 2, :])


@pytest.mark.parametrize("periodicity", [(1, 0), (1, 1)])
def test_periodic_wrap_equals_doubled_domain(periodicity):
    """Periodic wrap (Service/oc_algorithms.f90:50-135, as a decomposed run
    applies it: column mm is copied to mmm-1 and mmm to mm+1, likewise for
    rows): a periodic basin of width W advanced N steps equals, bit for bit,
    each copy inside a basin of width 2W (and height 2H when y is periodic
    too) holding the W basin's fields tiled twice, where the copies meet
    without a wrap.  A wrap that copies a column to its own neighbour, or is
    off by one, breaks the equality at the seam.  Bathymetry, state and
    water flux vary along x (no symmetry the wrap could hide behind)."""
    b, b2, (tx, ty) = periodic_pair(periodicity)
    steps = 15
    rd1 = oracle.OracleGrid(b, periodicity).rayleigh()
    rd2 = oracle.OracleGrid(b2, periodicity).rayleigh()
    o1 = oracle.run(b, steps, 2, out_steps=[steps], rdis=rd1, periodicity=periodicity)[steps]
    o2 = oracle.run(b2, steps, 2, out_steps=[steps], rdis=rd2, periodicity=periodicity)[steps]
    W, H = b.nx - 4, b.ny - 4
    for k in ("ssh", "sshp", "u", "up", "v", "vp", "ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady"):
        a = o1[k][2:-2, 2:-2]
        for i in range(tx):
            for j in range(ty):
                c = o2[k][2 + j * H:2 + (j + 1) * H, 2 + i * W:2 + (i + 1) * W]
                assert np.array_equal(a.view(np.uint64), c.view(np.uint64)), (k, i, j)
    assert np.abs(o1["u"][2:-2, 2:-2]).max() > 0


def test_reference_grid_closed_forms():
    """The restated grid set-up in the forms its definitions fix:
    Cartesian: every metric is dxst (dyst) * pi/180 * R on points
    2..nx-1 x 2..ny-1 and zero outside, rlh_s = 2 Omega / sqrt(2) there and
    2 Omega outside; spherical: y-metrics as Cartesian, x-metrics of T and U
    rows scaled by the cosine of the row latitude and of V and H rows by
    the cosine half a step north (so dx/dxh is the ratio of those cosines),
    rlh_s = 2 Omega sin(latitude of the H row)."""
    from inmsom_bt.cases import EARTH_OMEGA, EARTH_RADIUS, reference_grid

    nx, ny, lon0, lat0, dl, dp = 20, 14, 10.0, 30.0, 0.5, 0.25
    pif = float(np.float32(math.pi))
    c = reference_grid(nx, ny, lon0, lat0, dl, dp, curve_grid=0)
    inner = np.zeros((ny, nx), dtype=bool)
    inner[1:-1, 1:-1] = True
    for k in ("dx", "dxt", "dxh", "dxb"):
        np.testing.assert_allclose(c[k][inner], dl * pif / 180.0 * EARTH_RADIUS, rtol=1e-15)
        assert np.all(c[k][~inner] == 0.0)
    for k in ("dy", "dyt", "dyh", "dyb"):
        np.testing.assert_allclose(c[k][inner], dp * pif / 180.0 * EARTH_RADIUS, rtol=1e-15)
    np.testing.assert_allclose(c["rlh_s"][inner], 2 * EARTH_OMEGA / math.sqrt(2.0), rtol=1e-15)
    assert np.all(c["rlh_s"][~inner] == 2 * EARTH_OMEGA)
    s = reference_grid(nx, ny, lon0, lat0, dl, dp, curve_grid=1)
    lat_t = lat0 + (np.arange(ny) - 2) * dp
    lat_v = lat_t + dp / 2
    rad = pif / 180.0
    for k in ("dy", "dyt", "dyh", "dyb"):
        np.testing.assert_array_equal(s[k], c[k])
    for k, lat in (("dx", lat_t), ("dxt", lat_t), ("dxh", lat_v), ("dxb", lat_v)):
        want = c[k] * np.cos(lat * rad)[:, None]
        np.testing.assert_allclose(s[k][inner], want[inner], rtol=1e-14)
    np.testing.assert_allclose(s["rlh_s"][inner], (2 * EARTH_OMEGA * np.sin(lat_v * rad)[:, None]
                                                   * np.ones((1, nx)))[inner], rtol=1e-14)


def test_rayleigh_stretched_metrics_and_coasts():
    """Rayleigh coefficient (Control/init_pars.f90:459-471) on a grid whose
    eight metrics all differ, with a depth quadratic in the distance along
    x and y built from dxt and dyt alone: the slope on the U face between m
    and m+1 is then a (X_m + X_{m+1}) and on the V face b (Y_n + Y_{n+1}).
    Interior cells average the squared slopes of their two faces per
    direction, coastal cells use the one open face: a metric other than
    dxt / dyt, a face taken from the wrong side or lcu and lcv exchanged
    fail.  The basin has coasts with one and with both faces open."""
    b = _flat(22, 18, h=0.0)
    nyh, nxh = b.ny, b.nx
    mm = np.arange(nxh, dtype=np.float64)[None, :] * np.ones((nyh, 1))
    nn = np.arange(nyh, dtype=np.float64)[:, None] * np.ones((1, nxh))
    for k, name in enumerate(("dx", "dy", "dxt", "dyt", "dxh", "dyh", "dxb", "dyb")):
        setattr(b, name, 1000.0 * (1.0 + 0.05 * k + 0.2 * np.sin(0.7 * mm + k) * (name[1] == "x")
                                   + 0.2 * np.cos(0.9 * nn + k) * (name[1] == "y")))
    # dxt depends on m only, dyt on n only: distances X(m), Y(n) along the axes
    X = np.concatenate([[0.0], np.cumsum(b.dxt[0, :-1])])
    Y = np.concatenate([[0.0], np.cumsum(b.dyt[:-1, 0])])
    a, c = 2e-9, -3e-9
    wet = np.zeros((nyh, nxh), dtype=bool)
    wet[2:-2, 2:-2] = True
    wet[7:10, 9:12] = False        # an island: coasts with one open face
    wet[13, 5] = False              # a single land cell
    b.lbasins = np.where(wet, 0, 1).astype(np.int32)
    b.hhq_rest = np.where(wet, 50.0 + a * X[None, :] ** 2 + c * Y[:, None] ** 2, 0.0)
    r = oracle.OracleGrid(b).rayleigh(1e-4)
    lcu = wet[:, :-1] & wet[:, 1:]   # U face between m and m+1 (0-based m)
    lcv = wet[:-1, :] & wet[1:, :]
    checked = {1: 0, 2: 0}
    for j in range(2, nyh - 2):
        for i in range(2, nxh - 2):
            if not wet[j, i]:
                continue
            sx = ([(a * (X[i] + X[i + 1])) ** 2] * int(lcu[j, i]) +
                  [(a * (X[i - 1] + X[i])) ** 2] * int(lcu[j, i - 1]))
            sy = ([(c * (Y[j] + Y[j + 1])) ** 2] * int(lcv[j, i]) +
                  [(c * (Y[j - 1] + Y[j])) ** 2] * int(lcv[j - 1, i]))
            if not sx or not sy:
                continue
            want = 1e-4 * math.sqrt(sum(sx) / len(sx) + sum(sy) / len(sy))
            assert math.isclose(r[j, i], want, rel_tol=1e-10), (j, i, r[j, i], want)
            checked[min(len(sx), len(sy))] += 1
    assert checked[1] > 10 and checked[2] > 100, checked
