"""The dynamics of one 3-D model step in the reference's order
(Control/dyncall.f90:37-449 without the vertical mixing, tracers and
turbulence), every block on the GPU through the C ABI, against the same
chain of oracle routines, bit for bit at every hand-over: the density
definition -> pressure_gradients -> forcing of the barotropic step ->
barotropic_dynamics (the drop-in) -> removal of the barotropic part ->
vertical_velocity -> internal update (ssh_internal, re-add, pgrx) ->
baroclinic_dynamics."""
import ctypes as C

import numpy as np
import pytest

import oracle
from _util import assert_bitwise
from inmsom_bt import BarotropicModel, make_basin, randomize_state

pytestmark = pytest.mark.gpu

ACC = ("ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.mark.parametrize("nz,metrics", [(6, "uniform"), (20, "stretched")])
def test_dynamics_step_chain(nz, metrics):
    from test_gpu_forcing3d import _fields
    print(f"dynamics chain nz={nz} {metrics}", flush=True)
    b = randomize_state(make_basin(64, 48, seed=51, sigma=4.0, metrics=metrics), 7)
    og = oracle.OracleGrid(b)
    rd = og.rayleigh()
    rng = np.random.default_rng(12 + nz)
    sh = (b.ny, b.nx)
    s3 = (nz,) + sh
    f = _fields(b, nz, seed=nz + 3)
    wflux = np.ascontiguousarray(b.wflux, dtype=np.float64)
    z = np.sort(rng.random(nz))
    tt = 2.0 + 25.0 * rng.random(s3)
    ss = 5.0 + 30.0 * rng.random(s3)
    zero3, zero2 = np.zeros(s3), np.zeros(sh)
    coupling_in = {"slpr": f["slpr"], "wflux": wflux, "ssh_i": 0.2 * rng.standard_normal(sh) * b.wet,
                   "sshp_i": 0.2 * rng.standard_normal(sh) * b.wet}
    bc_in = {"uup": f["uup"], "vvp": f["vvp"], "rhsx_tran": 1e2 * rng.standard_normal(s3),
             "rhsy_tran": 1e2 * rng.standard_normal(s3), "nu": 1e-3 + 1e-2 * rng.random((nz + 1,) + sh),
             "rdis": rd, "tx_surf": 1e-4 * rng.standard_normal(sh), "ty_surf": 1e-4 * rng.standard_normal(sh),
             "tx_bot": 1e-5 * rng.standard_normal(sh), "ty_bot": 1e-5 * rng.standard_normal(sh),
             "hhup": f["hhu"] + 0.01, "hhvp": f["hhv"] - 0.01, "hhh": f["hhh"], "dz": f["dz"],
             "hzt": (0.5 + rng.random(nz + 1)) / nz}

    # ---- the oracle chain
    L = oracle.lib()
    eos = oracle.density(og, {"z": z, "tt": tt, "ss": ss, "hhq": f["hhq"], "den": zero3, "den_pot": zero3})
    pg = oracle.pressure_gradients(og, {"z": z, "den": eos["den"], "rhsx": zero3, "rhsy": zero3, "hhq": f["hhq"],
                                        "hhu": f["hhu"], "hhv": f["hhv"]})
    frc = oracle.forcing_3d(og, dict(f, pgx=pg["rhsx"], pgy=pg["rhsy"], rhsx2d=zero2, rhsy2d=zero2), b.cb_nl,
                            b.ebottom)
    st_o = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
    acc_o = {k: np.zeros(sh) for k in ACC}
    oracle.barotropic_dynamics(og, b.dt, 2, 0, st_o, acc_o, frc["rhsx2d"], frc["rhsy2d"], wflux, b.mu, b.mu4, rd)
    uu_r, vv_r = f["uu"].copy(), f["vv"].copy()
    u2_r, v2_r = f["uu2d"].copy(), f["vv2d"].copy()
    L.bto_remove_barotropic(C.byref(og.g), nz, _d(uu_r), _d(vv_r), _d(u2_r), _d(v2_r))
    ww_o = oracle.vertical_velocity(og, {"dz": f["dz"], "uu": uu_r, "vv": vv_r, "ww": np.zeros((nz + 1,) + sh),
                                         "hu": f["hhu"], "hv": f["hhv"], "wflux": wflux})
    cpl_o = oracle.coupling(og, b.dt, dict(coupling_in, uu=f["uu"], vv=f["vv"], uu2d=f["uu2d"], vv2d=f["vv2d"],
                                           hhu=f["hhu"], hhv=f["hhv"], ubrtr_i=acc_o["ubrtr_i"],
                                           vbrtr_i=acc_o["vbrtr_i"], pgrx=acc_o["ssh4gradx"],
                                           pgry=acc_o["ssh4grady"]))
    bc_o = oracle.baroclinic(og, b.dt, dict(bc_in, uu=cpl_o["uu"], vv=cpl_o["vv"], ww=ww_o, rhsx=pg["rhsx"],
                                            rhsy=pg["rhsy"], pgrx=cpl_o["pgrx"], pgry=cpl_o["pgry"],
                                            hhu=f["hhu"], hhun=cpl_o["hhun"], hhv=f["hhv"], hhvn=cpl_o["hhvn"]))

    # ---- the same chain on the GPU
    with BarotropicModel(b, rdis=rd) as m:
        den, den_pot = m.density(z, tt, ss, f["hhq"], zero3, zero3)
        pgx, pgy = m.pressure_gradients(z, den, zero3, zero3, f["hhq"], f["hhu"], f["hhv"])
        frc_g = m.forcing_3d(dict(f, taux=b.taux, tauy=b.tauy), f["dz"], zero2, zero2, pgx, pgy)
        st_g = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
        acc_g = {k: np.zeros(sh) for k in ACC}
        m.barotropic_dynamics(b.dt, 2, 0, st_g["u"], st_g["up"], st_g["v"], st_g["vp"], st_g["ssh"], st_g["sshp"],
                              acc_g["ubrtr_i"], acc_g["vbrtr_i"], acc_g["ssh4gradx"], acc_g["ssh4grady"],
                              frc_g["rhsx2d"], frc_g["rhsy2d"], wflux, np.ascontiguousarray(b.mu),
                              np.ascontiguousarray(b.mu4), rd)
        uu, vv = f["uu"].copy(), f["vv"].copy()
        u2, v2 = f["uu2d"].copy(), f["vv2d"].copy()
        m.remove_barotropic(uu, vv, u2, v2)
        ww = m.vertical_velocity(f["dz"], uu, vv, np.zeros((nz + 1,) + sh), f["hhu"], f["hhv"], wflux)
        cpl = m.internal_update(b.dt, dict(coupling_in, uu2d=u2, vv2d=v2, hhu=f["hhu"], hhv=f["hhv"],
                                           ubrtr_i=acc_g["ubrtr_i"], vbrtr_i=acc_g["vbrtr_i"],
                                           pgrx=acc_g["ssh4gradx"], pgry=acc_g["ssh4grady"], hhqn=zero2,
                                           hhun=zero2, hhvn=zero2, hhhn=zero2), uu, vv)
        bc = m.baroclinic_dynamics(b.dt, dict(bc_in, uu=uu, vv=vv, ww=ww, rhsx=pgx, rhsy=pgy, pgrx=cpl["pgrx"],
                                              pgry=cpl["pgry"], hhu=f["hhu"], hhun=cpl["hhun"], hhv=f["hhv"],
                                              hhvn=cpl["hhvn"]))

    assert_bitwise({"den": den, "den_pot": den_pot}, eos, ("den", "den_pot"), "density")
    assert_bitwise({"pgx": pgx, "pgy": pgy}, {"pgx": pg["rhsx"], "pgy": pg["rhsy"]}, ("pgx", "pgy"), "pressure")
    assert_bitwise(frc_g, frc, ("rhsx2d", "rhsy2d"), "forcing")
    assert_bitwise(st_g, st_o, tuple(st_o), "barotropic state")
    assert_bitwise(acc_g, acc_o, ACC, "barotropic time means")
    assert_bitwise({"ww": ww}, {"ww": ww_o}, ("ww",), "vertical velocity")
    assert_bitwise(dict(cpl, uu=uu, vv=vv), cpl_o, ("uu", "vv", "uu2d", "vv2d", "pgrx", "pgry", "ssh_i", "sshp_i",
                                                   "hhqn", "hhun", "hhvn", "hhhn"), "coupling")
    assert_bitwise(bc, bc_o, ("uu", "uup", "vv", "vvp"), "baroclinic")
    assert np.abs(bc["uu"] - cpl_o["uu"]).max() > 0.0
