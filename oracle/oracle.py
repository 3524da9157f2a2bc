"""ctypes front-end of the C restatement in bt_oracle.c.

TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.

Arrays follow inmsom_bt.cases: float64 (ny, nx) C order == the
reference's column-major (m, n).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

MASKS = ("lu", "luu", "luh", "lcu", "lcv", "llu", "llv")
METRICS = ("dx", "dy", "dxt", "dyt", "dxh", "dyh", "dxb", "dyb")
DEPTHS = ("hhq_e", "hhqp_e", "hhqn_e", "hhu_e", "hhup_e", "hhun_e",
          "hhv_e", "hhvp_e", "hhvn_e", "hhh_e", "hhhp_e", "hhhn_e")
OUT_FIELDS = ("ssh", "sshp", "u", "up", "v", "vp", "ubrtr_i", "vbrtr_i",
              "ssh4gradx", "ssh4grady")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class _Grid(C.Structure):
    _fields_ = ([("nx", C.c_int), ("ny", C.c_int), ("periodicity_x", C.c_int),
                 ("periodicity_y", C.c_int), ("full_free_surface", C.c_int),
                 ("free_fall_acc", C.c_double), ("ref_den", C.c_double),
                 ("time_smooth", C.c_double)]
                + [(k, _fp) for k in MASKS]
                + [(k, _dp) for k in METRICS]
                + [("rlh_s", _dp), ("hhq_rest", _dp)])


class _Depths(C.Structure):
    _fields_ = [("h", _dp * 12)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
        _lib = C.CDLL(LIB_PATH)
        _lib.bto_volume.restype = C.c_double
    return _lib


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _f(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def gridcon(lbasins: np.ndarray, px: int = 0, py: int = 0) -> Dict[str, np.ndarray]:
    """Service/grid_construction.f90:33-112."""
    ny, nx = lbasins.shape
    lb = np.ascontiguousarray(lbasins, dtype=np.int32)
    out = {k: np.zeros((ny, nx), dtype=np.float32) for k in MASKS}
    lib().bto_gridcon(nx, ny, px, py, lb.ctypes.data_as(_ip), *[_f(out[k]) for k in MASKS])
    return out


class OracleGrid:
    """Static fields of one basin, held for the C oracle."""

    def __init__(self, b, periodicity=(0, 0), full_free_surface=1):
        self.b = b
        self.masks = gridcon(b.lbasins, *periodicity)
        self.keep = {k: np.ascontiguousarray(getattr(b, k), dtype=np.float64)
                     for k in METRICS + ("rlh_s", "hhq_rest")}
        g = _Grid()
        g.nx, g.ny = b.nx, b.ny
        g.periodicity_x, g.periodicity_y = periodicity
        g.full_free_surface = full_free_surface
        g.free_fall_acc, g.ref_den, g.time_smooth = b.g, b.rho0, b.time_smooth
        for k in MASKS:
            setattr(g, k, _f(self.masks[k]))
        for k, a in self.keep.items():
            setattr(g, k, _d(a))
        self.g = g
        self.ref = C.byref(g)

    def rayleigh(self, r_fric: Optional[float] = None) -> np.ndarray:
        out = np.zeros((self.b.ny, self.b.nx))
        lib().bto_rayleigh_init(self.ref, C.c_double(self.b.r_fric if r_fric is None else r_fric),
                                _d(out))
        return out

    def volume(self, ssh: np.ndarray) -> float:
        return float(lib().bto_volume(self.ref, _d(np.ascontiguousarray(ssh))))

    def new_depths(self):
        arrs = [np.zeros((self.b.ny, self.b.nx)) for _ in range(12)]
        d = _Depths()
        for i, a in enumerate(arrs):
            d.h[i] = _d(a)
        return d, arrs


def run(b, nsteps: int, nstep: int = 1, ksw4: int = 0, out_steps: Sequence[int] = (),
        rdis: Optional[np.ndarray] = None, periodicity=(0, 0), grid: Optional[OracleGrid] = None):
    """Time loop of bto_model_step; returns {step: {field: array}} for the
    steps in out_steps (fields: OUT_FIELDS + DEPTHS)."""
    L = lib()
    og = grid or OracleGrid(b, periodicity)
    if rdis is None:
        rdis = og.rayleigh()
    st = {k: np.ascontiguousarray(getattr(b, k), dtype=np.float64).copy()
          for k in ("ssh", "sshp", "u", "up", "v", "vp")}
    acc = {k: np.zeros((b.ny, b.nx)) for k in ("ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")}
    dstruct, darr = og.new_depths()
    L.bto_hh_init(og.ref, C.byref(dstruct), _d(st["ssh"]), _d(st["sshp"]))
    work = np.zeros(16 * b.nx * b.ny)
    forc = [np.ascontiguousarray(getattr(b, k), dtype=np.float64)
            for k in ("taux", "tauy", "wflux", "mu", "mu4")]
    rd = np.ascontiguousarray(rdis, dtype=np.float64)
    outs = {}
    want = set(out_steps)
    for step in range(1, nsteps + 1):
        L.bto_model_step(og.ref, C.byref(dstruct), C.c_double(b.dt), nstep, ksw4,
                         C.c_double(b.cb_nl), C.c_double(b.ebottom),
                         _d(st["u"]), _d(st["up"]), _d(st["v"]), _d(st["vp"]),
                         _d(st["ssh"]), _d(st["sshp"]),
                         _d(acc["ubrtr_i"]), _d(acc["vbrtr_i"]), _d(acc["ssh4gradx"]),
                         _d(acc["ssh4grady"]), *[_d(a) for a in forc], _d(rd), _d(work))
        if step in want:
            o = {k: v.copy() for k, v in st.items()}
            o.update({k: v.copy() for k, v in acc.items()})
            o.update({k: a.copy() for k, a in zip(DEPTHS, darr)})
            outs[step] = o
    return outs


def barotropic_dynamics(og: OracleGrid, tau: float, nstep: int, ksw4: int, st: dict, acc: dict,
                        rhsx, rhsy, wflux, mu, mu4, rdis) -> Dict[str, np.ndarray]:
    """One call of Function/vel_ssh.f90:379-679 on depths hh_init(ssh, sshp);
    st (u, up, v, vp, ssh, sshp) and acc are updated in place; returns the
    depths after the call."""
    L = lib()
    b = og.b
    ds, da = og.new_depths()
    L.bto_hh_init(og.ref, C.byref(ds), _d(st["ssh"]), _d(st["sshp"]))
    z = lambda: np.zeros((b.ny, b.nx))
    scratch = [z() for _ in range(9)]
    ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (rhsx, rhsy, wflux, mu, mu4)]
    rd = np.ascontiguousarray(rdis, dtype=np.float64)
    L.bto_barotropic_dynamics(og.ref, C.byref(ds), C.c_double(tau), nstep, ksw4,
                              _d(st["u"]), _d(st["up"]), _d(st["v"]), _d(st["vp"]), _d(st["ssh"]),
                              _d(st["sshp"]), _d(acc["ubrtr_i"]), _d(acc["vbrtr_i"]),
                              _d(acc["ssh4gradx"]), _d(acc["ssh4grady"]), *[_d(a) for a in ins],
                              _d(scratch[0]), _d(scratch[1]), _d(scratch[2]), _d(scratch[3]),
                              _d(scratch[4]), _d(rd), _d(scratch[5]), _d(scratch[6]),
                              _d(scratch[7]), _d(scratch[8]))
    return {k: a.copy() for k, a in zip(DEPTHS, da)}


def units(b, nstep: int = 1, rdis: Optional[np.ndarray] = None) -> Dict[str, np.ndarray]:
    """The helper-routine sequence of the harness's mode 1."""
    L = lib()
    og = OracleGrid(b)
    nx, ny = b.nx, b.ny
    z = lambda: np.zeros((ny, nx))
    ssh = np.ascontiguousarray(b.ssh).copy()
    sshp = np.ascontiguousarray(b.sshp).copy()
    u, up = np.ascontiguousarray(b.u).copy(), np.ascontiguousarray(b.up).copy()
    v, vp = np.ascontiguousarray(b.v).copy(), np.ascontiguousarray(b.vp).copy()
    ds, da = og.new_depths()
    out = {}
    L.bto_hh_init(og.ref, C.byref(ds), _d(ssh), _d(sshp))
    out.update({"init_" + k: a.copy() for k, a in zip(DEPTHS, da)})
    L.bto_hh_update(og.ref, _d(da[2]), _d(da[5]), _d(da[8]), _d(da[11]), _d(u))
    for k, i in (("hhqn_e", 2), ("hhun_e", 5), ("hhvn_e", 8), ("hhhn_e", 11)):
        out["update_" + k] = da[i].copy()
    L.bto_hh_shift(og.ref, C.byref(ds), nstep)
    out.update({"shift_" + k: a.copy() for k, a in zip(DEPTHS, da)})
    strt, strs = z(), z()
    L.bto_stress_components(og.ref, _d(up), _d(vp), _d(strt), _d(strs))
    out["str_t"], out["str_s"] = strt.copy(), strs.copy()
    vort, radv, rady = z(), z(), z()
    L.bto_uv_trans(og.ref, _d(u), _d(v), _d(vort), _d(da[0]), _d(da[3]), _d(da[6]), _d(da[9]),
                   _d(radv), _d(rady))
    out["vort"], out["rhsx_adv"], out["rhsy_adv"] = vort, radv, rady
    rdfx, rdfy = z(), z()
    mu = np.ascontiguousarray(b.mu)
    L.bto_uv_diff2(og.ref, _d(mu), _d(strt), _d(strs), _d(da[0]), _d(da[3]), _d(da[6]),
                   _d(da[9]), _d(rdfx), _d(rdfy))
    out["rhsx_dif2"], out["rhsy_dif2"] = rdfx.copy(), rdfy.copy()
    fx, fy = z(), z()
    mu4 = np.ascontiguousarray(b.mu4)
    L.bto_uv_diff4(og.ref, _d(mu4), _d(strt), _d(strs), _d(fx), _d(fy), _d(da[0]), _d(da[3]),
                   _d(da[6]), _d(da[9]), _d(rdfx), _d(rdfy))
    out.update({"d4_fx": fx, "d4_fy": fy, "d4_str_t": strt.copy(), "d4_str_s": strs.copy(),
                "rhsx_dif4": rdfx.copy(), "rhsy_dif4": rdfy.copy()})
    sx, sy, bf, bx, by = z(), z(), z(), z(), z()
    taux, tauy = np.ascontiguousarray(b.taux), np.ascontiguousarray(b.tauy)
    L.bto_surface_stress(og.ref, _d(taux), _d(tauy), _d(sx), _d(sy))
    L.bto_bottom_stress(og.ref, C.c_double(b.cb_nl), C.c_double(b.ebottom), _d(u), _d(v),
                        _d(up), _d(vp), _d(da[3]), _d(da[6]), _d(da[0]), _d(bf), _d(bx), _d(by))
    out.update({"surf_x": sx, "surf_y": sy, "bottom_friction": bf, "bot_x": bx, "bot_y": by})
    return out


# ------------------------------------------------------- reference harness
UNIT_LAYOUT = ([("init_" + k) for k in DEPTHS]
               + ["update_hhqn_e", "update_hhun_e", "update_hhvn_e", "update_hhhn_e"]
               + [("shift_" + k) for k in DEPTHS]
               + ["str_t", "str_s", "vort", "rhsx_adv", "rhsy_adv", "rhsx_dif2", "rhsy_dif2",
                  "d4_fx", "d4_fy", "d4_str_t", "d4_str_s", "rhsx_dif4", "rhsy_dif4",
                  "surf_x", "surf_y", "bottom_friction", "bot_x", "bot_y"])


# 2-D -> 3-D coupling (Control/dyncall.f90:169-231, ssh_internal)
COUPLING_IN = ("ubrtr_i", "vbrtr_i", "hhu", "hhv", "slpr", "wflux", "uu2d", "vv2d", "pgrx", "pgry",
               "ssh_i", "sshp_i")
COUPLING_OUT = ("uu2d", "vv2d", "pgrx", "pgry", "ssh_i", "sshp_i", "hhqn", "hhun", "hhvn", "hhhn")


def coupling(og: "OracleGrid", tau: float, inp: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """bto_remove_barotropic then bto_internal_update on copies of ``inp``
    (the fields of COUPLING_IN plus 3-D "uu", "vv" of shape (nz, ny, nx))."""
    L = lib()
    w = {k: np.ascontiguousarray(inp[k], dtype=np.float64).copy() for k in inp}
    nz = w["uu"].shape[0]
    hn = {k: np.zeros_like(w["ssh_i"]) for k in ("hhqn", "hhun", "hhvn", "hhhn")}
    L.bto_remove_barotropic(C.byref(og.g), nz, _d(w["uu"]), _d(w["vv"]), _d(w["uu2d"]), _d(w["vv2d"]))
    L.bto_internal_update(C.byref(og.g), C.c_double(tau), nz, _d(w["ubrtr_i"]), _d(w["vbrtr_i"]), _d(w["hhu"]),
                          _d(w["hhv"]), _d(w["slpr"]), _d(w["wflux"]), _d(w["uu2d"]), _d(w["vv2d"]),
                          _d(w["pgrx"]), _d(w["pgry"]), _d(w["ssh_i"]), _d(w["sshp_i"]), _d(hn["hhqn"]),
                          _d(hn["hhun"]), _d(hn["hhvn"]), _d(hn["hhhn"]), _d(w["uu"]), _d(w["vv"]))
    out = {k: w[k] for k in ("uu", "vv", "uu2d", "vv2d", "pgrx", "pgry", "ssh_i", "sshp_i")}
    out.update(hn)
    return out


# forcing of the barotropic step (Control/dyncall.f90:37-130)
FORCING_IN_3D = ("uu", "vv", "uup", "vvp", "amuv", "pgx", "pgy")
FORCING_IN_2D = ("uu2d", "vv2d", "uup2d", "vvp2d", "amuv2d", "hhq", "hhu", "hhv", "hhh", "slpr", "rhsx2d",
                 "rhsy2d")


def forcing_3d(og: "OracleGrid", inp: Dict[str, np.ndarray], cb_nl: float, ebottom: float,
               pressure: bool = True) -> Dict[str, np.ndarray]:
    """bto_forcing_3d on copies of ``inp`` (dz, FORCING_IN_3D as (nz, ny, nx),
    FORCING_IN_2D, and the basin's taux/tauy); returns rhsx2d, rhsy2d."""
    L = lib()
    b = og.b
    w = {k: np.ascontiguousarray(inp[k], dtype=np.float64).copy() for k in inp}
    nz = w["uu"].shape[0]
    taux = np.ascontiguousarray(b.taux, dtype=np.float64)
    tauy = np.ascontiguousarray(b.tauy, dtype=np.float64)
    pg = (_d(w["pgx"]), _d(w["pgy"])) if pressure else (None, None)
    L.bto_forcing_3d(C.byref(og.g), nz, _d(w["dz"]), C.c_double(cb_nl), C.c_double(ebottom),
                     *[_d(w[k]) for k in ("uu", "vv", "uup", "vvp", "amuv", "uu2d", "vv2d", "uup2d", "vvp2d",
                                          "amuv2d", "hhq", "hhu", "hhv", "hhh")],
                     _d(taux), _d(tauy), _d(w["slpr"]), *pg, _d(w["rhsx2d"]), _d(w["rhsy2d"]))
    return {"rhsx2d": w["rhsx2d"], "rhsy2d": w["rhsy2d"]}


BAROCLINIC_IN_3D = ("uu", "uup", "vv", "vvp", "rhsx", "rhsy", "rhsx_tran", "rhsy_tran")
BAROCLINIC_IN_W = ("ww", "nu")                      # nz + 1 levels
BAROCLINIC_IN_2D = ("pgrx", "pgry", "rdis", "tx_surf", "ty_surf", "tx_bot", "ty_bot", "hhu", "hhup", "hhun",
                    "hhv", "hhvp", "hhvn", "hhh")
BAROCLINIC_OUT = ("uu", "uup", "vv", "vvp")


def baroclinic(og: "OracleGrid", tau: float, inp: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """bto_baroclinic_dynamics on copies of ``inp`` (dz, hzt, BAROCLINIC_IN_3D
    as (nz, ny, nx), ww/nu as (nz+1, ny, nx), BAROCLINIC_IN_2D); returns
    the updated uu, uup, vv, vvp."""
    L = lib()
    w = {k: np.ascontiguousarray(inp[k], dtype=np.float64).copy() for k in inp}
    nz = w["uu"].shape[0]
    L.bto_baroclinic_dynamics(C.byref(og.g), nz, C.c_double(tau), _d(w["dz"]), _d(w["hzt"]),
                              *[_d(w[k]) for k in ("uu", "uup", "vv", "vvp", "ww", "nu", "rhsx", "rhsy", "rhsx_tran",
                                                   "rhsy_tran") + BAROCLINIC_IN_2D])
    return {k: w[k] for k in BAROCLINIC_OUT}


def vertical_velocity(og: "OracleGrid", inp: Dict[str, np.ndarray]) -> np.ndarray:
    """bto_vertical_velocity on copies of ``inp`` (dz, uu, vv (nz, ny, nx),
    ww (nz+1, ny, nx), hu, hv, wflux); returns the new ww."""
    L = lib()
    w = {k: np.ascontiguousarray(inp[k], dtype=np.float64).copy() for k in inp}
    nz = w["uu"].shape[0]
    L.bto_vertical_velocity(C.byref(og.g), nz, _d(w["dz"]), _d(w["uu"]), _d(w["vv"]), _d(w["ww"]), _d(w["hu"]),
                            _d(w["hv"]), _d(w["wflux"]))
    return w["ww"]


def stress_components_3d(og: "OracleGrid", u: np.ndarray, v: np.ndarray, str_t: np.ndarray,
                         str_s: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Function/mixing.f90:115-168 with nlev = nz levels (u, v, str_t, str_s
    (nz, ny, nx)): every level is the one-level bto_stress_components (the
    routine's loops and its wrap are per level); cells it does not write keep
    the values of the str_t, str_s passed.  Returns new arrays."""
    L = lib()
    uc = np.ascontiguousarray(u, dtype=np.float64)
    vc = np.ascontiguousarray(v, dtype=np.float64)
    st = np.array(str_t, dtype=np.float64, order="C")
    ss = np.array(str_s, dtype=np.float64, order="C")
    for k in range(uc.shape[0]):
        t, s = np.ascontiguousarray(st[k]), np.ascontiguousarray(ss[k])
        L.bto_stress_components(og.ref, _d(uc[k]), _d(vc[k]), _d(t), _d(s))
        st[k], ss[k] = t, s
    return st, ss


def smagorinsky(og: "OracleGrid", inp: Dict[str, np.ndarray], visc2: float, dift: float) -> Dict[str, np.ndarray]:
    """bto_smagorinsky on copies of ``inp`` (dz, stress_t, stress_s, amts,
    amuv (nz, ny, nx), amuv2d); returns the new amts, amuv, amuv2d."""
    L = lib()
    w = {k: np.ascontiguousarray(inp[k], dtype=np.float64).copy() for k in inp}
    nz = w["amuv"].shape[0]
    L.bto_smagorinsky(C.byref(og.g), nz, C.c_double(visc2), C.c_double(dift), _d(w["stress_t"]),
                      _d(w["stress_s"]), _d(w["dz"]), _d(w["amts"]), _d(w["amuv"]), _d(w["amuv2d"]))
    return {k: w[k] for k in ("amts", "amuv", "amuv2d")}


def pressure_gradients(og: "OracleGrid", inp: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """bto_pressure_gradients on copies of ``inp`` (z, den, rhsx, rhsy as
    (nz, ny, nx), hhq, hhu, hhv); returns the updated rhsx, rhsy."""
    L = lib()
    w = {k: np.ascontiguousarray(inp[k], dtype=np.float64).copy() for k in inp}
    nz = w["den"].shape[0]
    L.bto_pressure_gradients(C.byref(og.g), nz, _d(w["z"]), _d(w["den"]), _d(w["hhq"]), _d(w["hhu"]),
                             _d(w["hhv"]), _d(w["rhsx"]), _d(w["rhsy"]))
    return {"rhsx": w["rhsx"], "rhsy": w["rhsy"]}


def density(og: "OracleGrid", inp: Dict[str, np.ndarray]) -> Dict[str, np.ndarray]:
    """bto_density on copies of ``inp`` (z, tt, ss, den, den_pot as
    (nz, ny, nx), hhq); returns the new den, den_pot."""
    L = lib()
    w = {k: np.ascontiguousarray(inp[k], dtype=np.float64).copy() for k in inp}
    nz = w["tt"].shape[0]
    L.bto_density(C.byref(og.g), nz, _d(w["z"]), _d(w["tt"]), _d(w["ss"]), _d(w["hhq"]), _d(w["den"]),
                  _d(w["den_pot"]))
    return {"den": w["den"], "den_pot": w["den_pot"]}


def have_reference() -> bool:
    return os.path.exists(REF_HARNESS)


def run_reference(case_path: str, out_path: str, nx: int, ny: int, mode: int,
                  out_steps: Sequence[int] = (), threads: int = 4) -> dict:
    """Run the compiled reference harness on a case file and parse its output."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([REF_HARNESS, case_path, out_path], check=True, env=env,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, cwd=os.path.dirname(out_path))
    raw = np.fromfile(out_path, dtype="<f8")
    per = nx * ny
    if mode == 8:
        # masks, then dx, dy, dxt, dyt, dxh, dyh, dxb, dyb, rlh_s of basinpar
        assert raw.size == 16 * per
        names = ("dx", "dy", "dxt", "dyt", "dxh", "dyh", "dxb", "dyb", "rlh_s")
        return {k: raw[(7 + i) * per:(8 + i) * per].reshape(ny, nx).copy() for i, k in enumerate(names)}
    if mode == 7:
        # masks, z (nz), tt, ss, den, den_pot (nz planes), hhq, den, den_pot out
        nz = (raw.size - 8 * per) // (1 + 6 * per)
        res = {"masks": {k: raw[i * per:(i + 1) * per].reshape(ny, nx) for i, k in enumerate(MASKS)}}
        o = 7 * per
        inp = {"z": raw[o:o + nz].copy()}
        o += nz
        for k in ("tt", "ss", "den", "den_pot"):
            inp[k] = raw[o:o + nz * per].reshape(nz, ny, nx).copy()
            o += nz * per
        inp["hhq"] = raw[o:o + per].reshape(ny, nx).copy()
        o += per
        out = {}
        for k in ("den", "den_pot"):
            out[k] = raw[o:o + nz * per].reshape(nz, ny, nx).copy()
            o += nz * per
        assert o == raw.size
        res["eos_in"], res["eos_out"] = inp, out
        return res
    if mode == 6:
        # masks, z (nz), den, rhsx, rhsy (nz planes), hhq, hhu, hhv, rhsx, rhsy out
        nz = (raw.size - 7 * per - 3 * per) // (1 + 5 * per)
        res = {"masks": {k: raw[i * per:(i + 1) * per].reshape(ny, nx) for i, k in enumerate(MASKS)}}
        o = 7 * per
        inp = {"z": raw[o:o + nz].copy()}
        o += nz
        for k in ("den", "rhsx", "rhsy"):
            inp[k] = raw[o:o + nz * per].reshape(nz, ny, nx).copy()
            o += nz * per
        for k in ("hhq", "hhu", "hhv"):
            inp[k] = raw[o:o + per].reshape(ny, nx).copy()
            o += per
        out = {}
        for k in ("rhsx", "rhsy"):
            out[k] = raw[o:o + nz * per].reshape(nz, ny, nx).copy()
            o += nz * per
        assert o == raw.size
        res["pg_in"], res["pg_out"] = inp, out
        return res
    if mode == 10:
        # masks, dz (nz), uup, vvp (the velocities stress_components read),
        # stress_t, stress_s, amts, amuv (nz planes), amuv2d, then amts, amuv
        # (nz planes), amuv2d out
        nz = (raw.size - 7 * per - 2 * per) // (1 + 8 * per)
        res = {"masks": {k: raw[i * per:(i + 1) * per].reshape(ny, nx) for i, k in enumerate(MASKS)}}
        o = 7 * per
        inp = {"dz": raw[o:o + nz].copy()}
        o += nz
        for k in ("uup", "vvp", "stress_t", "stress_s", "amts", "amuv"):
            inp[k] = raw[o:o + nz * per].reshape(nz, ny, nx).copy()
            o += nz * per
        inp["amuv2d"] = raw[o:o + per].reshape(ny, nx).copy()
        o += per
        out = {}
        for k in ("amts", "amuv"):
            out[k] = raw[o:o + nz * per].reshape(nz, ny, nx).copy()
            o += nz * per
        out["amuv2d"] = raw[o:o + per].reshape(ny, nx).copy()
        o += per
        assert o == raw.size
        res["smag_in"], res["smag_out"] = inp, out
        return res
    if mode == 5:
        # masks, dz (nz), uu, vv (nz planes), ww (nz+1), hu, hv, wflux, ww out (nz+1)
        nz = (raw.size - 7 * per - 3 * per - 2 * per) // (1 + 4 * per)
        res = {"masks": {k: raw[i * per:(i + 1) * per].reshape(ny, nx) for i, k in enumerate(MASKS)}}
        o = 7 * per
        inp = {"dz": raw[o:o + nz].copy()}
        o += nz
        for k, n_ in (("uu", nz), ("vv", nz), ("ww", nz + 1)):
            inp[k] = raw[o:o + n_ * per].reshape(n_, ny, nx).copy()
            o += n_ * per
        for k in ("hu", "hv", "wflux"):
            inp[k] = raw[o:o + per].reshape(ny, nx).copy()
            o += per
        res["vv_in"] = inp
        res["ww_out"] = raw[o:o + (nz + 1) * per].reshape(nz + 1, ny, nx).copy()
        assert o + (nz + 1) * per == raw.size
        return res
    if mode == 4:
        # masks, dz (nz), hzt (nz+1), 8 3-D inputs, 2 of nz+1 levels, 14 planes, 4 3-D outputs
        nin2 = len(BAROCLINIC_IN_2D)
        # raw = 7 per + nz + (nz+1) + 8 nz per + 2 (nz+1) per + nin2 per + 4 nz per
        nz = (raw.size - (7 + 2 + nin2) * per - 1) // (2 + 14 * per)
        res = {"masks": {k: raw[i * per:(i + 1) * per].reshape(ny, nx) for i, k in enumerate(MASKS)}}
        o = 7 * per
        inp = {"dz": raw[o:o + nz].copy()}
        o += nz
        inp["hzt"] = raw[o:o + nz + 1].copy()
        o += nz + 1
        for k in BAROCLINIC_IN_3D:
            inp[k] = raw[o:o + nz * per].reshape(nz, ny, nx).copy()
            o += nz * per
        for k in BAROCLINIC_IN_W:
            inp[k] = raw[o:o + (nz + 1) * per].reshape(nz + 1, ny, nx).copy()
            o += (nz + 1) * per
        for k in BAROCLINIC_IN_2D:
            inp[k] = raw[o:o + per].reshape(ny, nx).copy()
            o += per
        out = {}
        for k in BAROCLINIC_OUT:
            out[k] = raw[o:o + nz * per].reshape(nz, ny, nx).copy()
            o += nz * per
        assert o == raw.size
        res["baroclinic_in"], res["baroclinic_out"] = inp, out
        return res
    if mode == 3:
        # masks, dz (nz values), 7 3-D fields, 12 planes of inputs, 2 of outputs
        nz = (raw.size - (7 + len(FORCING_IN_2D) + 2) * per) // (1 + len(FORCING_IN_3D) * per)
        res = {"masks": {k: raw[i * per:(i + 1) * per].reshape(ny, nx) for i, k in enumerate(MASKS)}}
        o = 7 * per
        inp = {"dz": raw[o:o + nz].copy()}
        o += nz
        for k in FORCING_IN_3D:
            inp[k] = raw[o:o + nz * per].reshape(nz, ny, nx)
            o += nz * per
        for k in FORCING_IN_2D:
            inp[k] = raw[o:o + per].reshape(ny, nx)
            o += per
        res["forcing_in"] = inp
        res["forcing_out"] = {"rhsx2d": raw[o:o + per].reshape(ny, nx), "rhsy2d": raw[o + per:o + 2 * per].reshape(ny, nx)}
        return res
    arrs = [raw[i * per:(i + 1) * per].reshape(ny, nx) for i in range(raw.size // per)]
    res = {"masks": dict(zip(MASKS, arrs[:7]))}
    arrs = arrs[7:]
    if mode == 2:
        nz = (len(arrs) - len(COUPLING_IN) - len(COUPLING_OUT)) // 4
        k = len(COUPLING_IN)
        inp = dict(zip(COUPLING_IN, arrs[:k]))
        inp["uu"] = np.stack(arrs[k:k + nz]); inp["vv"] = np.stack(arrs[k + nz:k + 2 * nz])
        k += 2 * nz
        out = {"uu": np.stack(arrs[k:k + nz]), "vv": np.stack(arrs[k + nz:k + 2 * nz])}
        out.update(zip(COUPLING_OUT, arrs[k + 2 * nz:]))
        res["coupling_in"], res["coupling_out"] = inp, out
    elif mode == 0:
        nf = len(OUT_FIELDS) + len(DEPTHS)
        for k, step in enumerate(out_steps):
            blk = arrs[k * nf:(k + 1) * nf]
            res[step] = dict(zip(OUT_FIELDS + DEPTHS, blk))
    else:
        res["units"] = dict(zip(UNIT_LAYOUT, arrs))
    return res
