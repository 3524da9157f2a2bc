"""Host-side mirror of the reference's barotropic interface.

``BarotropicModel`` owns one device context (one basin on one GPU).  Its
methods follow the reference's routines:

* ``barotropic_dynamics(...)``  - Function/vel_ssh.f90:379-679, same
  argument order and meaning, numpy arrays updated in place;
* ``model_steps(n, nstep, ksw4)`` - the 2-D part of
  Control/dyncall.f90:ocean_model_step (wind stress, bottom drag, RHS
  assembly, barotropic_dynamics), n times, in the reference's operation
  order;
* ``run(n, mode)`` - the same step with nstep = 1 through one fused kernel
  per step (modes 'fp64_exact', 'fp64', 'fp32').

All numpy fields are float64 (ny, nx) C-order arrays, i.e. the reference's
column-major (m, n) layout (see inmsom_bt.cases).
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Sequence

import numpy as np

from . import _capi
from ._capi import check, lib
from .cases import Basin

MASKS = ("lu", "luu", "luh", "lcu", "lcv", "llu", "llv")
METRICS = ("dx", "dy", "dxt", "dyt", "dxh", "dyh", "dxb", "dyb")
STATE = ("u", "up", "v", "vp", "ssh", "sshp")
DEPTHS = ("hhq_e", "hhqp_e", "hhqn_e", "hhu_e", "hhup_e", "hhun_e",
          "hhv_e", "hhvp_e", "hhvn_e", "hhh_e", "hhhp_e", "hhhn_e")


def _d(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_capi._dp)


def _f(a: np.ndarray):
    return a.ctypes.data_as(_capi._fp)


def gridcon(lbasins: np.ndarray, periodicity=(0, 0)) -> Dict[str, np.ndarray]:
    """Service/grid_construction.f90:33-133 (host code of the library)."""
    ny, nx = lbasins.shape
    lb = np.ascontiguousarray(lbasins, dtype=np.int32)
    out = {k: np.zeros((ny, nx), dtype=np.float32) for k in MASKS}
    check(lib().inmsom_bt_gridcon(nx, ny, periodicity[0], periodicity[1],
                                  lb.ctypes.data_as(_capi._ip), *[_f(out[k]) for k in MASKS]),
          "inmsom_bt_gridcon")
    return out


def _grid_struct(b: Basin, keep: list, periodicity=(0, 0), full_free_surface: int = 1,
                 masks: Optional[Dict[str, np.ndarray]] = None) -> _capi.Grid:
    g = _capi.Grid()
    g.nx, g.ny = b.nx, b.ny
    g.periodicity_x, g.periodicity_y = periodicity
    g.full_free_surface = full_free_surface
    g.free_fall_acc, g.ref_den, g.time_smooth = b.g, b.rho0, b.time_smooth
    lb = np.ascontiguousarray(b.lbasins, dtype=np.int32)
    keep.append(lb)
    g.lbasins = lb.ctypes.data_as(_capi._ip)
    if masks is not None:
        for k in MASKS:
            a = np.ascontiguousarray(masks[k], dtype=np.float32)
            keep.append(a)
            setattr(g, k, _f(a))
    for k in METRICS + ("rlh_s", "hhq_rest"):
        a = np.ascontiguousarray(getattr(b, k), dtype=np.float64)
        keep.append(a)
        setattr(g, k, _d(a))
    return g


def rayleigh(b: Basin, r_fric: Optional[float] = None, periodicity=(0, 0)) -> np.ndarray:
    """Control/init_pars.f90:459-481 (host code of the library)."""
    keep: list = []
    g = _grid_struct(b, keep, periodicity)
    out = np.zeros((b.ny, b.nx))
    check(lib().inmsom_bt_rayleigh(C.byref(g), b.r_fric if r_fric is None else r_fric, _d(out)),
          "inmsom_bt_rayleigh")
    return out


class BarotropicModel:
    """One basin resident on one GPU."""

    def __init__(self, b: Basin, device: int = 0, rdis: Optional[np.ndarray] = None,
                 periodicity=(0, 0), full_free_surface: int = 1,
                 masks: Optional[Dict[str, np.ndarray]] = None):
        self.b = b
        self.device = device
        self.nx, self.ny = b.nx, b.ny
        self._keep: list = []
        g = _grid_struct(b, self._keep, periodicity, full_free_surface, masks)
        h = C.c_void_p()
        check(lib().inmsom_bt_init(C.byref(g), device, C.byref(h)), "inmsom_bt_init")
        self._h = h
        self.rdis = rayleigh(b, periodicity=periodicity) if rdis is None else np.ascontiguousarray(rdis)
        self.set_forcing()
        self.set_state(b.u, b.up, b.v, b.vp, b.ssh, b.sshp)

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().inmsom_bt_finalize(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ state
    def set_forcing(self, taux=None, tauy=None, wflux=None, mu=None, mu4=None, rdis=None,
                    cb_nl=None, ebottom=None):
        b = self.b
        arrs = {k: np.ascontiguousarray(v if v is not None else getattr(b, k), dtype=np.float64)
                for k, v in (("taux", taux), ("tauy", tauy), ("wflux", wflux), ("mu", mu),
                             ("mu4", mu4))}
        arrs["rdis"] = np.ascontiguousarray(self.rdis if rdis is None else rdis, dtype=np.float64)
        f = _capi.Forcing()
        for k, a in arrs.items():
            setattr(f, k, _d(a))
        f.cb_nl = b.cb_nl if cb_nl is None else cb_nl
        f.ebottom = b.ebottom if ebottom is None else ebottom
        check(lib().inmsom_bt_set_forcing(self._h, C.byref(f)), "inmsom_bt_set_forcing")
        self._forcing = arrs

    def set_state(self, u, up, v, vp, ssh, sshp):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (u, up, v, vp, ssh, sshp)]
        check(lib().inmsom_bt_set_state(self._h, *[_d(x) for x in a]), "inmsom_bt_set_state")

    def save_state(self):
        """Keeps a copy of the current state on the device."""
        check(lib().inmsom_bt_save_state(self._h), "inmsom_bt_save_state")

    def restore_state(self):
        """The saved state back as the current one (set_state without host
        copies)."""
        check(lib().inmsom_bt_restore_state(self._h), "inmsom_bt_restore_state")

    def get_state(self) -> Dict[str, np.ndarray]:
        out = {k: np.empty((self.ny, self.nx)) for k in STATE}
        check(lib().inmsom_bt_get_state(self._h, *[_d(out[k]) for k in STATE]), "inmsom_bt_get_state")
        return out

    def means(self) -> Dict[str, np.ndarray]:
        keys = ("ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")
        out = {k: np.empty((self.ny, self.nx)) for k in keys}
        check(lib().inmsom_bt_get_means(self._h, *[_d(out[k]) for k in keys]), "inmsom_bt_get_means")
        return out

    def depths(self) -> Dict[str, np.ndarray]:
        out = {k: np.empty((self.ny, self.nx)) for k in DEPTHS}
        arr = (_capi._dp * 12)(*[_d(out[k]) for k in DEPTHS])
        check(lib().inmsom_bt_get_depths(self._h, arr), "inmsom_bt_get_depths")
        return out

    # ------------------------------------------------------------ steps
    def barotropic_dynamics(self, tau, nstep, ksw4, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e,
                            sshp_e, ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady, RHSx, RHSy, wflux,
                            mu, mu4, rdis):
        """Function/vel_ssh.f90:379-679; the first ten arrays are updated in place."""
        arrs = [ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, sshp_e, ubrtr_i, vbrtr_i, ssh4gradx,
                ssh4grady, RHSx, RHSy, wflux, mu, mu4, rdis]
        check(lib().inmsom_bt_barotropic_dynamics(self._h, float(tau), int(nstep), int(ksw4),
                                                  *[_d(a) for a in arrs]),
              "inmsom_bt_barotropic_dynamics")

    def barotropic_dynamics_dev(self, tau, nstep, ksw4, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e,
                                sshp_e, ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady, RHSx, RHSy, wflux,
                                mu, mu4, rdis):
        """As barotropic_dynamics on DeviceArray arguments (inmsom_bt.hipmem)
        of this model's shape and GPU: the first ten are updated in place,
        no host copies."""
        arrs = [ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, sshp_e, ubrtr_i, vbrtr_i, ssh4gradx,
                ssh4grady, RHSx, RHSy, wflux, mu, mu4, rdis]
        s2 = (self.ny, self.nx)
        check(lib().inmsom_bt_barotropic_dynamics_dev(self._h, float(tau), int(nstep), int(ksw4),
                                                      *[self._dev(a, s2) for a in arrs]),
              "inmsom_bt_barotropic_dynamics_dev")

    def barotropic_dynamics_noalloc(self, tau, nstep, ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, sshp_e,
                                    ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady, RHSx, RHSy, wflux, mu, rdis):
        """As barotropic_dynamics_dev with ksw4 = 0 through the entry that
        allocates no device memory (its scratch is a fixed array of the
        library); queued on the model's stream: call synchronize() before
        reading the results."""
        arrs = [ubrtr_e, ubrtrp_e, vbrtr_e, vbrtrp_e, ssh_e, sshp_e, ubrtr_i, vbrtr_i, ssh4gradx, ssh4grady,
                RHSx, RHSy, wflux, mu, rdis]
        s2 = (self.ny, self.nx)
        check(lib().inmsom_bt_barotropic_dynamics_noalloc(self._h, float(tau), int(nstep),
                                                          *[self._dev(a, s2) for a in arrs]),
              "inmsom_bt_barotropic_dynamics_noalloc")

    DROPIN_PATHS = {"auto": 0, "literal": 1, "fused": 2}

    def set_dropin_path(self, path: str = "auto"):
        """Path of barotropic_dynamics: 'auto' (fused where the call allows
        it), 'literal' (one kernel per reference sweep) or 'fused'."""
        check(lib().inmsom_bt_set_dropin_path(self._h, self.DROPIN_PATHS[path]), "inmsom_bt_set_dropin_path")

    def dropin_path(self, ksw4: int = 0) -> str:
        """The path ('literal' or 'fused') a barotropic_dynamics call with
        this ksw4 takes."""
        k = C.c_int()
        check(lib().inmsom_bt_get_dropin_path(self._h, int(ksw4), C.byref(k)), "inmsom_bt_get_dropin_path")
        return {1: "literal", 2: "fused"}[k.value]

    def dropin_timed(self, ncalls: int, nstep: int = 1, ksw4: int = 0, tau: Optional[float] = None) -> float:
        """ncalls device-resident barotropic_dynamics calls on the model's
        state (RHS of its wind stress and drag); HIP-event time in ms."""
        ms = C.c_double()
        check(lib().inmsom_bt_dropin_timed(self._h, float(self.b.dt if tau is None else tau), int(nstep),
                                           int(ksw4), int(ncalls), C.byref(ms)), "inmsom_bt_dropin_timed")
        return ms.value

    def model_steps(self, nsteps: int, nstep: int = 1, ksw4: int = 0, tau: Optional[float] = None):
        check(lib().inmsom_bt_model_steps(self._h, float(self.b.dt if tau is None else tau),
                                          int(nstep), int(ksw4), int(nsteps)),
              "inmsom_bt_model_steps")

    def run(self, nsteps: int, mode: str = "fp64", nstep: int = 1, tau: Optional[float] = None):
        """nsteps outer steps of tau, each sub-cycled in nstep fused launches."""
        check(lib().inmsom_bt_run(self._h, float(self.b.dt if tau is None else tau), int(nstep),
                                  int(nsteps), _capi.MODES[mode]), "inmsom_bt_run")

    def run_timed(self, nsteps: int, mode: str = "fp64", nstep: int = 1,
                  tau: Optional[float] = None) -> float:
        """As run(); returns the HIP-event time of the launches in ms."""
        ms = C.c_double()
        check(lib().inmsom_bt_run_timed(self._h, float(self.b.dt if tau is None else tau),
                                        int(nstep), int(nsteps), _capi.MODES[mode], C.byref(ms)),
              "inmsom_bt_run_timed")
        return ms.value

    def set_substeps_per_launch(self, k: int):
        """Fused step: 2 = substep pairs in one launch, 1 = one launch per
        substep, 0 = by mode (default); results do not depend on it."""
        check(lib().inmsom_bt_set_substeps_per_launch(self._h, int(k)), "inmsom_bt_set_substeps_per_launch")

    def set_pipeline(self, stages: int):
        """FP64 / FP32 fused step: stages = K in 2..4 runs K substeps per
        launch as a K-wave pipeline, 1 = off, 0 = automatic; results agree
        to rounding."""
        check(lib().inmsom_bt_set_pipeline(self._h, int(stages)), "inmsom_bt_set_pipeline")

    def pipeline(self, mode: str) -> int:
        """Pipeline stages the next run in this mode uses (0: none)."""
        k = C.c_int()
        check(lib().inmsom_bt_get_pipeline(self._h, _capi.MODES[mode], C.byref(k)), "inmsom_bt_get_pipeline")
        return k.value

    def substeps_per_launch(self, mode: str) -> int:
        k = C.c_int()
        check(lib().inmsom_bt_get_substeps_per_launch(self._h, _capi.MODES[mode], C.byref(k)),
              "inmsom_bt_get_substeps_per_launch")
        return k.value

    def set_launch(self, rows_per_wave: int):
        check(lib().inmsom_bt_set_launch(self._h, int(rows_per_wave)), "inmsom_bt_set_launch")

    def get_launch(self) -> int:
        r = C.c_int()
        check(lib().inmsom_bt_get_launch(self._h, C.byref(r)), "inmsom_bt_get_launch")
        return r.value

    def synchronize(self):
        check(lib().inmsom_bt_synchronize(self._h), "inmsom_bt_synchronize")

    # ------------------------------------------------------------ 2-D -> 3-D coupling
    def remove_barotropic(self, uu: np.ndarray, vv: np.ndarray, uu2d: np.ndarray, vv2d: np.ndarray):
        """Control/dyncall.f90:169-184 on (nz, ny, nx) arrays, in place."""
        for a in (uu, vv):
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.shape[1:] != (self.ny, self.nx):
                raise ValueError("uu, vv must be C-contiguous float64 (nz, ny, nx)")
        u2, v2 = (np.ascontiguousarray(a, dtype=np.float64) for a in (uu2d, vv2d))
        check(lib().inmsom_bt_remove_barotropic(self._h, uu.shape[0], _d(uu), _d(vv), _d(u2), _d(v2)),
              "inmsom_bt_remove_barotropic")

    def internal_update(self, tau: float, fields: dict, uu: np.ndarray, vv: np.ndarray) -> dict:
        """Control/dyncall.f90:188-231 with ssh_internal.  ``fields`` holds
        ubrtr_i, vbrtr_i, hhu, hhv, slpr, wflux (inputs) and uu2d, vv2d,
        pgrx, pgry, ssh_i, sshp_i, hhqn, hhun, hhvn, hhhn (updated; copies
        are returned); uu, vv (nz, ny, nx) are updated in place."""
        ins = {k: np.ascontiguousarray(fields[k], dtype=np.float64)
               for k in ("ubrtr_i", "vbrtr_i", "hhu", "hhv", "slpr", "wflux")}
        io = {k: np.array(fields[k], dtype=np.float64, order="C")
              for k in ("uu2d", "vv2d", "pgrx", "pgry", "ssh_i", "sshp_i", "hhqn", "hhun", "hhvn", "hhhn")}
        hh = (_capi._dp * 4)(*[_d(io[k]) for k in ("hhqn", "hhun", "hhvn", "hhhn")])
        for a in (uu, vv):
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.shape[1:] != (self.ny, self.nx):
                raise ValueError("uu, vv must be C-contiguous float64 (nz, ny, nx)")
        check(lib().inmsom_bt_internal_update(
            self._h, float(tau), uu.shape[0], *[_d(ins[k]) for k in ("ubrtr_i", "vbrtr_i", "hhu", "hhv", "slpr",
                                                                      "wflux")],
            *[_d(io[k]) for k in ("uu2d", "vv2d", "pgrx", "pgry", "ssh_i", "sshp_i")], hh, _d(uu), _d(vv)),
            "inmsom_bt_internal_update")
        return io

    def _dev(self, t, shape):
        """Device address of a DeviceArray (inmsom_bt.hipmem) of the given
        shape on this model's GPU."""
        from .hipmem import DeviceArray
        if not isinstance(t, DeviceArray) or t.ptr is None:
            raise ValueError("device arrays must be inmsom_bt.hipmem.DeviceArray objects")
        if t.shape != tuple(shape) or t.device != self.device:
            raise ValueError(f"device array must be on GPU {self.device} with shape {tuple(shape)}, "
                             f"not GPU {t.device}, {t.shape}")
        return C.cast(C.c_void_p(t.ptr), _capi._dp)

    def remove_barotropic_dev(self, uu, vv, uu2d, vv2d):
        """As remove_barotropic on DeviceArray arguments (no host copies);
        the arrays must be ready (synchronise the producing stream first)."""
        nz = uu.shape[0]
        s3, s2 = (nz, self.ny, self.nx), (self.ny, self.nx)
        check(lib().inmsom_bt_remove_barotropic_dev(self._h, nz, self._dev(uu, s3), self._dev(vv, s3),
                                                    self._dev(uu2d, s2), self._dev(vv2d, s2)),
              "inmsom_bt_remove_barotropic_dev")

    def internal_update_dev(self, tau: float, fields: dict, uu, vv):
        """As internal_update on device arrays, all updated in place."""
        nz = uu.shape[0]
        s3, s2 = (nz, self.ny, self.nx), (self.ny, self.nx)
        hh = (_capi._dp * 4)(*[self._dev(fields[k], s2) for k in ("hhqn", "hhun", "hhvn", "hhhn")])
        check(lib().inmsom_bt_internal_update_dev(
            self._h, float(tau), nz,
            *[self._dev(fields[k], s2) for k in ("ubrtr_i", "vbrtr_i", "hhu", "hhv", "slpr", "wflux", "uu2d",
                                                 "vv2d", "pgrx", "pgry", "ssh_i", "sshp_i")],
            hh, self._dev(uu, s3), self._dev(vv, s3)), "inmsom_bt_internal_update_dev")

    FORCING_3D = ("uu", "vv", "uup", "vvp", "amuv")
    FORCING_2D = ("uu2d", "vv2d", "uup2d", "vvp2d", "amuv2d", "hhq", "hhu", "hhv", "hhh", "taux", "tauy",
                  "slpr")

    def forcing_3d(self, fields: dict, dz, rhsx2d, rhsy2d, pgx=None, pgy=None) -> dict:
        """Control/dyncall.f90:37-130 (ksw_vert = ksw_lat = ksw_lat4 = 0):
        ``fields`` holds the 3-D (nz, ny, nx) uu, vv, uup, vvp, amuv and the
        2-D uu2d, vv2d, uup2d, vvp2d, amuv2d, hhq, hhu, hhv, hhh, taux, tauy,
        slpr; pgx, pgy are optional 3-D pressure-gradient terms.  Returns
        the updated copies of rhsx2d, rhsy2d."""
        a3 = {k: np.ascontiguousarray(fields[k], dtype=np.float64) for k in self.FORCING_3D}
        nz = a3["uu"].shape[0]
        for k, a in a3.items():
            if a.shape != (nz, self.ny, self.nx):
                raise ValueError(f"{k} must be (nz, ny, nx)")
        a2 = {k: np.ascontiguousarray(fields[k], dtype=np.float64) for k in self.FORCING_2D}
        dzc = np.ascontiguousarray(dz, dtype=np.float64)
        if dzc.shape != (nz,):
            raise ValueError("dz must have nz values")
        pg = [np.ascontiguousarray(p, dtype=np.float64) if p is not None else None for p in (pgx, pgy)]
        out = {"rhsx2d": np.array(rhsx2d, dtype=np.float64, order="C"),
               "rhsy2d": np.array(rhsy2d, dtype=np.float64, order="C")}
        check(lib().inmsom_bt_forcing_3d(
            self._h, nz, _d(dzc), *[_d(a3[k]) for k in self.FORCING_3D],
            *[_d(a2[k]) for k in ("uu2d", "vv2d", "uup2d", "vvp2d", "amuv2d", "hhq", "hhu", "hhv", "hhh", "taux",
                                  "tauy", "slpr")],
            *[(_d(p) if p is not None else None) for p in pg], _d(out["rhsx2d"]), _d(out["rhsy2d"])),
            "inmsom_bt_forcing_3d")
        return out

    def forcing_3d_dev(self, fields: dict, dz, rhsx2d, rhsy2d, pgx=None, pgy=None):
        """As forcing_3d on DeviceArray arguments (no host copies; dz a host
        sequence): rhsx2d, rhsy2d are updated in place."""
        dzc = np.ascontiguousarray(dz, dtype=np.float64)
        nz = dzc.shape[0]
        s3, s2 = (nz, self.ny, self.nx), (self.ny, self.nx)
        if (pgx is None) != (pgy is None):
            raise ValueError("pgx and pgy: both or neither")
        check(lib().inmsom_bt_forcing_3d_dev(
            self._h, nz, _d(dzc), *[self._dev(fields[k], s3) for k in self.FORCING_3D],
            *[self._dev(fields[k], s2) for k in self.FORCING_2D],
            *[(self._dev(q, s3) if q is not None else None) for q in (pgx, pgy)],
            self._dev(rhsx2d, s2), self._dev(rhsy2d, s2)), "inmsom_bt_forcing_3d_dev")

    def write_snapshot(self, path: str, nrec: int, xt0: float = 0.0, dxst: float = 1.0, yt0: float = 0.0,
                       dyst: float = 1.0, date=(2000, 1, 1, 0, 0), calendar: int = 1,
                       tstep: float = 86400.0) -> float:
        """ssh.dat, ubrtr.dat, vbrtr.dat (+ .ctl) in ``path`` with record
        ``nrec`` = the current state (reference output formats); returns the
        device time (ms) of packing and copying the fields."""
        ms = C.c_double()
        y, mo, d, h, mi = date
        check(lib().inmsom_bt_write_snapshot(self._h, path.encode(), int(nrec), float(xt0), float(dxst), float(yt0),
                                             float(dyst), int(calendar), y, mo, d, h, mi, float(tstep),
                                             C.byref(ms)), "inmsom_bt_write_snapshot")
        return ms.value

    BAROCLINIC_3D = ("rhsx", "rhsy", "rhsx_tran", "rhsy_tran")
    BAROCLINIC_W = ("ww", "nu")
    BAROCLINIC_2D = ("pgrx", "pgry", "rdis", "tx_surf", "ty_surf", "tx_bot", "ty_bot")
    BAROCLINIC_DEPTHS = ("hhu", "hhup", "hhun", "hhv", "hhvp", "hhvn", "hhh")

    def baroclinic_dynamics(self, tau: float, fields: dict) -> dict:
        """Function/vel_ssh.f90:794-1027 on host arrays: ``fields`` holds uu,
        uup, vv, vvp and rhsx, rhsy, rhsx_tran, rhsy_tran (nz, ny, nx), ww
        and nu (nz+1, ny, nx), the 2-D pgrx, pgry, rdis, tx_surf, ty_surf,
        tx_bot, ty_bot, the depths hhu .. hhh, and dz[nz], hzt[nz+1].
        Returns the updated uu, uup, vv, vvp."""
        out = {k: np.array(fields[k], dtype=np.float64, order="C") for k in ("uu", "uup", "vv", "vvp")}
        nz = out["uu"].shape[0]
        s3, sw, s2 = (nz, self.ny, self.nx), (nz + 1, self.ny, self.nx), (self.ny, self.nx)
        arr = {}
        for keys, shape in ((self.BAROCLINIC_3D, s3), (self.BAROCLINIC_W, sw),
                            (self.BAROCLINIC_2D + self.BAROCLINIC_DEPTHS, s2)):
            for k in keys:
                arr[k] = np.ascontiguousarray(fields[k], dtype=np.float64)
                if arr[k].shape != shape:
                    raise ValueError(f"{k} must have shape {shape}")
        for k in out:
            if out[k].shape != s3:
                raise ValueError(f"{k} must have shape {s3}")
        dz = np.ascontiguousarray(fields["dz"], dtype=np.float64)
        hzt = np.ascontiguousarray(fields["hzt"], dtype=np.float64)
        if dz.shape != (nz,) or hzt.shape != (nz + 1,):
            raise ValueError("dz must have nz values and hzt nz + 1")
        check(lib().inmsom_bt_baroclinic_dynamics(
            self._h, float(tau), nz, *[_d(out[k]) for k in ("uu", "uup", "vv", "vvp")],
            *[_d(arr[k]) for k in self.BAROCLINIC_W + self.BAROCLINIC_3D + self.BAROCLINIC_2D],
            _d(dz), _d(hzt), *[_d(arr[k]) for k in self.BAROCLINIC_DEPTHS]), "inmsom_bt_baroclinic_dynamics")
        return out

    def baroclinic_dynamics_dev(self, tau: float, fields: dict):
        """As baroclinic_dynamics on DeviceArray fields (dz, hzt host
        sequences); uu, uup, vv, vvp updated in place."""
        dz = np.ascontiguousarray(fields["dz"], dtype=np.float64)
        hzt = np.ascontiguousarray(fields["hzt"], dtype=np.float64)
        nz = dz.shape[0]
        s3, sw, s2 = (nz, self.ny, self.nx), (nz + 1, self.ny, self.nx), (self.ny, self.nx)
        check(lib().inmsom_bt_baroclinic_dynamics_dev(
            self._h, float(tau), nz, *[self._dev(fields[k], s3) for k in ("uu", "uup", "vv", "vvp")],
            *[self._dev(fields[k], sw) for k in self.BAROCLINIC_W],
            *[self._dev(fields[k], s3) for k in self.BAROCLINIC_3D],
            *[self._dev(fields[k], s2) for k in self.BAROCLINIC_2D],
            _d(dz), _d(hzt), *[self._dev(fields[k], s2) for k in self.BAROCLINIC_DEPTHS]),
            "inmsom_bt_baroclinic_dynamics_dev")

    def baroclinic_timed(self, tau: float, nz: int, nrep: int) -> float:
        """HIP-event time (ms) of nrep repetitions of baroclinic_dynamics's
        device work on the fields the last host-array call left staged."""
        ms = C.c_double()
        check(lib().inmsom_bt_baroclinic_timed(self._h, float(tau), int(nz), int(nrep), C.byref(ms)),
              "inmsom_bt_baroclinic_timed")
        return ms.value

    def vertical_velocity(self, dz, uu, vv, ww, hu, hv, wflux) -> np.ndarray:
        """Function/vel_ssh.f90:1030-1076 on host arrays: uu, vv (nz, ny, nx),
        ww (nz+1, ny, nx) (cells not computed keep their values), hu, hv,
        wflux (ny, nx), dz[nz].  Returns the new ww."""
        u = np.ascontiguousarray(uu, dtype=np.float64)
        v = np.ascontiguousarray(vv, dtype=np.float64)
        nz = u.shape[0]
        w = np.array(ww, dtype=np.float64, order="C")
        if u.shape != (nz, self.ny, self.nx) or v.shape != u.shape or w.shape != (nz + 1, self.ny, self.nx):
            raise ValueError("uu, vv must be (nz, ny, nx) and ww (nz+1, ny, nx)")
        two = [np.ascontiguousarray(a, dtype=np.float64) for a in (hu, hv, wflux)]
        dzc = np.ascontiguousarray(dz, dtype=np.float64)
        if dzc.shape != (nz,):
            raise ValueError("dz must have nz values")
        check(lib().inmsom_bt_vertical_velocity(self._h, nz, _d(dzc), _d(u), _d(v), _d(w), *[_d(a) for a in two]),
              "inmsom_bt_vertical_velocity")
        return w

    def vertical_velocity_dev(self, dz, uu, vv, ww, hu, hv, wflux):
        """As vertical_velocity on DeviceArray arguments; ww updated in place."""
        dzc = np.ascontiguousarray(dz, dtype=np.float64)
        nz = dzc.shape[0]
        s3, sw, s2 = (nz, self.ny, self.nx), (nz + 1, self.ny, self.nx), (self.ny, self.nx)
        check(lib().inmsom_bt_vertical_velocity_dev(self._h, nz, _d(dzc), self._dev(uu, s3), self._dev(vv, s3),
                                                    self._dev(ww, sw), self._dev(hu, s2), self._dev(hv, s2),
                                                    self._dev(wflux, s2)), "inmsom_bt_vertical_velocity_dev")

    def vertical_velocity_timed(self, nz: int, nrep: int) -> float:
        """HIP-event time (ms) of nrep repetitions of vertical_velocity's
        device work on the fields the last host-array call staged."""
        ms = C.c_double()
        check(lib().inmsom_bt_vertical_velocity_timed(self._h, int(nz), int(nrep), C.byref(ms)),
              "inmsom_bt_vertical_velocity_timed")
        return ms.value

    def pressure_gradients(self, z, den, rhsx, rhsy, hhq, hhu, hhv):
        """Function/vel_ssh.f90:320-376 on host arrays: den, rhsx, rhsy
        (nz, ny, nx), z[nz], hhq, hhu, hhv (ny, nx).  Returns the updated
        (rhsx, rhsy) as new arrays (cells off lcu, lcv keep their values)."""
        d = np.ascontiguousarray(den, dtype=np.float64)
        nz = d.shape[0]
        rx = np.array(rhsx, dtype=np.float64, order="C")
        ry = np.array(rhsy, dtype=np.float64, order="C")
        if d.shape != (nz, self.ny, self.nx) or rx.shape != d.shape or ry.shape != d.shape:
            raise ValueError("den, rhsx, rhsy must be (nz, ny, nx)")
        two = [np.ascontiguousarray(a, dtype=np.float64) for a in (hhq, hhu, hhv)]
        if any(a.shape != (self.ny, self.nx) for a in two):
            raise ValueError("hhq, hhu, hhv must be (ny, nx)")
        zc = np.ascontiguousarray(z, dtype=np.float64)
        if zc.shape != (nz,):
            raise ValueError("z must have nz values")
        check(lib().inmsom_bt_pressure_gradients(self._h, nz, _d(zc), _d(d), *[_d(a) for a in two], _d(rx), _d(ry)),
              "inmsom_bt_pressure_gradients")
        return rx, ry

    def pressure_gradients_dev(self, z, den, rhsx, rhsy, hhq, hhu, hhv):
        """As pressure_gradients on DeviceArray arguments; rhsx, rhsy updated
        in place."""
        zc = np.ascontiguousarray(z, dtype=np.float64)
        nz = zc.shape[0]
        s3, s2 = (nz, self.ny, self.nx), (self.ny, self.nx)
        check(lib().inmsom_bt_pressure_gradients_dev(self._h, nz, _d(zc), self._dev(den, s3), self._dev(hhq, s2),
                                                     self._dev(hhu, s2), self._dev(hhv, s2), self._dev(rhsx, s3),
                                                     self._dev(rhsy, s3)), "inmsom_bt_pressure_gradients_dev")

    def pressure_gradients_timed(self, nz: int, nrep: int) -> float:
        """HIP-event time (ms) of nrep repetitions of pressure_gradients's
        device work on the fields the last host-array call staged."""
        ms = C.c_double()
        check(lib().inmsom_bt_pressure_gradients_timed(self._h, int(nz), int(nrep), C.byref(ms)),
              "inmsom_bt_pressure_gradients_timed")
        return ms.value

    def density(self, z, tt, ss, hhq, den, den_pot):
        """Control/dyncall.f90:437-449 with Service/oc_algorithms.f90:138-161
        on host arrays: tt, ss, den, den_pot (nz, ny, nx), hhq (ny, nx), z[nz].
        Returns the new (den, den_pot) as new arrays (cells the reference
        does not write keep their values)."""
        t = np.ascontiguousarray(tt, dtype=np.float64)
        nz = t.shape[0]
        s = np.ascontiguousarray(ss, dtype=np.float64)
        d = np.array(den, dtype=np.float64, order="C")
        dp = np.array(den_pot, dtype=np.float64, order="C")
        if t.shape != (nz, self.ny, self.nx) or any(a.shape != t.shape for a in (s, d, dp)):
            raise ValueError("tt, ss, den, den_pot must be (nz, ny, nx)")
        h = np.ascontiguousarray(hhq, dtype=np.float64)
        if h.shape != (self.ny, self.nx):
            raise ValueError("hhq must be (ny, nx)")
        zc = np.ascontiguousarray(z, dtype=np.float64)
        if zc.shape != (nz,):
            raise ValueError("z must have nz values")
        check(lib().inmsom_bt_density(self._h, nz, _d(zc), _d(t), _d(s), _d(h), _d(d), _d(dp)),
              "inmsom_bt_density")
        return d, dp

    def density_dev(self, z, tt, ss, hhq, den, den_pot):
        """As density on DeviceArray arguments; den, den_pot updated in place."""
        zc = np.ascontiguousarray(z, dtype=np.float64)
        nz = zc.shape[0]
        s3, s2 = (nz, self.ny, self.nx), (self.ny, self.nx)
        check(lib().inmsom_bt_density_dev(self._h, nz, _d(zc), self._dev(tt, s3), self._dev(ss, s3),
                                          self._dev(hhq, s2), self._dev(den, s3), self._dev(den_pot, s3)),
              "inmsom_bt_density_dev")

    def density_timed(self, nz: int, nrep: int) -> float:
        """HIP-event time (ms) of nrep repetitions of the density
        definition's device work on the fields the last host-array call
        staged."""
        ms = C.c_double()
        check(lib().inmsom_bt_density_timed(self._h, int(nz), int(nrep), C.byref(ms)), "inmsom_bt_density_timed")
        return ms.value

    def stress_components(self, u, v, str_t, str_s):
        """Function/mixing.f90:115-168 on host arrays of 3-D velocities
        u, v and stress components str_t, str_s (nz, ny, nx) (cells not
        computed keep their values).  Returns the new (str_t, str_s)."""
        uc = np.ascontiguousarray(u, dtype=np.float64)
        nz = uc.shape[0]
        vc = np.ascontiguousarray(v, dtype=np.float64)
        st = np.array(str_t, dtype=np.float64, order="C")
        ss = np.array(str_s, dtype=np.float64, order="C")
        if uc.shape != (nz, self.ny, self.nx) or any(a.shape != uc.shape for a in (vc, st, ss)):
            raise ValueError("u, v, str_t, str_s must be (nz, ny, nx)")
        check(lib().inmsom_bt_stress_components(self._h, nz, _d(uc), _d(vc), _d(st), _d(ss)),
              "inmsom_bt_stress_components")
        return st, ss

    def stress_components_dev(self, nz: int, u, v, str_t, str_s):
        """As stress_components on DeviceArray arguments; str_t, str_s
        updated in place."""
        s3 = (int(nz), self.ny, self.nx)
        check(lib().inmsom_bt_stress_components_dev(self._h, int(nz), self._dev(u, s3), self._dev(v, s3),
                                                    self._dev(str_t, s3), self._dev(str_s, s3)),
              "inmsom_bt_stress_components_dev")

    def stress_components_timed(self, nz: int, nrep: int) -> float:
        """HIP-event time (ms) of nrep repetitions of stress_components'
        device work on the fields the last host-array call staged."""
        ms = C.c_double()
        check(lib().inmsom_bt_stress_components_timed(self._h, int(nz), int(nrep), C.byref(ms)),
              "inmsom_bt_stress_components_timed")
        return ms.value

    def smagorinsky(self, dz, str_t, str_s, amts, amuv, amuv2d, visc2: float, dift: float):
        """Control/dyncall.f90:40-43 on host arrays: smagorinsky_coeff
        (Function/mixing.f90:171-219) then depth_ave(amuv, amuv2d, lu, 0)
        (Function/depth.f90:223-254).  str_t, str_s, amts, amuv (nz, ny, nx),
        amuv2d (ny, nx), dz[nz]; visc2, dift are lvisc_2 and ldiff_ts.
        Returns the new (amts, amuv, amuv2d) as new arrays (cells the
        reference does not write keep their values)."""
        t = np.ascontiguousarray(str_t, dtype=np.float64)
        nz = t.shape[0]
        s = np.ascontiguousarray(str_s, dtype=np.float64)
        ts = np.array(amts, dtype=np.float64, order="C")
        uv = np.array(amuv, dtype=np.float64, order="C")
        uv2 = np.array(amuv2d, dtype=np.float64, order="C")
        if t.shape != (nz, self.ny, self.nx) or any(a.shape != t.shape for a in (s, ts, uv)):
            raise ValueError("str_t, str_s, amts, amuv must be (nz, ny, nx)")
        if uv2.shape != (self.ny, self.nx):
            raise ValueError("amuv2d must be (ny, nx)")
        dzc = np.ascontiguousarray(dz, dtype=np.float64)
        if dzc.shape != (nz,):
            raise ValueError("dz must have nz values")
        check(lib().inmsom_bt_smagorinsky(self._h, nz, float(visc2), float(dift), _d(t), _d(s), _d(dzc), _d(ts),
                                          _d(uv), _d(uv2)), "inmsom_bt_smagorinsky")
        return ts, uv, uv2

    def smagorinsky_dev(self, dz, str_t, str_s, amts, amuv, amuv2d, visc2: float, dift: float):
        """As smagorinsky on DeviceArray arguments; amts, amuv, amuv2d
        updated in place."""
        dzc = np.ascontiguousarray(dz, dtype=np.float64)
        nz = dzc.shape[0]
        s3, s2 = (nz, self.ny, self.nx), (self.ny, self.nx)
        check(lib().inmsom_bt_smagorinsky_dev(self._h, nz, float(visc2), float(dift), self._dev(str_t, s3),
                                              self._dev(str_s, s3), _d(dzc), self._dev(amts, s3),
                                              self._dev(amuv, s3), self._dev(amuv2d, s2)),
              "inmsom_bt_smagorinsky_dev")

    def smagorinsky_timed(self, nz: int, nrep: int) -> float:
        """HIP-event time (ms) of nrep repetitions of the Smagorinsky
        viscosity's device work on the fields the last host-array call
        staged."""
        ms = C.c_double()
        check(lib().inmsom_bt_smagorinsky_timed(self._h, int(nz), int(nrep), C.byref(ms)),
              "inmsom_bt_smagorinsky_timed")
        return ms.value

    def forcing_3d_timed(self, nz: int, nrep: int) -> float:
        """HIP-event time (ms) of nrep repetitions of the forcing's device
        work on the fields the last forcing_3d call left on the device."""
        ms = C.c_double()
        check(lib().inmsom_bt_forcing_3d_timed(self._h, int(nz), int(nrep), C.byref(ms)),
              "inmsom_bt_forcing_3d_timed")
        return ms.value

    def coupling_timed(self, tau: float, nz: int, nrep: int) -> float:
        """HIP-event time (ms) of nrep repetitions of the coupling's device
        kernels on the fields the last internal_update left on the device."""
        ms = C.c_double()
        check(lib().inmsom_bt_coupling_timed(self._h, float(tau), int(nz), int(nrep), C.byref(ms)),
              "inmsom_bt_coupling_timed")
        return ms.value

    # ------------------------------------------------------------ diagnostics
    def coarsen(self, factor: int = 4) -> dict:
        """Block average of the current state onto a grid ``factor`` times
        coarser: {"u", "v", "ssh"} arrays of ((ny-4)/f+4, (nx-4)/f+4)."""
        nxc, nyc = (self.nx - 4) // factor + 4, (self.ny - 4) // factor + 4
        out = {k: np.zeros((nyc, nxc)) for k in ("u", "v", "ssh")}
        check(lib().inmsom_bt_coarsen_state(self._h, factor, _d(out["u"]), _d(out["v"]), _d(out["ssh"])),
              "inmsom_bt_coarsen_state")
        return out

    def volume(self) -> float:
        v = C.c_double()
        check(lib().inmsom_bt_volume(self._h, C.byref(v)), "inmsom_bt_volume")
        return v.value

    def range_check(self) -> int:
        """Number of wet cells failing Control/dyncall.f90:452-516's bounds."""
        bad = C.c_longlong()
        code = lib().inmsom_bt_range_check(self._h, C.byref(bad))
        if code not in (_capi.OK, _capi.ERANGE):
            check(code, "inmsom_bt_range_check")
        return bad.value
