"""Batched ensembles of uneven basins (host side of inmsom_bt_ens_*).

``Ensemble`` advances many closed basins together: one persistent launch
per substep over all members, replayed from a HIP graph per outer step.
``autotune`` picks the launch shape (rows marched per wave, waves per XCD
of the persistent grid) by timing candidates on the ensemble itself, and
keeps the choice in a JSON cache keyed by the ensemble's shape.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _capi
from ._capi import check, lib
from .cases import Basin
from .model import STATE, _d, _grid_struct, rayleigh

DEFAULT_ROWS = (0, 12, 16, 24, 32, 40, 48, 64)
DEFAULT_WAVES = (0, 256, 384, 512)
# two-wave workgroups per XCD when substep pairs run as a pipeline
DEFAULT_GROUPS = (0, 96, 128, 160, 192)


class Ensemble:
    def __init__(self, basins: Sequence[Basin], mode: str = "fp32", device: int = 0,
                 rdis: Optional[Sequence[np.ndarray]] = None):
        self.basins = list(basins)
        self.mode = mode
        self._keep: list = []
        M = len(self.basins)
        grids = (_capi.Grid * M)()
        forc = (_capi.Forcing * M)()
        self.rdis = []
        for k, b in enumerate(self.basins):
            grids[k] = _grid_struct(b, self._keep)
            rd = rayleigh(b) if rdis is None else np.ascontiguousarray(rdis[k], dtype=np.float64)
            self.rdis.append(rd)
            arrs = [np.ascontiguousarray(getattr(b, n), dtype=np.float64)
                    for n in ("taux", "tauy", "wflux", "mu", "mu4")] + [rd]
            self._keep.extend(arrs)
            f = _capi.Forcing()
            f.taux, f.tauy, f.wflux, f.mu, f.mu4, f.rdis = [_d(a) for a in arrs]
            f.cb_nl, f.ebottom = b.cb_nl, b.ebottom
            forc[k] = f
        h = C.c_void_p()
        check(lib().inmsom_bt_ens_create(M, grids, forc, _capi.MODES[mode], device, C.byref(h)),
              "inmsom_bt_ens_create")
        self._h = h
        for k, b in enumerate(self.basins):
            self.set_state(k, b.u, b.up, b.v, b.vp, b.ssh, b.sshp)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().inmsom_bt_ens_destroy(self._h)
            self._h = None
        self.closure = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def size(self) -> Tuple[int, int, int]:
        n, w, c = C.c_int(), C.c_longlong(), C.c_longlong()
        check(lib().inmsom_bt_ens_size(self._h, C.byref(n), C.byref(w), C.byref(c)), "inmsom_bt_ens_size")
        return n.value, w.value, c.value

    def set_state(self, k, u, up, v, vp, ssh, sshp):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (u, up, v, vp, ssh, sshp)]
        check(lib().inmsom_bt_ens_set_state(self._h, k, *[_d(x) for x in a]), "inmsom_bt_ens_set_state")

    def reset(self):
        for k, b in enumerate(self.basins):
            self.set_state(k, b.u, b.up, b.v, b.vp, b.ssh, b.sshp)

    def get_state(self, k):
        b = self.basins[k]
        out = {n: np.empty((b.ny, b.nx)) for n in STATE}
        check(lib().inmsom_bt_ens_get_state(self._h, k, *[_d(out[n]) for n in STATE]),
              "inmsom_bt_ens_get_state")
        return out

    def set_launch(self, rows_per_wave: int = 0, waves_per_xcd: int = 0):
        check(lib().inmsom_bt_ens_set_launch(self._h, rows_per_wave, waves_per_xcd),
              "inmsom_bt_ens_set_launch")
        self.launch = (rows_per_wave, waves_per_xcd)

    def set_substeps_per_launch(self, k: int):
        """2 = substep pairs in one launch, 1 = one launch per substep,
        0 = by mode (default); results do not depend on it."""
        check(lib().inmsom_bt_ens_set_substeps_per_launch(self._h, int(k)),
              "inmsom_bt_ens_set_substeps_per_launch")

    def set_pipeline(self, stages: int):
        """2 = substep pairs as a two-wave pipeline, 1 = off, 0 = by mode
        (default); results agree to rounding."""
        check(lib().inmsom_bt_ens_set_pipeline(self._h, int(stages)), "inmsom_bt_ens_set_pipeline")

    def pipeline(self) -> int:
        k = C.c_int()
        check(lib().inmsom_bt_ens_get_pipeline(self._h, C.byref(k)), "inmsom_bt_ens_get_pipeline")
        return k.value

    def substeps_per_launch(self) -> int:
        k = C.c_int()
        check(lib().inmsom_bt_ens_get_substeps_per_launch(self._h, C.byref(k)),
              "inmsom_bt_ens_get_substeps_per_launch")
        return k.value

    # Graph replays by default; False makes run()/run_timed() without an
    # explicit graph argument launch directly (same results bit for bit),
    # for tools that do not follow graph launches.
    USE_GRAPHS = True

    def run(self, nsteps: int, nstep: int = 2, tau: Optional[float] = None, graph: Optional[bool] = None):
        tau = self.basins[0].dt if tau is None else tau
        graph = self.USE_GRAPHS if graph is None else graph
        check(lib().inmsom_bt_ens_run(self._h, float(tau), nstep, nsteps, int(graph)), "inmsom_bt_ens_run")

    def run_timed(self, nsteps: int, nstep: int = 2, tau: Optional[float] = None,
                  graph: Optional[bool] = None) -> float:
        tau = self.basins[0].dt if tau is None else tau
        graph = self.USE_GRAPHS if graph is None else graph
        ms = C.c_double()
        check(lib().inmsom_bt_ens_run_timed(self._h, float(tau), nstep, nsteps, int(graph), C.byref(ms)),
              "inmsom_bt_ens_run_timed")
        return ms.value

    def save(self):
        check(lib().inmsom_bt_ens_save(self._h), "inmsom_bt_ens_save")

    def restore(self):
        check(lib().inmsom_bt_ens_restore(self._h), "inmsom_bt_ens_restore")

    def range_check(self) -> int:
        bad = C.c_longlong()
        code = lib().inmsom_bt_ens_range_check(self._h, C.byref(bad))
        if code not in (_capi.OK, _capi.ERANGE):
            check(code, "inmsom_bt_ens_range_check")
        return bad.value

    def volume(self, k: int) -> float:
        v = C.c_double()
        check(lib().inmsom_bt_ens_volume(self._h, k, C.byref(v)), "inmsom_bt_ens_volume")
        return v.value

    # ------------------------------------------------------------ closure
    def set_closure(self, net, ratio: float = 3.0, interval: int = 1):
        """Attach a learned closure (a ``ClosureNet``; None detaches): after
        every ``interval``-th outer step the network's tendency times
        ``interval`` steps is added to each member's u and v, except for
        members whose closure-tendency norm exceeds ``ratio`` times their
        plain viscous tendency (plain step instead)."""
        h = None if net is None else net._h
        check(lib().inmsom_bt_ens_set_closure(self._h, h, float(ratio)), "inmsom_bt_ens_set_closure")
        check(lib().inmsom_bt_ens_set_closure_interval(self._h, int(interval)),
              "inmsom_bt_ens_set_closure_interval")
        self.closure = net  # the network must outlive the attachment

    def closure_stats(self):
        """(fallback step counts [M], last-step norms [M, 2] = closure, viscous)."""
        M = len(self.basins)
        nf = np.zeros(M, dtype=np.int32)
        nr = np.zeros((M, 2), dtype=np.float64)
        check(lib().inmsom_bt_ens_closure_stats(self._h, nf.ctypes.data_as(_capi._ip),
                                                 nr.ctypes.data_as(_capi._dp)),
              "inmsom_bt_ens_closure_stats")
        return nf, nr

    # ------------------------------------------------------------ autotuner
    def signature(self) -> str:
        shapes = ",".join(f"{b.nx}x{b.ny}" for b in self.basins)
        return hashlib.sha1(f"{self.mode}|pipeline {self.pipeline()}|{shapes}".encode()).hexdigest()[:16]

    def autotune(self, nstep: int = 2, trial_steps: int = 20,
                 rows: Sequence[int] = DEFAULT_ROWS, waves: Optional[Sequence[int]] = None,
                 cache: Optional[str] = None) -> dict:
        """Times every (rows per wave, waves per XCD) candidate for
        trial_steps outer steps from the same saved state (restored after
        each trial) and keeps the fastest.  Results never depend on the
        shape; only the time does.  waves defaults to DEFAULT_WAVES, or to
        DEFAULT_GROUPS (two-wave workgroups) under the pipeline."""
        if waves is None:
            waves = DEFAULT_GROUPS if self.pipeline() == 2 else DEFAULT_WAVES
        key = self.signature()
        if cache and os.path.exists(cache):
            with open(cache) as f:
                entries = json.load(f)
            if key in entries:
                best = entries[key]
                self.set_launch(best["rows_per_wave"], best["waves_per_xcd"])
                return dict(best, cached=True)
        self.save()
        trials: List[dict] = []
        for r in rows:
            for w in waves:
                self.set_launch(r, w)
                self.run(1, nstep)  # builds tables and the graph outside the timing
                self.restore()
                ms = self.run_timed(trial_steps, nstep)
                self.restore()
                trials.append({"rows_per_wave": r, "waves_per_xcd": w, "ms": ms})
        best = min(trials, key=lambda t: t["ms"])
        self.set_launch(best["rows_per_wave"], best["waves_per_xcd"])
        result = {"rows_per_wave": best["rows_per_wave"], "waves_per_xcd": best["waves_per_xcd"],
                  "ms_per_step": best["ms"] / trial_steps, "trials": trials, "cached": False}
        if cache:
            entries = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    entries = json.load(f)
            entries[key] = {k: result[k] for k in ("rows_per_wave", "waves_per_xcd", "ms_per_step")}
            with open(cache, "w") as f:
                json.dump(entries, f, indent=1)
        return result
