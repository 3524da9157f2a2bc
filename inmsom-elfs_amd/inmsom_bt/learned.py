"""Learned closure workflow: snapshots of a fine run, coarse targets,
training on the device, and rollouts judged against the coarsened fine run.

1. ``FineTruth`` advances the fine basin (FP64 fused step) and block-averages
   its state onto the coarse grid on the device at chosen fine steps.
2. ``samples`` turns coarsened snapshots into training samples.  The coarse
   model's state is (level t, level t - dt_c / nstep): its leap-frog
   substep is dt_c / nstep.  So a snapshot holds C(t - dt_c / nstep), C(t)
   and C(t + dt_c); the coarse model (an ensemble of copies of the coarse
   basin, FP64) takes one outer step from (C(t), C(t - dt_c / nstep)), and
   the target is the subgrid lateral-momentum tendency
       (C(t + dt_c) - coarse_step(C(t), C(t - dt_c / nstep))) / dt_c
   on U and V points, i.e. the coarsened fine tendency minus the coarse-grid
   tendency, and the inputs are the plain coarse step's output state, which
   is what the closure sees when it is applied after each outer step.
   (The coarse model's previous level is Robert-Asselin filtered, the
   coarsened one is not; the difference is time_smooth / (2 nstep) times a
   second time difference and is left to the closure.)
3. ``train`` runs Adam on the device-resident samples.
4. ``rollout`` runs the coarse model with or without the closure and scores
   SSH, u and v against the coarsened truth.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from .cases import Basin, coarsen_basin
from .closure import ClosureNet, features
from .ensemble import Ensemble
from .model import BarotropicModel


def uv_masks(b: Basin):
    """U and V point masks (both neighbouring T cells wet)."""
    wet = b.wet
    lcu = np.zeros_like(wet)
    lcv = np.zeros_like(wet)
    lcu[1:-2, 1:-2] = wet[1:-2, 1:-2] & wet[1:-2, 2:-1]
    lcv[1:-2, 1:-2] = wet[1:-2, 1:-2] & wet[2:-1, 1:-2]
    return lcu, lcv


def mask_coarse(state: Dict[str, np.ndarray], cb: Basin) -> Dict[str, np.ndarray]:
    """Coarsened fields restricted to the coarse basin's wet points."""
    lcu, lcv = uv_masks(cb)
    return {"u": state["u"] * lcu, "v": state["v"] * lcv, "ssh": state["ssh"] * cb.wet}


class FineTruth:
    """The fine run and its coarsened view."""

    def __init__(self, fine: Basin, factor: int = 4, nstep: int = 2, mode: str = "fp64",
                 coarse_nstep: int = 2):
        self.fine, self.factor, self.nstep, self.mode = fine, factor, nstep, mode
        self.coarse = coarsen_basin(fine, factor)
        self.coarse_nstep = coarse_nstep
        # fine outer steps per coarse outer step, and per coarse leap-frog substep
        self.per_step = factor
        sub = self.coarse.dt / coarse_nstep / fine.dt
        if abs(sub - round(sub)) > 1e-9 or round(sub) < 1:
            raise ValueError("the coarse substep must be a whole number of fine steps")
        self.per_sub = int(round(sub))
        self.model = BarotropicModel(fine)
        self.step = 0

    def advance_to(self, step: int):
        if step < self.step:
            raise ValueError("the fine run only moves forward")
        if step > self.step:
            self.model.run(step - self.step, self.mode, nstep=self.nstep)
            self.step = step

    def coarse_state(self) -> Dict[str, np.ndarray]:
        return mask_coarse(self.model.coarsen(self.factor), self.coarse)

    def close(self):
        self.model.close()


@dataclasses.dataclass
class Snapshot:
    step: int                       # fine step of C(t)
    prev: Dict[str, np.ndarray]     # C(t - dt_c / nstep)
    now: Dict[str, np.ndarray]      # C(t)
    nxt: Dict[str, np.ndarray]      # C(t + dt_c)


def snapshots(truth: FineTruth, steps: Sequence[int]) -> List[Snapshot]:
    """Coarsened states around each fine step in ``steps`` (ascending, at
    least per_sub apart from the previous snapshot's t + dt_c)."""
    out = []
    for s in steps:
        truth.advance_to(s - truth.per_sub)
        prev = truth.coarse_state()
        truth.advance_to(s)
        now = truth.coarse_state()
        truth.advance_to(s + truth.per_step)
        out.append(Snapshot(s, prev, now, truth.coarse_state()))
    return out


@dataclasses.dataclass
class Norms:
    shift: np.ndarray
    scale: np.ndarray
    out_scale: np.ndarray


class CoarseStepper:
    """An FP64 ensemble of ``width`` copies of the coarse basin that takes one
    plain outer step from each of a batch of snapshots."""

    def __init__(self, cb: Basin, width: int = 64, nstep: int = 2):
        self.cb, self.width, self.nstep = cb, width, nstep
        self.ens = Ensemble([cb] * width, mode="fp64")

    def __call__(self, snaps: Sequence[Snapshot]) -> List[dict]:
        out: List[dict] = []
        for i0 in range(0, len(snaps), self.width):
            chunk = snaps[i0:i0 + self.width]
            for k, s in enumerate(chunk):
                self.ens.set_state(k, s.now["u"], s.prev["u"], s.now["v"], s.prev["v"], s.now["ssh"],
                                   s.prev["ssh"])
            self.ens.run(1, nstep=self.nstep, tau=self.cb.dt)
            out.extend(self.ens.get_state(k) for k in range(len(chunk)))
        return out

    def close(self):
        self.ens.close()


def coarse_steps(cb: Basin, snaps: Sequence[Snapshot], width: int = 64, nstep: int = 2) -> List[dict]:
    """One plain FP64 outer step of the coarse basin from each snapshot's
    (C(t), C(t - dt_c / nstep)), ``width`` snapshots per batched launch."""
    st = CoarseStepper(cb, min(width, len(snaps)), nstep)
    try:
        return st(snaps)
    finally:
        st.close()


def samples(cb: Basin, snaps: Sequence[Snapshot], plain: Sequence[dict]):
    """(inputs-before-normalisation [n, 4, ny, nx], targets [n, ny, nx, 2], mask [ny, nx])."""
    lcu, lcv = uv_masks(cb)
    dt = cb.dt
    n = len(snaps)
    raw = np.zeros((n, 4, cb.ny, cb.nx), dtype=np.float64)
    tgt = np.zeros((n, cb.ny, cb.nx, 2), dtype=np.float32)
    for k, (s, p) in enumerate(zip(snaps, plain)):
        raw[k, 0], raw[k, 1], raw[k, 2], raw[k, 3] = p["u"], p["v"], p["ssh"], cb.hhq_rest
        tgt[k, ..., 0] = (s.nxt["u"] - p["u"]) / dt * lcu
        tgt[k, ..., 1] = (s.nxt["v"] - p["v"]) / dt * lcv
    mask = (lcu.astype(np.uint8) | (lcv.astype(np.uint8) << 1))
    return raw, tgt, mask


def fit_norms(cb: Basin, raw: np.ndarray, tgt: np.ndarray) -> Norms:
    """Input shift/scale from the wet points' mean/std; output scale = the
    targets' RMS per component."""
    lcu, lcv = uv_masks(cb)
    wet = cb.wet
    shift = np.zeros(4, dtype=np.float32)
    scale = np.ones(4, dtype=np.float32)
    for c, m in enumerate((lcu, lcv, wet, wet)):
        vals = raw[:, c][:, m]
        shift[c] = vals.mean() if c == 3 else 0.0
        sd = float(np.sqrt(((vals - shift[c]) ** 2).mean()))
        scale[c] = 1.0 / sd if sd > 0 else 1.0
    out = np.array([np.sqrt((tgt[..., 0][:, lcu] ** 2).mean()), np.sqrt((tgt[..., 1][:, lcv] ** 2).mean())],
                   dtype=np.float32)
    out[out == 0] = 1.0
    return Norms(shift, scale, out)


def load_dataset(net: ClosureNet, raw: np.ndarray, tgt: np.ndarray, mask: np.ndarray, nm: Norms,
                 offset: int = 0):
    for k in range(raw.shape[0]):
        x = features(raw[k, 0], raw[k, 1], raw[k, 2], raw[k, 3], nm.shift, nm.scale)
        net.dataset_set(offset + k, x, tgt[k] / nm.out_scale, mask)


def train(net: ClosureNet, ntrain: int, steps: int, lr: float = 1e-3, seed: int = 0,
          log_every: int = 0) -> List[float]:
    """Adam on random batches of the first ``ntrain`` device samples."""
    rng = np.random.default_rng(seed)
    losses = []
    for it in range(steps):
        idx = rng.choice(ntrain, size=net.batch, replace=net.batch > ntrain)
        losses.append(net.train_step(idx, lr))
        if log_every and (it + 1) % log_every == 0:
            print(f"  train step {it + 1}: loss {np.mean(losses[-log_every:]):.4f}", flush=True)
    return losses


def rel(a: np.ndarray, b: np.ndarray, m: np.ndarray) -> float:
    d = np.linalg.norm((a - b)[m])
    n = np.linalg.norm(b[m])
    return float(d / n) if n > 0 else float(d)


def rollout(cb: Basin, start: Snapshot, nsteps: int, truth: Optional[Dict[int, Dict[str, np.ndarray]]] = None,
            net: Optional[ClosureNet] = None, ratio: float = 3.0, mode: str = "fp64", nstep: int = 2,
            interval: int = 1, timed: bool = False) -> dict:
    """Coarse run of ``nsteps`` outer steps from ``start`` (C(t), C(t - dt_c / nstep)).
    ``truth`` maps coarse step -> coarsened fine state; errors are scored at
    those steps.  ``interval``: the closure after every interval-th step
    (with interval steps' worth of its tendency).  Returns errors, the final state, fallback counts and the
    run's wall time (ms) when ``timed`` (after the ensemble's autotuner has
    picked the launch shape from the same state; results do not depend on
    the shape)."""
    lcu, lcv = uv_masks(cb)
    wet = cb.wet
    res: dict = {"steps": [], "ssh": [], "u": [], "v": []}
    with Ensemble([cb], mode=mode) as ens:
        ens.set_state(0, start.now["u"], start.prev["u"], start.now["v"], start.prev["v"], start.now["ssh"],
                      start.prev["ssh"])
        if net is not None:
            ens.set_closure(net, ratio, interval)
        if timed:
            ens.autotune(nstep=nstep, trial_steps=20, rows=(0, 2, 3, 4, 6, 8), waves=(0, 256, 384, 512))
        marks = sorted(truth) if truth else []
        marks = [k for k in marks if 0 < k <= nsteps]
        if not marks or marks[-1] != nsteps:
            marks.append(nsteps)
        done, ms = 0, 0.0
        for k in marks:
            if timed:
                ms += ens.run_timed(k - done, nstep=nstep, tau=cb.dt)
            else:
                ens.run(k - done, nstep=nstep, tau=cb.dt)
            done = k
            if truth and k in truth:
                st = ens.get_state(0)
                t = truth[k]
                res["steps"].append(k)
                res["ssh"].append(rel(st["ssh"], t["ssh"], wet))
                res["u"].append(rel(st["u"], t["u"], lcu))
                res["v"].append(rel(st["v"], t["v"], lcv))
        res["final"] = ens.get_state(0)
        res["bad"] = ens.range_check()
        res["ms"] = ms
        if net is not None:
            nf, nr = ens.closure_stats()
            res["fallback_steps"] = int(nf[0])
            res["norms"] = nr[0].tolist()
    return res
