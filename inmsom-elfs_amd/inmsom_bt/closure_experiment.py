"""Learned-closure experiment on the 2048 x 1536 basin (python -m inmsom_bt.closure_experiment).

1. Fine truth: the 1 km basin (single_basin('cfg1')), FP64 fused step,
   dt = 10 s, nstep_barotrop = 2.  Snapshots every ``every`` fine steps
   (default 2000 x 50), each block-averaged 4x to 512 x 384 at 4 km at
   t - 20 s, t and t + 40 s.
2. Targets: one plain coarse step (dt_c = 40 s, two 20 s leap-frog
   substeps) from (C(t), C(t - 20 s)); target = (C(t + 40 s) - plain) / 40 s
   on U/V points.
3. Training on the device: Adam, lr 1e-3, batch 32, on the first
   ``train_frac`` of the snapshots (chronological split); the rest is the
   validation set.
4. Rollout: from the fine run's state after the last snapshot, 10^4 coarse
   steps of 40 s with and without the closure, scored against the
   coarsened fine run; both coarse runs and the fine run are also timed.
5. Ensemble of 128 coarse members (sizes 128x128 .. 512x384, 4 km) in FP32
   with the closure fused into the graph-replayed step, against a plain
   FP64 run of the same members.

Writes <out>/closure_params.npz (parameters + normalisation) and
<out>/closure_results.json.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from .closure import ClosureNet
from .ensemble import Ensemble
from .learned import (CoarseStepper, FineTruth, Norms, fit_norms, load_dataset, rel, rollout, samples, snapshots,
                      train)
from .workloads import NSTEP_BAROTROP, coarse_ensemble, single_basin


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def init_params(seed: int, nparam: int, out_scale: float) -> np.ndarray:
    """He-normal on used entries (input channels 0..3 of layer 0), zero
    biases, last layer scaled by out_scale; layout of inmsom_bt_cnn_set_params."""
    rng = np.random.default_rng(seed)
    p = np.zeros(nparam, dtype=np.float32)
    off = 0
    for l, cin in enumerate((8, 32, 32, 32)):
        kp = (9 * cin + 15) // 16 * 16
        used = 4 if l == 0 else cin
        w = np.zeros((32, 9, cin))
        w[:, :, :used] = rng.standard_normal((32, 9, used)) * np.sqrt(2.0 / (9 * used))
        if l == 3:
            w *= out_scale
        full = np.zeros((32, kp))
        full[:, :9 * cin] = w.reshape(32, -1)
        p[off:off + 32 * kp] = full.reshape(-1)
        off += 32 * kp + 32
    assert off == nparam
    return p


def save_params(path: str, net: ClosureNet, nm: Norms, meta: dict):
    np.savez(path, params=net.params(), shift=nm.shift, scale=nm.scale, out_scale=nm.out_scale,
             meta=json.dumps(meta))


def load_params(path: str):
    z = np.load(path, allow_pickle=False)
    return z["params"], Norms(z["shift"], z["scale"], z["out_scale"]), json.loads(str(z["meta"]))


def build_dataset(truth: FineTruth, steps, chunk: int = 64, keep=()):
    """float32 raw inputs [n, 4, ny, nx], targets [n, ny, nx, 2], mask, and
    the snapshots whose indices are in ``keep``."""
    cb = truth.coarse
    raws, tgts = [], []
    mask = None
    kept = {}
    stepper = CoarseStepper(cb, min(chunk, len(steps)))
    try:
        for i0 in range(0, len(steps), chunk):
            snaps = snapshots(truth, steps[i0:i0 + chunk])
            raw, tgt, mask = samples(cb, snaps, stepper(snaps))
            raws.append(raw.astype(np.float32))
            tgts.append(tgt)
            for k, sn in enumerate(snaps):
                if i0 + k in keep:
                    kept[i0 + k] = sn
            if (i0 // chunk) % 8 == 0:
                log(f"  snapshots {i0 + len(snaps)}/{len(steps)} (fine step {truth.step})")
    finally:
        stepper.close()
    return np.concatenate(raws), np.concatenate(tgts), mask, kept


def choose_strength(cb, net: ClosureNet, nm: Norms, start, truth: dict, ratio: float,
                    alphas=(0.0625, 0.125, 0.25, 0.5, 1.0)):
    """Closure strength alpha (tendency x alpha) from rollouts inside the
    training period: the largest alpha whose SSH error, averaged over the
    checkpoints in ``truth`` (coarse step -> coarsened state), is no larger
    than the plain model's and whose run stays in range; the smallest alpha
    when none qualifies.  Returns (alpha, per-alpha scores, plain score)."""
    if start is None or not truth:
        return 1.0, {}, None
    steps = max(truth)

    def score(r):
        return float(np.mean(r["ssh"])) if r["bad"] == 0 and np.all(np.isfinite(r["ssh"])) else float("inf")
    plain = score(rollout(cb, start, steps, truth, None, ratio, "fp64"))
    errs = {}
    for a in alphas:
        net.set_norm(nm.shift, nm.scale, nm.out_scale * np.float32(a))
        r = rollout(cb, start, steps, truth, net, ratio, "fp64")
        errs[a] = {"ssh_mean": score(r), "u_mean": float(np.mean(r["u"])), "v_mean": float(np.mean(r["v"]))}
    net.set_norm(nm.shift, nm.scale, nm.out_scale)
    ok = [a for a in alphas if errs[a]["ssh_mean"] <= plain]
    return (max(ok) if ok else min(alphas)), errs, plain


def val_loss(net: ClosureNet, first: int, n: int) -> float:
    """Mean loss over held-out device samples first..first+n-1 (batches of net.batch)."""
    losses = []
    for i0 in range(first, first + n - net.batch + 1, net.batch):
        losses.append(net.eval(np.arange(i0, i0 + net.batch)))
    return float(np.mean(losses)) if losses else float("nan")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--snapshots", type=int, default=2000)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--train-frac", type=float, default=0.9)
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--rollout", type=int, default=10000, help="coarse steps of 40 s")
    ap.add_argument("--score-every", type=int, default=500)
    ap.add_argument("--ratio", type=float, default=3.0,
                    help="ensemble fallback: closure norm above ratio x plain viscous tendency norm")
    ap.add_argument("--rollout-ratio", type=float, default=0.0,
                    help="fallback ratio of the single-basin rollouts (0 = closure always applied)")
    ap.add_argument("--tune-steps", type=int, default=4000, help="coarse steps of the strength-selection rollout")
    ap.add_argument("--members", type=int, default=128)
    ap.add_argument("--member-steps", type=int, default=2000)
    ap.add_argument("--workload", default="cfg1")
    ap.add_argument("--out", default="gpurun_out/closure")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--output", default="uv", choices=["uv", "psi"],
                    help="network outputs: u, v tendencies, or a stream-function tendency (divergence-free)")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    res: dict = {"args": vars(args)}
    T = time.perf_counter

    fine = single_basin(args.workload)
    truth = FineTruth(fine)
    cb = truth.coarse
    f = truth.factor
    log(f"fine {fine.nx - 4}x{fine.ny - 4} ({fine.n_wet} wet), coarse {cb.nx - 4}x{cb.ny - 4} "
        f"({cb.n_wet} wet), dt_c {cb.dt} s")
    res["fine"] = {"interior": [fine.nx - 4, fine.ny - 4], "wet": fine.n_wet, "dt": fine.dt}
    res["coarse"] = {"interior": [cb.nx - 4, cb.ny - 4], "wet": cb.n_wet, "dt": cb.dt}

    # ---- 1-2: snapshots and targets
    t0 = T()
    steps = [f + k * args.every for k in range(args.snapshots)]
    ntrain = max(args.batch, int(round(args.train_frac * len(steps))))
    # strength selection: a rollout from the middle of the training period,
    # scored at later training snapshots every `score_every` coarse steps
    # (snapshots are every `every` fine steps)
    v0 = ntrain // 2
    per_mark = args.score_every * f // args.every if (args.score_every * f) % args.every == 0 else 0
    nmarks = min(args.tune_steps // args.score_every, (ntrain - 1 - v0) // per_mark) if per_mark else 0
    tune_idx = [v0 + per_mark * j for j in range(1, nmarks + 1)]
    raw, tgt, mask, kept = build_dataset(truth, steps, keep={v0, *tune_idx} if tune_idx else ())
    n = raw.shape[0]
    nm = fit_norms(cb, raw[:ntrain], tgt[:ntrain])
    if args.output == "psi":
        # one scale for psi: both tendencies are its differences
        nm.out_scale[:] = np.float32(np.sqrt(np.mean(nm.out_scale.astype(np.float64) ** 2)))
    res["dataset"] = {"samples": n, "train": ntrain, "validation": n - ntrain, "seconds": T() - t0,
                      "norm_shift": nm.shift.tolist(), "norm_scale": nm.scale.tolist(),
                      "target_rms": nm.out_scale.tolist()}
    log(f"dataset: {n} samples in {T() - t0:.1f} s; target rms {nm.out_scale}")

    # ---- 3: training
    net = ClosureNet(args.batch, cb.ny, cb.nx)
    net.set_output(args.output)
    net.set_norm(nm.shift, nm.scale, nm.out_scale)
    net.set_params(init_params(args.seed, net.nparam, out_scale=0.1))
    net.dataset_alloc(n)
    load_dataset(net, raw, tgt, mask, nm)
    del raw
    t0 = T()
    losses = train(net, ntrain, args.train_steps, args.lr, seed=args.seed, log_every=100)
    net.synchronize()
    train_s = T() - t0
    vl = val_loss(net, ntrain, n - ntrain)
    rr = args.rollout_ratio if args.rollout_ratio > 0 else 1e300
    tune_truth = {args.score_every * (j + 1): kept[ix].now for j, ix in enumerate(tune_idx)}
    alpha, aerr, perr = choose_strength(cb, net, nm, kept.get(v0) if tune_idx else None, tune_truth, rr)
    res["strength"] = {"alpha": alpha, "tune_steps": max(tune_truth) if tune_truth else 0,
                       "by_alpha": {str(k): v for k, v in aerr.items()}, "ssh_mean_plain": perr,
                       "start_step": steps[v0] if tune_idx else None,
                       "rule": "largest alpha whose mean SSH error over the window is <= plain's"}
    log(f"strength: alpha {alpha} (scores {aerr}, plain ssh mean {perr}) over {res['strength']['tune_steps']} coarse steps")
    nm = Norms(nm.shift, nm.scale, (nm.out_scale * np.float32(alpha)).astype(np.float32))
    net.set_norm(nm.shift, nm.scale, nm.out_scale)
    res["training"] = {"steps": args.train_steps, "seconds": train_s, "ms_per_step": 1e3 * train_s / args.train_steps,
                       "loss_first": float(np.mean(losses[:10])), "loss_last": float(np.mean(losses[-50:])),
                       "val_loss": vl, "loss_curve": [float(np.mean(losses[i:i + 50]))
                                                     for i in range(0, len(losses), 50)]}
    log(f"training: {args.train_steps} steps in {train_s:.1f} s, loss {res['training']['loss_first']:.4f} -> "
        f"{res['training']['loss_last']:.4f}, validation {vl:.4f} (1.0 = no skill)")
    # the held-out rollout starts after the last training snapshot
    t_start = steps[-1] + args.every
    save_params(os.path.join(args.out, "closure_params.npz"), net, nm,
                {"workload": args.workload, "snapshots": n, "every": args.every, "train": ntrain,
                 "train_steps": args.train_steps, "lr": args.lr, "batch": args.batch, "t_start": t_start,
                 "val_loss": vl, "seed": args.seed, "alpha": alpha, "output": args.output})

    # ---- 4: rollout against the coarsened truth
    start = snapshots(truth, [t_start])[0]
    marks = list(range(args.score_every, args.rollout + 1, args.score_every))
    tr = {}
    t0 = T()
    for k in marks:
        truth.advance_to(t_start + f * k)
        tr[k] = truth.coarse_state()
    log(f"truth: {args.rollout * f} fine steps in {T() - t0:.1f} s")
    runs = {}
    for name, mode, use in (("plain_fp64", "fp64", False), ("learned_fp64", "fp64", True),
                            ("plain_fp32", "fp32", False), ("learned_fp32", "fp32", True)):
        r = rollout(cb, start, args.rollout, tr, net if use else None, rr, mode)
        r.pop("final")
        runs[name] = r
        log(f"rollout {name}: ssh rel err @end {r['ssh'][-1]:.4f}, mean {np.mean(r['ssh']):.4f}; "
            f"u {r['u'][-1]:.4f}, v {r['v'][-1]:.4f}" +
            (f", fallback steps {r['fallback_steps']}, last-step norms closure/viscous {r['norms']}" if use else ""))
    res["rollout"] = runs

    # ---- throughput: fine FP64 vs coarse with/without closure (simulated seconds per wall second)
    tm = {}
    with truth.model as m:
        m.run(20, "fp64", nstep=NSTEP_BAROTROP)
        ms = m.run_timed(2000, "fp64", nstep=NSTEP_BAROTROP)
        tm["fine_fp64"] = {"ms_per_step": ms / 2000, "sim_s_per_s": fine.dt * 2000 / (ms * 1e-3),
                           "wet_cell_steps_per_s": fine.n_wet * 2000 * NSTEP_BAROTROP / (ms * 1e-3)}
    for name, mode, use in (("coarse_plain_fp64", "fp64", False), ("coarse_learned_fp64", "fp64", True),
                            ("coarse_plain_fp32", "fp32", False), ("coarse_learned_fp32", "fp32", True)):
        r = rollout(cb, start, 2000, None, net if use else None, rr, mode, timed=True)
        ms = r["ms"]
        tm[name] = {"ms_per_step": ms / 2000, "sim_s_per_s": cb.dt * 2000 / (ms * 1e-3),
                    "wet_cell_steps_per_s": cb.n_wet * 2000 * NSTEP_BAROTROP / (ms * 1e-3)}
    for k, v in tm.items():
        v["speedup_vs_fine"] = v["sim_s_per_s"] / tm["fine_fp64"]["sim_s_per_s"]
        log(f"throughput {k}: {v['ms_per_step'] * 1e3:.1f} us/step, {v['sim_s_per_s']:.0f} simulated s per s "
            f"({v['speedup_vs_fine']:.1f}x fine)")
    res["throughput"] = tm

    # ---- 5: 128-member coarse ensemble, closure fused into the FP32 step
    if args.members > 0:
        t0 = T()
        members = coarse_ensemble(args.members)
        log(f"members: {len(members)} generated in {T() - t0:.1f} s")
        with Ensemble(members, mode="fp32") as ens, Ensemble(members, mode="fp64") as ref:
            nmem, wet, cells = ens.size()
            ens.set_closure(net, args.ratio)
            ens.run(10, NSTEP_BAROTROP)
            ens.reset()
            ms = ens.run_timed(args.member_steps, NSTEP_BAROTROP)
            ref.run(args.member_steps, NSTEP_BAROTROP)
            nf, nr = ens.closure_stats()
            errs = []
            for k, bk in enumerate(members):
                a, b = ens.get_state(k), ref.get_state(k)
                errs.append(rel(a["ssh"], b["ssh"], bk.wet))
        res["ensemble"] = {"members": nmem, "wet_cells": wet, "array_cells": cells, "steps": args.member_steps,
                           "ms_per_step": ms / args.member_steps,
                           "wet_cell_steps_per_s": wet * args.member_steps * NSTEP_BAROTROP / (ms * 1e-3),
                           "ssh_rel_err_vs_plain_fp64": {"median": float(np.median(errs)),
                                                         "max": float(np.max(errs)),
                                                         "per_member": [float(e) for e in errs]},
                           "fallback_steps": {"total": int(nf.sum()), "members_with_fallback": int((nf > 0).sum()),
                                              "per_member": nf.tolist()}}
        log(f"ensemble: {nmem} members, {ms / args.member_steps * 1e3:.0f} us/step, "
            f"{res['ensemble']['wet_cell_steps_per_s']:.3e} wet-cell steps/s, ssh err vs fp64 median "
            f"{np.median(errs):.3e} max {np.max(errs):.3e}, fallback members {(nf > 0).sum()}")
    with open(os.path.join(args.out, "closure_results.json"), "w") as fo:
        json.dump(res, fo, indent=1)
    net.close()
    print(json.dumps({k: res[k] for k in res if k in ("training", "throughput")}), flush=True)


if __name__ == "__main__":
    main()
