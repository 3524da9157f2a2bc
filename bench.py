#!/usr/bin/env python3
"""Throughput of the fused barotropic step on one MI355X.

Workload (inmsom_bt.workloads.single_basin('cfg1')): a 2048 x 1536 closed
basin at 1 km, 10 s outer steps, barotropic mode sub-cycled twice
(5 s leap-frog substeps).  A bench "step" is one outer step = two
leap-frog substeps, which FP64 and FP32 advance in one fused launch (the
second substep takes the first's rows from registers) and FP64_EXACT in
two; the reported value counts state-advancing leap-frog substeps:

    value = wet cells * steps * nstep_barotrop / timed seconds

Printed (stdout, one JSON line): FP64 (the reference's precision) as the
headline value, with FP32 and FP64_EXACT alongside, each with its roofline
figure, the FP32 / FP64 errors after the timed run, and the oracle (the
C restatement of the reference, OpenMP) timed on a bounded sample of the
same workload on this host's cores.  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "inmsom-elfs_amd"))

import numpy as np  # noqa: E402

from inmsom_bt import BarotropicModel  # noqa: E402
from inmsom_bt.model import gridcon  # noqa: E402
from inmsom_bt.workloads import NSTEP_BAROTROP, single_basin  # noqa: E402

METRIC = "wet-cell steps/s on 1 GPU: plain FP64 vs FP32/batched/learned at fixed SSH err"
PEAK_HBM_GBS = 8000.0

# Algorithmic bytes per streamed cell per substep (the minimal count of
# what a substep must move): level n and filtered level n-1 of (u, v, ssh) read and written
# (FP64: 12 x 8 B; FP32 stencils: level n 6 x 4 B, the FP64 filtered level
# 6 x 8 B), the depth at rest and the Rayleigh coefficient read (2 x 8 B;
# FP32 2 x 4 B).  Wind stress, Coriolis and viscosity are row tables or
# constants (0 B).  A launch of two substeps is credited twice this; it
# moves far less, since the intermediate level never leaves the registers
# (roofline.traffic).
BYTES_PER_CELL = {
    "fp64": 12 * 8 + 8 + 8,                # 112
    "fp64_exact": 12 * 8 + 8 + 8,          # 112
    "fp32": 6 * 4 + 6 * 8 + 4 + 4,         # 80
}
# The same count with the two RHS components frozen over a call, which a
# substep of the reference reads (written by the first substep of a call,
# read by the second: one access per component per substep on average with
# nstep_barotrop = 2); reported beside the minimal count, never as it.
BYTES_PER_CELL_WITH_RHS = {"fp64": 128, "fp64_exact": 128, "fp32": 88}
# Warm-up before a timed region: at least the requested warm-up steps and
# at least this many seconds of them, so that the clocks have settled.
WARMUP_SECONDS = 2.0
# Warm-up of the side legs' timed repetitions.  A timed region that starts
# after the GPU idled for 5 ms or more (host copies, set-up) runs 5-20 %
# slower for its first ~60 ms while the clocks come back (measured on
# MI355X: a 20-step FP64 region after 10-50 ms idle 133-141 us/step, after
# <= 2 ms 116-120 us/step), so every timed region follows at least this
# much of the same work.
LEG_WARMUP_SECONDS = 0.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "fused_step_cfg1_traffic.json")
# the kernels each leg launches (two substeps per launch in FP64 and FP32; the
# last template argument marks the call's first, forcing-forming substep)
TRAFFIC_KERNELS = {"fp64": ("k_fused_pipe<double, 2>",),
                   "fp32": ("k_fused_pipe<float, 2>",),
                   "fp64_exact": ("k_fused_step<double, true, true>", "k_fused_step<double, true, false>"),
                   "ensemble_fp32": ("k_fused_batch_pipe<float>",),
                   "forcing": ("k_forcing<",),
                   "dropin": ("k_dropin<true, true, false>", "k_dropin<false, true, false>"),
                   "vertical_velocity": ("k_vertical(",),
                   "pressure_gradients": ("k_pressure(",),
                   "density": ("k_density(",),
                   "smagorinsky": ("k_smag(",)}


# kernels of one 2-D -> 3-D coupling and their launches per coupling
COUPLING_SEQUENCE = {"k_barotropic_part(": 2, "k_internal_uv(": 1, "k_ssh_internal(ibt::DevGrid, double": 1,
                     "k_ssh_internal_shift(": 1, "k_hh_update(": 1}
# kernels of one baroclinic_dynamics call (nz = 20)
BAROCLINIC_SEQUENCE = {"k_bc_solve<20, true>": 1, "k_bc_frame(": 1}
SEQUENCES = {"coupling": COUPLING_SEQUENCE, "baroclinic": BAROCLINIC_SEQUENCE}


def write_traffic(fetch_path, write_path, out=TRAFFIC_JSON):
    """HBM bytes per launch of the fused kernels (and per coupling) from two
    rocprofv3 --pmc passes of this script (FETCH_SIZE, then WRITE_SIZE; KiB
    per dispatch; counter_collection.csv or the rocpd results.db), each
    scaled by what the counter reads for k_copy_d in the same run: the FP64
    set-up copies 2052 x 1540 doubles (read once, written once).  Entries
    found in these files replace the same entries of ``out``; the others
    are kept."""
    import collections
    import csv
    import sqlite3

    def load(path, counter):
        agg = collections.defaultdict(list)
        if path.endswith(".db"):
            con = sqlite3.connect(path)
            rows = con.execute("select name, counter_value from pmc_events where counter_name = ?", (counter,))
            for name, v in rows:
                agg[name].append(float(v))
            con.close()
        else:
            with open(path) as f:
                for r in csv.DictReader(f):
                    if r["Counter_Name"] == counter:
                        agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        return {k: (sum(v) / len(v), len(v)) for k, v in agg.items()}

    def find(d, key):
        ks = [k for k in d if key in k]
        if len(ks) != 1:
            raise KeyError(f"{key}: {ks}")
        return d[ks[0]], ks[0]

    F, W = load(fetch_path, "FETCH_SIZE"), load(write_path, "WRITE_SIZE")
    kib = 2052 * 1540 * 8 / 1024
    (fc, _), _ = find(F, "k_copy_d")
    (wc, _), _ = find(W, "k_copy_d")
    ff, wf = kib / fc, kib / wc
    src = {"fetch_file": os.path.relpath(fetch_path, ROOT), "write_file": os.path.relpath(write_path, ROOT),
           "calibration": {"kernel": "k_copy_d", "bytes_each_way": 2052 * 1540 * 8, "fetch_kib": fc,
                           "write_kib": wc, "fetch_factor": ff, "write_factor": wf}}
    try:
        with open(out) as f:
            res = json.load(f)
    except (OSError, ValueError):
        res = {}
    for mode, keys in TRAFFIC_KERNELS.items():
        try:
            parts = [(find(F, k), find(W, k)) for k in keys]
        except KeyError:
            continue
        # the mean launch over the mode's kernels, weighted by their dispatches
        nf = sum(pf[0][1] for pf, _ in parts)
        f = sum(pf[0][0] * pf[0][1] for pf, _ in parts) / nf
        w = sum(pw[0][0] * pf[0][1] for pf, pw in parts) / nf
        res[mode] = dict(src, kernel=" + ".join(pf[1] for pf, _ in parts), dispatches=nf, fetch_kib=f,
                         write_kib=w, fetch_bytes_per_launch=f * ff * 1024, write_bytes_per_launch=w * wf * 1024)
    for leg, seq in SEQUENCES.items():
        try:
            parts = {}
            for key, times in seq.items():
                (f, nf), name = find(F, key)
                (w, _), _ = find(W, key)
                parts[name] = {"per_call": times, "dispatches": nf, "fetch_kib": f, "write_kib": w}
        except KeyError:
            continue
        res[leg] = dict(src, kernels=parts,
                        fetch_bytes_per_launch=sum(p["per_call"] * p["fetch_kib"] for p in parts.values()) * ff * 1024,
                        write_bytes_per_launch=sum(p["per_call"] * p["write_kib"] for p in parts.values()) * wf * 1024)
    with open(out, "w") as fo:
        json.dump(res, fo, indent=1)
    return res


def traffic_from_profiles(mode):
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
        e = t[mode]
        return float(e["fetch_bytes_per_launch"] + e["write_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def warm_up(run, sync, steps, seconds=WARMUP_SECONDS):
    """Runs at least ``steps`` warm-up steps and keeps running them, in
    doubling batches, until ``seconds`` have passed; returns (steps run,
    seconds)."""
    t0 = time.perf_counter()
    done, batch = 0, max(int(steps), 1)
    while True:
        run(batch)
        sync()
        done += batch
        if done >= steps and time.perf_counter() - t0 >= seconds:
            return done, time.perf_counter() - t0
        batch = min(batch * 2, 4096)


def timed_mode(m, b, mode, steps, warmup, ncell):
    # the initial state kept on the device: after the warm-up it comes back
    # by a device copy, so the GPU does not idle between the two
    m.set_state(b.u, b.up, b.v, b.vp, b.ssh, b.sshp)
    m.save_state()
    warm_steps, warm_s = warm_up(lambda k: m.run(k, mode, nstep=NSTEP_BAROTROP), m.synchronize, warmup)
    m.restore_state()
    vol0 = m.volume()
    m.synchronize()
    t0 = time.perf_counter()
    ms = m.run_timed(steps, mode, nstep=NSTEP_BAROTROP)
    wall = time.perf_counter() - t0
    bad = m.range_check()
    vol1 = m.volume()
    st = m.get_state()
    subs = m.substeps_per_launch(mode)
    substeps = steps * NSTEP_BAROTROP
    launches = substeps // subs
    avg_launch_s = ms * 1e-3 / launches
    bpc = BYTES_PER_CELL[mode]
    achieved = bpc * subs * ncell / avg_launch_s / 1e9
    traffic = traffic_from_profiles(mode)
    res = {
        "value": b.n_wet * substeps / (ms * 1e-3),
        "ms_per_step": ms / steps,
        "wall_s": wall,
        "avg_launch_us": avg_launch_s * 1e6,
        "substeps_per_launch": subs,
        "warmup_steps_run": warm_steps, "warmup_s": warm_s,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "bytes_per_cell_substep": bpc, "cells_per_launch": ncell, "substeps_per_launch": subs,
                     "frac_with_frozen_rhs": BYTES_PER_CELL_WITH_RHS[mode] * subs * ncell / avg_launch_s / 1e9
                     / PEAK_HBM_GBS},
        "range_check_bad_cells": bad,
        "volume_rel_drift": abs(vol1 - vol0) / max(abs(vol0), 1e-300),
    }
    log(f"{mode}: {ms:.1f} ms for {steps} steps ({avg_launch_s*1e6:.1f} us/launch of {subs} substeps), "
        f"{res['value']:.3e} wet-cell steps/s, {achieved:.0f} GB/s")
    return res, st


def cpu_baseline(b, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    og = oracle.OracleGrid(b)
    rdis = og.rayleigh()
    threads = int(oracle.lib().bto_num_threads())
    t0 = time.perf_counter()
    oracle.run(b, steps, NSTEP_BAROTROP, out_steps=[], rdis=rdis, grid=og)
    dt = time.perf_counter() - t0
    # the oracle runs the reference's barotropic_dynamics: 2 * nstep_barotrop
    # substeps per outer step (nstep state-advancing, nstep look-ahead), each
    # the work of one fused substep; every substep it runs is credited
    subs = 2 * NSTEP_BAROTROP
    return {"value": b.n_wet * steps * subs / dt, "unit": "wet-cell steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{steps} outer steps of the same 2048x1536 basin through the oracle's "
                      f"barotropic_dynamics (oracle/bt_oracle.c, OpenMP), {dt:.1f} s; each outer step runs "
                      f"{subs} leap-frog substeps ({NSTEP_BAROTROP} state-advancing + {NSTEP_BAROTROP} "
                      f"look-ahead) and all {subs} are counted"}


def cpu_reference_fortran(steps):
    """The reference's own Fortran (oracle/_ref/ref_harness, built from its
    sources; the timing mode uses its start_timer / end_timer) and the C
    restatement on the same basin at the reference's compiled 289 x 163,
    one rank x the host's threads (at most 16): the calibration of the
    port-based cpu_baseline.  None where the harness is not built."""
    if steps <= 0:
        return None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import time_reference

    if not oracle.have_reference():
        log("cpu reference: oracle/_ref/ref_harness not built, skipped")
        return None
    threads = min(os.cpu_count() or 1, 16)
    r = time_reference.measure(steps, [threads])
    t = r["threads"][str(threads)]
    log(f"cpu reference: Fortran {t['fortran_rate']:.3e}, C {t['c_rate']:.3e} wet-cell steps/s on {threads} "
        f"threads ({r['basin']})")
    return {"value": t["fortran_rate"], "unit": r["unit"], "cores": threads, "kind": "reference",
            "ranks": 1, "c_port_value": t["c_rate"], "c_over_fortran": t["c_over_fortran"],
            "sample": f"{steps} outer steps of the {r['basin']} (the reference's compiled grid) through "
                      f"oracle/_ref/ref_harness mode 9, timed with the reference's start_timer/end_timer; "
                      f"{r['substeps_per_outer_step']} leap-frog substeps counted per outer step; the C "
                      f"restatement on the same steps and threads beside it"}


def ensemble_leg(steps, warmup, nmembers):
    """cfg2: uneven basins advanced together in FP32 (one persistent,
    graph-replayed launch per substep), launch shape from the autotuner."""
    from inmsom_bt.ensemble import Ensemble
    from inmsom_bt.workloads import ensemble

    t0 = time.perf_counter()
    basins = ensemble(nmembers)
    log(f"ensemble: {nmembers} members generated in {time.perf_counter() - t0:.1f} s")
    with Ensemble(basins, mode="fp32") as ens:
        nmem, wet, cells = ens.size()
        tune = ens.autotune(nstep=NSTEP_BAROTROP, trial_steps=10)
        log(f"ensemble autotune: rows/wave {tune['rows_per_wave']}, waves/XCD {tune['waves_per_xcd']}")
        # the members' initial volumes first (one short reduction and host
        # sync each): the timed steps follow the warm-up and the device-side
        # reset without the GPU idling in between
        ens.reset()
        ens.save()
        v0 = [ens.volume(k) for k in range(nmem)]
        warm_steps, warm_s = warm_up(lambda k: ens.run(k, NSTEP_BAROTROP), lambda: None, warmup)
        ens.restore()
        ms = ens.run_timed(steps, NSTEP_BAROTROP)
        bad = ens.range_check()
        drift = max(abs(ens.volume(k) - v0[k]) / abs(v0[k]) for k in range(nmem))
        subs = ens.substeps_per_launch()
    substeps = steps * NSTEP_BAROTROP
    launches = substeps // subs
    avg = ms * 1e-3 / launches
    bpc = BYTES_PER_CELL["fp32"]
    achieved = bpc * subs * cells / avg / 1e9
    res = {"value": wet * substeps / (ms * 1e-3), "unit": "wet-cell steps/s",
           "members": nmem, "wet_cells": wet, "array_cells": cells, "steps": steps,
           "ms_per_step": ms / steps, "avg_launch_us": avg * 1e6, "substeps_per_launch": subs,
           "launch": {"rows_per_wave": tune["rows_per_wave"], "waves_per_xcd": tune["waves_per_xcd"]},
           "warmup_steps_run": warm_steps, "warmup_s": warm_s,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_from_profiles("ensemble_fp32"),
                        "bytes_per_cell_substep": bpc, "cells_per_launch": cells, "substeps_per_launch": subs,
                        "frac_with_frozen_rhs": BYTES_PER_CELL_WITH_RHS["fp32"] * subs * cells / avg / 1e9
                        / PEAK_HBM_GBS},
           "range_check_bad_cells": bad, "max_volume_rel_drift": drift}
    log(f"ensemble fp32: {ms:.1f} ms for {steps} steps ({avg * 1e6:.1f} us/launch), "
        f"{res['value']:.3e} wet-cell steps/s, {achieved:.0f} GB/s")
    return res


# nz of the reference's configuration (Inc/basinpar.fi:4-6)
COUPLING_NZ = 20
# Algorithmic bytes of one coupling (remove + update) per array cell for
# the 2-D fields, each plane touched once per kernel that uses it (float
# masks 4 B, fields 8 B): removal 24 (lcu, lcv, uu2d, vv2d), 2-D velocity
# and pgrx/pgry 112, continuity 76, hh_update 128 (depth at rest, ssh_new,
# eight metrics, four masks read; four depths written), Asselin shift 44,
# re-add 24.  The 3-D part is 2 passes x (uu, vv) x (read + write) x 8 B
# x nz per open U/V face.
COUPLING_2D_BYTES = 24 + 112 + 76 + 128 + 44 + 24


def coupling_leg(m, b, nz, nrep):
    """The 2-D -> 3-D coupling after the step (Control/dyncall.f90:169-231)
    on the same basin with nz levels: device kernels timed with HIP events
    on the library's stream, and the whole host-array call (PCIe
    included), next to the oracle's C on one host core."""
    rng = np.random.default_rng(7)
    shape = (b.ny, b.nx)
    wet = b.wet.astype(np.float64)
    f = {"ubrtr_i": 5.0 * rng.standard_normal(shape), "vbrtr_i": 5.0 * rng.standard_normal(shape),
         "hhu": 50.0 + np.abs(rng.standard_normal(shape)), "hhv": 50.0 + np.abs(rng.standard_normal(shape)),
         "slpr": 1.0e5 + 100.0 * rng.standard_normal(shape), "wflux": 1e-6 * rng.standard_normal(shape) * wet,
         "uu2d": 0.1 * rng.standard_normal(shape), "vv2d": 0.1 * rng.standard_normal(shape),
         "pgrx": rng.standard_normal(shape), "pgry": rng.standard_normal(shape),
         "ssh_i": 0.2 * rng.standard_normal(shape) * wet, "sshp_i": 0.2 * rng.standard_normal(shape) * wet}
    z = np.zeros(shape)
    f.update(hhqn=z, hhun=z, hhvn=z, hhhn=z)
    uu = rng.standard_normal((nz,) + shape)
    vv = rng.standard_normal((nz,) + shape)
    inp = dict(f, uu=uu.copy(), vv=vv.copy())
    m.remove_barotropic(uu, vv, f["uu2d"], f["vv2d"])          # first call allocates the staging
    m.internal_update(b.dt, f, uu, vv)
    uu[:], vv[:] = inp["uu"], inp["vv"]
    m.synchronize()
    t0 = time.perf_counter()
    m.remove_barotropic(uu, vv, f["uu2d"], f["vv2d"])
    m.internal_update(b.dt, f, uu, vv)
    host_s = time.perf_counter() - t0
    # the device-address entry points on HBM-resident arrays (what a caller
    # whose 3-D fields live on the GPU pays per step, launch and sync included)
    from inmsom_bt.hipmem import DeviceArray
    dev = {k: DeviceArray.from_host(a, m.device) for k, a in f.items()}
    duu, dvv = DeviceArray.from_host(uu, m.device), DeviceArray.from_host(vv, m.device)
    m.remove_barotropic_dev(duu, dvv, dev["uu2d"], dev["vv2d"])
    m.internal_update_dev(b.dt, dev, duu, dvv)
    t0 = time.perf_counter()
    m.remove_barotropic_dev(duu, dvv, dev["uu2d"], dev["vv2d"])
    m.internal_update_dev(b.dt, dev, duu, dvv)
    dev_s = time.perf_counter() - t0
    for a in list(dev.values()) + [duu, dvv]:
        a.free()
    warm_up(lambda k: m.coupling_timed(b.dt, nz, k), lambda: None, 2, LEG_WARMUP_SECONDS)
    ms = m.coupling_timed(b.dt, nz, nrep)
    avg_s = ms * 1e-3 / nrep
    masks = gridcon(b.lbasins)
    faces = int((masks["lcu"] > 0.5).sum() + (masks["lcv"] > 0.5).sum())
    alg = 2 * 2 * 8 * nz * faces + COUPLING_2D_BYTES * b.nx * b.ny
    achieved = alg / avg_s / 1e9
    res = {"levels": nz, "array_cells": b.nx * b.ny, "open_faces": faces,
           "value": b.n_wet * nz / avg_s, "unit": "wet-cell levels coupled/s (device-resident)",
           "device_ms": avg_s * 1e3, "device_array_call_ms": dev_s * 1e3, "host_call_ms": host_s * 1e3,
           "host_call_rate": b.n_wet * nz / host_s,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_from_profiles("coupling"),
                        "bytes_per_launch_sequence": alg}}
    log(f"coupling nz={nz}: device {avg_s * 1e3:.3f} ms ({achieved:.0f} GB/s), device-array calls "
        f"{dev_s * 1e3:.2f} ms, host-array calls {host_s * 1e3:.1f} ms")
    return res, inp


def coupling_cpu(b, inp):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    og = oracle.OracleGrid(b)
    t0 = time.perf_counter()
    oracle.coupling(og, b.dt, inp)
    dt = time.perf_counter() - t0
    nz = inp["uu"].shape[0]
    return {"value": b.n_wet * nz / dt, "unit": "wet-cell levels coupled/s", "cores": 1, "kind": "port",
            "sample": f"one coupling of the same basin with {nz} levels, {dt:.2f} s, oracle/bt_oracle.c "
                      "(bto_remove_barotropic + bto_internal_update, one thread)"}


# Algorithmic bytes of the forcing per array cell: each 3-D input (uu, vv,
# uup, vvp, amuv, pgx, pgy) read once per level (56 B x nz); the 2-D inputs
# read once (uu2d, vv2d, uup2d, vvp2d, amuv2d, hhq, hhu, hhv, hhh, taux,
# tauy, slpr: 96 B; eight metrics: 64 B; masks lu, luu, lcu, lcv, llu, llv
# as float: 24 B; dz is negligible); RHSx2d and RHSy2d read and written
# (32 B).
FORCING_3D_BYTES = 56
FORCING_2D_BYTES = 96 + 64 + 24 + 32


def forcing_fields(b, nz, seed=11):
    """Synthetic level-(n, n-1) 3-D and 2-D fields of the forcing on basin b
    (velocities zero on closed faces, positive viscosities and depths)."""
    rng = np.random.default_rng(seed)
    sh = (b.ny, b.nx)
    masks = gridcon(b.lbasins)
    lcu, lcv = (masks["lcu"] > 0.5).astype(np.float64), (masks["lcv"] > 0.5).astype(np.float64)
    f = {}
    for k, msk in (("uu", lcu), ("vv", lcv), ("uup", lcu), ("vvp", lcv)):
        f[k] = 0.1 * rng.standard_normal((nz,) + sh) * msk
    f["amuv"] = 10.0 + np.abs(rng.standard_normal((nz,) + sh))
    f["pgx"] = 1e-3 * rng.standard_normal((nz,) + sh)
    f["pgy"] = 1e-3 * rng.standard_normal((nz,) + sh)
    for k, msk in (("uu2d", lcu), ("vv2d", lcv), ("uup2d", lcu), ("vvp2d", lcv)):
        f[k] = 0.1 * rng.standard_normal(sh) * msk
    f["amuv2d"] = 10.0 + np.abs(rng.standard_normal(sh))
    for k in ("hhq", "hhu", "hhv", "hhh"):
        f[k] = 50.0 + np.abs(rng.standard_normal(sh))
    f["slpr"] = 1.0e5 + 100.0 * rng.standard_normal(sh)
    f["taux"], f["tauy"] = b.taux, b.tauy
    f["rhsx2d"], f["rhsy2d"] = np.zeros(sh), np.zeros(sh)
    w = np.arange(1, nz + 1) + 0.5
    f["dz"] = w / w.sum()
    return f


def forcing_leg(m, b, nz, nrep):
    """The forcing of the barotropic step from the 3-D model (Control/
    dyncall.f90:37-130) on the same basin with nz levels: the device work
    timed with HIP events on the library's stream, and the whole host-array
    call (PCIe included)."""
    f = forcing_fields(b, nz)
    pg = (f["pgx"], f["pgy"])
    m.forcing_3d(f, f["dz"], f["rhsx2d"], f["rhsy2d"], *pg)          # first call allocates the staging
    m.synchronize()
    t0 = time.perf_counter()
    m.forcing_3d(f, f["dz"], f["rhsx2d"], f["rhsy2d"], *pg)
    host_s = time.perf_counter() - t0
    warm_up(lambda k: m.forcing_3d_timed(nz, k), lambda: None, 2, LEG_WARMUP_SECONDS)
    ms = m.forcing_3d_timed(nz, nrep)
    avg_s = ms * 1e-3 / nrep
    ncell = b.nx * b.ny
    alg = (FORCING_3D_BYTES * nz + FORCING_2D_BYTES) * ncell
    achieved = alg / avg_s / 1e9
    res = {"levels": nz, "array_cells": ncell,
           "value": b.n_wet * nz / avg_s, "unit": "wet-cell levels of forcing/s (device-resident)",
           "device_ms": avg_s * 1e3, "host_call_ms": host_s * 1e3, "host_call_rate": b.n_wet * nz / host_s,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_from_profiles("forcing"),
                        "bytes_per_call": alg}}
    log(f"forcing nz={nz}: device {avg_s * 1e3:.3f} ms ({achieved:.0f} GB/s), host-array call "
        f"{host_s * 1e3:.1f} ms")
    return res, f


def forcing_cpu(b, f):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    og = oracle.OracleGrid(b)
    t0 = time.perf_counter()
    oracle.forcing_3d(og, f, b.cb_nl, b.ebottom)
    dt = time.perf_counter() - t0
    nz = f["uu"].shape[0]
    return {"value": b.n_wet * nz / dt, "unit": "wet-cell levels of forcing/s", "cores": 1, "kind": "port",
            "sample": f"one forcing of the same basin with {nz} levels, {dt:.2f} s, oracle/bt_oracle.c "
                      "(bto_forcing_3d, one thread)"}


# Algorithmic bytes of baroclinic_dynamics per array cell: uu, uup, vv,
# vvp read and written and RHSx, RHSy, RHSx_tran, RHSy_tran read (96 B x
# nz), ww and nu read (16 B x (nz + 1)); 2-D: pgrx, pgry, rdis, four
# stresses, seven depths (112 B), eight metrics and rlh_s (72 B), lcu, lcv
# as float (8 B).
BAROCLINIC_BYTES_NZ = 96 + 16
BAROCLINIC_BYTES_2D = 16 + 112 + 72 + 8


def baroclinic_fields(b, nz, seed=13):
    rng = np.random.default_rng(seed)
    sh = (b.ny, b.nx)
    f = {k: 0.1 * rng.standard_normal((nz,) + sh) for k in ("uu", "uup", "vv", "vvp")}
    f["ww"] = 1e-4 * rng.standard_normal((nz + 1,) + sh)
    f["nu"] = 1e-3 + 1e-2 * rng.random((nz + 1,) + sh)
    for k in ("rhsx", "rhsy", "rhsx_tran", "rhsy_tran"):
        f[k] = 1e3 * rng.standard_normal((nz,) + sh)
    for k in ("pgrx", "pgry"):
        f[k] = 1e3 * rng.standard_normal(sh)
    f["rdis"] = 1e-5 * rng.random(sh)
    for k in ("tx_surf", "ty_surf", "tx_bot", "ty_bot"):
        f[k] = 1e-4 * rng.standard_normal(sh)
    for k in ("hhu", "hhup", "hhun", "hhv", "hhvp", "hhvn", "hhh"):
        f[k] = 40.0 + rng.random(sh)
    dz = 0.5 + rng.random(nz)
    f["dz"] = dz / dz.sum()
    f["hzt"] = (0.5 + rng.random(nz + 1)) / nz
    return f


def baroclinic_leg(m, b, nz, nrep, tau=40.0):
    """baroclinic_dynamics (Function/vel_ssh.f90:794-1027) on the same basin
    with nz levels: device time (HIP events on the library's stream) and
    the whole host-array call (PCIe included)."""
    f = baroclinic_fields(b, nz)
    m.baroclinic_dynamics(tau, f)                                   # stages the fields
    m.synchronize()
    t0 = time.perf_counter()
    m.baroclinic_dynamics(tau, f)
    host_s = time.perf_counter() - t0
    warm_up(lambda k: m.baroclinic_timed(tau, nz, k), lambda: None, 2, LEG_WARMUP_SECONDS)
    ms = m.baroclinic_timed(tau, nz, nrep)
    avg_s = ms * 1e-3 / nrep
    ncell = b.nx * b.ny
    alg = (BAROCLINIC_BYTES_NZ * nz + BAROCLINIC_BYTES_2D) * ncell
    achieved = alg / avg_s / 1e9
    res = {"levels": nz, "array_cells": ncell,
           "value": b.n_wet * nz / avg_s, "unit": "wet-cell levels of baroclinic_dynamics/s (device-resident)",
           "device_ms": avg_s * 1e3, "host_call_ms": host_s * 1e3,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_from_profiles("baroclinic"),
                        "bytes_per_call": alg}}
    log(f"baroclinic nz={nz}: device {avg_s * 1e3:.3f} ms ({achieved:.0f} GB/s), host-array call "
        f"{host_s * 1e3:.1f} ms")
    return res, f


def baroclinic_cpu(b, f, tau=40.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    og = oracle.OracleGrid(b)
    t0 = time.perf_counter()
    oracle.baroclinic(og, tau, f)
    dt = time.perf_counter() - t0
    nz = f["uu"].shape[0]
    return {"value": b.n_wet * nz / dt, "unit": "wet-cell levels of baroclinic_dynamics/s", "cores": 1,
            "kind": "port", "sample": f"one call on the same basin with {nz} levels, {dt:.2f} s, "
                                      "oracle/bt_oracle.c (bto_baroclinic_dynamics, one thread)"}


# Algorithmic bytes of vertical_velocity per array cell: u, v read (16 B x
# nz), w written (8 B x (nz + 1)); hu, hv, wflux, dx, dy, dxh, dyh read
# (56 B) and lu as float (4 B).
def vertical_leg(m, b, nz, nrep):
    """vertical_velocity (Function/vel_ssh.f90:1030-1076) on the same basin
    with nz levels: device time (HIP events) and the host-array call."""
    rng = np.random.default_rng(17)
    sh = (b.ny, b.nx)
    dz = 0.5 + rng.random(nz)
    f = {"dz": dz / dz.sum(), "uu": 0.1 * rng.standard_normal((nz,) + sh),
         "vv": 0.1 * rng.standard_normal((nz,) + sh), "ww": np.zeros((nz + 1,) + sh),
         "hu": 40.0 + rng.random(sh), "hv": 40.0 + rng.random(sh), "wflux": 1e-5 * rng.standard_normal(sh)}
    args = (f["dz"], f["uu"], f["vv"], f["ww"], f["hu"], f["hv"], f["wflux"])
    m.vertical_velocity(*args)
    t0 = time.perf_counter()
    m.vertical_velocity(*args)
    host_s = time.perf_counter() - t0
    warm_up(lambda k: m.vertical_velocity_timed(nz, k), lambda: None, 2, LEG_WARMUP_SECONDS)
    avg_s = m.vertical_velocity_timed(nz, nrep) * 1e-3 / nrep
    alg = (16 * nz + 8 * (nz + 1) + 60) * b.nx * b.ny
    achieved = alg / avg_s / 1e9
    log(f"vertical velocity nz={nz}: device {avg_s * 1e3:.3f} ms ({achieved:.0f} GB/s), host-array call "
        f"{host_s * 1e3:.1f} ms")
    return f, {"levels": nz, "value": b.n_wet * nz / avg_s, "unit": "wet-cell levels of vertical_velocity/s",
            "device_ms": avg_s * 1e3, "host_call_ms": host_s * 1e3,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_from_profiles("vertical_velocity"),
                         "bytes_per_call": alg}}


# Algorithmic bytes of the density definition per array cell: tt, ss read
# and den, den_pot written (32 B x nz); hhq read (8 B) and lu as float (4 B).
def density_leg(m, b, nz, nrep):
    """The density definition (Control/dyncall.f90:437-449) on the same
    basin with nz levels: device time (HIP events) and the host-array call;
    the oracle's one-thread C on the same call as the CPU baseline."""
    rng = np.random.default_rng(23)
    sh = (b.ny, b.nx)
    s3 = (nz,) + sh
    f = {"z": np.sort(rng.random(nz)), "tt": 2.0 + 25.0 * rng.random(s3), "ss": 5.0 + 30.0 * rng.random(s3),
         "hhq": 10.0 + 90.0 * rng.random(sh), "den": np.zeros(s3), "den_pot": np.zeros(s3)}
    args = (f["z"], f["tt"], f["ss"], f["hhq"], f["den"], f["den_pot"])
    m.density(*args)
    t0 = time.perf_counter()
    m.density(*args)
    host_s = time.perf_counter() - t0
    warm_up(lambda k: m.density_timed(nz, k), lambda: None, 2, LEG_WARMUP_SECONDS)
    avg_s = m.density_timed(nz, nrep) * 1e-3 / nrep
    alg = (32 * nz + 12) * b.nx * b.ny
    achieved = alg / avg_s / 1e9
    log(f"density nz={nz}: device {avg_s * 1e3:.3f} ms ({achieved:.0f} GB/s), host-array call "
        f"{host_s * 1e3:.1f} ms")
    return f, {"levels": nz, "value": b.n_wet * nz / avg_s, "unit": "wet-cell levels of the density definition/s",
               "device_ms": avg_s * 1e3, "host_call_ms": host_s * 1e3,
               "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_from_profiles("density"),
                            "bytes_per_call": alg}}


# Algorithmic bytes of the Smagorinsky viscosity per array cell: str_t,
# str_s read and amts, amuv written (32 B x nz); amuv2d written, dx, dy read
# (24 B) and lu as float (4 B).
def smagorinsky_leg(m, b, nz, nrep):
    """The Smagorinsky viscosity and its depth integral
    (Control/dyncall.f90:40-43) on the same basin with nz levels: device
    time (HIP events) and the host-array call; the oracle's one-thread C on
    the same call as the CPU baseline."""
    rng = np.random.default_rng(29)
    sh = (b.ny, b.nx)
    s3 = (nz,) + sh
    dz = 0.5 + rng.random(nz)
    f = {"dz": dz / dz.sum(), "stress_t": 1e-5 * rng.standard_normal(s3), "stress_s": 1e-5 * rng.standard_normal(s3),
         "amts": np.zeros(s3), "amuv": np.zeros(s3), "amuv2d": np.zeros(sh)}
    args = (f["dz"], f["stress_t"], f["stress_s"], f["amts"], f["amuv"], f["amuv2d"], 1.0e3, 5.0e2)
    m.smagorinsky(*args)
    t0 = time.perf_counter()
    m.smagorinsky(*args)
    host_s = time.perf_counter() - t0
    warm_up(lambda k: m.smagorinsky_timed(nz, k), lambda: None, 2, LEG_WARMUP_SECONDS)
    avg_s = m.smagorinsky_timed(nz, nrep) * 1e-3 / nrep
    alg = (32 * nz + 28) * b.nx * b.ny
    achieved = alg / avg_s / 1e9
    log(f"smagorinsky nz={nz}: device {avg_s * 1e3:.3f} ms ({achieved:.0f} GB/s), host-array call "
        f"{host_s * 1e3:.1f} ms")
    # its input, stress_components of the 3-D velocities (Control/dyncall.f90:36):
    # u, v read and str_t, str_s written (32 B x nz); eight metrics read
    # (64 B) and lu, luu as float (8 B)
    uv = {"u": 0.1 * rng.standard_normal(s3), "v": 0.1 * rng.standard_normal(s3)}
    m.stress_components(uv["u"], uv["v"], f["stress_t"], f["stress_s"])
    warm_up(lambda k: m.stress_components_timed(nz, k), lambda: None, 2, LEG_WARMUP_SECONDS)
    st_s = m.stress_components_timed(nz, nrep) * 1e-3 / nrep
    st_alg = (32 * nz + 72) * b.nx * b.ny
    log(f"stress_components nz={nz}: device {st_s * 1e3:.3f} ms ({st_alg / st_s / 1e9:.0f} GB/s)")
    return f, {"levels": nz, "value": b.n_wet * nz / avg_s, "unit": "wet-cell levels of the Smagorinsky viscosity/s",
               "device_ms": avg_s * 1e3, "host_call_ms": host_s * 1e3,
               "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_from_profiles("smagorinsky"),
                            "bytes_per_call": alg},
               "stress_components": {"device_ms": st_s * 1e3,
                                     "roofline": {"bound": "hbm", "achieved": st_alg / st_s / 1e9,
                                                  "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                                  "frac": st_alg / st_s / 1e9 / PEAK_HBM_GBS,
                                                  "bytes_per_call": st_alg}}}


def smagorinsky_cpu(b, f):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    og = oracle.OracleGrid(b)
    t0 = time.perf_counter()
    oracle.smagorinsky(og, f, 1.0e3, 5.0e2)
    dt = time.perf_counter() - t0
    nz = f["stress_t"].shape[0]
    return {"value": b.n_wet * nz / dt, "unit": "wet-cell levels of the Smagorinsky viscosity/s", "cores": 1,
            "kind": "port", "sample": f"one call on the same basin with {nz} levels, {dt:.2f} s, "
                                      "oracle/bt_oracle.c (bto_smagorinsky, one thread)"}


def density_cpu(b, f):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    og = oracle.OracleGrid(b)
    t0 = time.perf_counter()
    oracle.density(og, f)
    dt = time.perf_counter() - t0
    nz = f["tt"].shape[0]
    return {"value": b.n_wet * nz / dt, "unit": "wet-cell levels of the density definition/s", "cores": 1,
            "kind": "port", "sample": f"one call on the same basin with {nz} levels, {dt:.2f} s, "
                                      "oracle/bt_oracle.c (bto_density, one thread)"}


def vertical_cpu(b, f):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    og = oracle.OracleGrid(b)
    t0 = time.perf_counter()
    oracle.vertical_velocity(og, f)
    dt = time.perf_counter() - t0
    nz = f["uu"].shape[0]
    return {"value": b.n_wet * nz / dt, "unit": "wet-cell levels of vertical_velocity/s", "cores": 1,
            "kind": "port", "sample": f"one call on the same basin with {nz} levels, {dt:.2f} s, "
                                      "oracle/bt_oracle.c (bto_vertical_velocity, one thread)"}


def pressure_cpu(b, f):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    og = oracle.OracleGrid(b)
    t0 = time.perf_counter()
    oracle.pressure_gradients(og, f)
    dt = time.perf_counter() - t0
    nz = f["den"].shape[0]
    return {"value": b.n_wet * nz / dt, "unit": "wet-cell levels of pressure_gradients/s", "cores": 1,
            "kind": "port", "sample": f"one call on the same basin with {nz} levels, {dt:.2f} s, "
                                      "oracle/bt_oracle.c (bto_pressure_gradients, one thread)"}


# Algorithmic bytes of pressure_gradients per array cell: den read (8 B x
# nz), rhsx, rhsy read and written (32 B x nz); hhq, hhu, hhv, dxh, dyh read
# (40 B) and lcu, lcv as float (8 B).
def pressure_leg(m, b, nz, nrep):
    """pressure_gradients (Function/vel_ssh.f90:320-376) on the same basin
    with nz levels: device time (HIP events) and the host-array call."""
    rng = np.random.default_rng(19)
    sh = (b.ny, b.nx)
    f = {"z": np.sort(rng.random(nz)),
         "den": 1025.0 + 0.3 * np.arange(1, nz + 1)[:, None, None] + rng.standard_normal((nz,) + sh),
         "rhsx": rng.standard_normal((nz,) + sh), "rhsy": rng.standard_normal((nz,) + sh),
         "hhq": 30.0 + 20.0 * rng.random(sh), "hhu": 30.0 + rng.random(sh), "hhv": 30.0 + rng.random(sh)}
    args = (f["z"], f["den"], f["rhsx"], f["rhsy"], f["hhq"], f["hhu"], f["hhv"])
    m.pressure_gradients(*args)
    t0 = time.perf_counter()
    m.pressure_gradients(*args)
    host_s = time.perf_counter() - t0
    warm_up(lambda k: m.pressure_gradients_timed(nz, k), lambda: None, 2, LEG_WARMUP_SECONDS)
    avg_s = m.pressure_gradients_timed(nz, nrep) * 1e-3 / nrep
    alg = (40 * nz + 48) * b.nx * b.ny
    achieved = alg / avg_s / 1e9
    log(f"pressure gradients nz={nz}: device {avg_s * 1e3:.3f} ms ({achieved:.0f} GB/s), host-array call "
        f"{host_s * 1e3:.1f} ms")
    return f, {"levels": nz, "value": b.n_wet * nz / avg_s, "unit": "wet-cell levels of pressure_gradients/s",
            "device_ms": avg_s * 1e3, "host_call_ms": host_s * 1e3,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_from_profiles("pressure_gradients"),
                         "bytes_per_call": alg}}


def dropin_leg(m, b, calls=3):
    """The reference's own boundary as a host-resident caller uses it: one
    call of barotropic_dynamics (Function/vel_ssh.f90:379-679,
    nstep_barotrop substeps with their look-ahead) on host arrays, the
    state, forcing and outputs crossing PCIe (pageable memory) every call.
    The PCIe-inclusive rate, reported beside the headline, never as it."""
    sh = (b.ny, b.nx)
    st = [np.ascontiguousarray(getattr(b, k), dtype=np.float64).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")]
    acc = [np.zeros(sh) for _ in range(4)]
    rhs = [np.zeros(sh), np.zeros(sh)]
    frc = [np.ascontiguousarray(a, dtype=np.float64) for a in (b.wflux, b.mu, b.mu4, m.rdis)]
    args = (b.dt, NSTEP_BAROTROP, 0, *st, *acc, *rhs, *frc)
    m.barotropic_dynamics(*args)
    times = []
    for _ in range(calls):
        t0 = time.perf_counter()
        m.barotropic_dynamics(*args)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    log(f"drop-in host-array call: {t * 1e3:.1f} ms per call ({NSTEP_BAROTROP} substeps), "
        f"{b.n_wet * NSTEP_BAROTROP / t:.3e} wet-cell steps/s PCIe included")
    return {"value": b.n_wet * NSTEP_BAROTROP / t, "unit": "wet-cell steps/s (host arrays, PCIe included)",
            "call_ms": t * 1e3, "calls": calls, "substeps_per_call": NSTEP_BAROTROP,
            "entry": "inmsom_bt_barotropic_dynamics"}


def dropin_bytes_per_cell(nstep, row_metrics=True):
    """Algorithmic HBM bytes per array cell of one fused barotropic_dynamics
    call (2 * nstep launches): every launch reads the entering state (6 x 8 B),
    the depth at rest, Rayleigh coefficient, viscosity, water flux and the
    caller's two RHS components (6 x 8 B); all but the last write the state
    (48 B) and the two filtered U/V depths they carry (16 B; read by all but
    the first); the four averages are written by every launch and read by all
    but the first (32 B each way); launch nstep + 1 writes the call's result
    (48 B).  Per-cell metrics and Coriolis parameter add 9 x 8 B per launch
    (row tables: 0)."""
    n = 2 * nstep
    per = 6 * 8 + 6 * 8
    total = n * per + (n - 1) * (48 + 16) + (n - 1) * 16 + n * 32 + (n - 1) * 32 + 48
    if not row_metrics:
        total += n * 9 * 8
    return total


def dropin_device_leg(m, b, calls=20, nstep=NSTEP_BAROTROP):
    """The reference's boundary, device-resident: calls of barotropic_dynamics
    (Function/vel_ssh.f90:379-679: nstep state-advancing and nstep look-ahead
    substeps, the four averages, the state after substep nstep returned) on
    the model's state with the RHS of its wind stress and drag, through the
    fused kernel and through one kernel per reference sweep.  A call is
    credited its nstep state-advancing substeps."""
    out = {}
    m.set_state(b.u, b.up, b.v, b.vp, b.ssh, b.sshp)
    for path in ("fused", "literal"):
        m.set_dropin_path(path)
        n = calls if path == "fused" else max(calls // 4, 2)
        warm_up(lambda k: m.dropin_timed(k, nstep=nstep), lambda: None, 2, LEG_WARMUP_SECONDS)
        ms = m.dropin_timed(n, nstep=nstep)
        call_s = ms * 1e-3 / n
        out[path] = {"call_ms": call_s * 1e3, "calls": n,
                     "value": b.n_wet * nstep / call_s, "unit": "wet-cell steps/s (device-resident drop-in)"}
    m.set_dropin_path("auto")
    m.set_state(b.u, b.up, b.v, b.vp, b.ssh, b.sshp)
    bpc = dropin_bytes_per_cell(nstep, row_metrics=True)
    ncell = b.nx * b.ny
    achieved = bpc * ncell / (out["fused"]["call_ms"] * 1e-3) / 1e9
    out["fused"]["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                "frac": achieved / PEAK_HBM_GBS, "bytes_per_cell_call": bpc,
                                "cells": ncell, "launches_per_call": 2 * nstep,
                                "traffic": (lambda t: None if t is None else t * 2 * nstep)(
                                    traffic_from_profiles("dropin"))}
    out["speedup_fused_vs_literal"] = out["literal"]["call_ms"] / out["fused"]["call_ms"]
    # the same calls with the biharmonic viscosity on (ksw4 = 1): uv_diff4 by
    # reference-order kernels before each fused substep, or stage by stage
    for path in ("fused", "literal"):
        m.set_dropin_path(path)
        n = max(calls // 2, 2) if path == "fused" else max(calls // 4, 2)
        warm_up(lambda k: m.dropin_timed(k, nstep=nstep, ksw4=1), lambda: None, 2, LEG_WARMUP_SECONDS)
        ms = m.dropin_timed(n, nstep=nstep, ksw4=1)
        call_s = ms * 1e-3 / n
        out[path + "_ksw4"] = {"call_ms": call_s * 1e3, "calls": n,
                               "value": b.n_wet * nstep / call_s, "unit": "wet-cell steps/s (ksw4 = 1)"}
    m.set_dropin_path("auto")
    m.set_state(b.u, b.up, b.v, b.vp, b.ssh, b.sshp)
    log(f"drop-in device-resident, nstep={nstep}: fused {out['fused']['call_ms']:.3f} ms/call "
        f"({achieved:.0f} GB/s), stage-by-stage {out['literal']['call_ms']:.3f} ms/call; ksw4 = 1: fused "
        f"{out['fused_ksw4']['call_ms']:.3f}, stage-by-stage {out['literal_ksw4']['call_ms']:.3f} ms/call")
    return out


def snapshot_leg(m, b, count):
    """The state written in the reference's output formats (ssh, ubrtr,
    vbrtr as DIOGIN float32 records with GrADS descriptors, one record per
    snapshot) from the device: packing and the copy to pinned memory timed
    with HIP events, and the whole call including the file writes into a
    temporary directory; the host conversion and the host record writer on
    the same state as the CPU baseline."""
    import shutil
    import tempfile

    from inmsom_bt import fileio
    from inmsom_bt.model import gridcon

    nx, ny = b.nx, b.ny
    cells = (nx - 4) * (ny - 4) + (nx - 3) * (ny - 4) + (nx - 4) * (ny - 3)
    out_bytes = 4 * cells
    tmp = tempfile.mkdtemp(prefix="inmsom_snap_")
    try:
        m.write_snapshot(tmp, 1)                                     # allocates the buffers
        dev_ms, wall = [], []
        for r in range(count):
            t0 = time.perf_counter()
            dev_ms.append(m.write_snapshot(tmp, r + 2))
            wall.append(time.perf_counter() - t0)
        st = m.get_state()
        masks = gridcon(b.lbasins)
        t0 = time.perf_counter()
        for name, f, lu, box in (("h_ssh.dat", st["ssh"], masks["lu"], (3, nx - 2, 3, ny - 2)),
                                 ("h_u.dat", st["u"], masks["llu"], (2, nx - 2, 3, ny - 2)),
                                 ("h_v.dat", st["v"], masks["llv"], (3, nx - 2, 2, ny - 2))):
            fileio.wdstd(tmp, name, 1, np.ascontiguousarray(f, dtype=np.float32), lu, box + (1, 1))
        host_s = time.perf_counter() - t0
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    dev_s = float(np.mean(dev_ms)) * 1e-3
    wall_s = float(np.mean(wall))
    res = {"snapshots": count, "bytes_per_snapshot": out_bytes,
           "value": out_bytes / wall_s / 1e9, "unit": "GB/s of snapshot records written (device state to files)",
           "call_ms": wall_s * 1e3, "device_pack_and_copy_ms": dev_s * 1e3,
           "device_pack_and_copy_GBps": out_bytes / dev_s / 1e9,
           "cpu_baseline": {"value": out_bytes / host_s / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
                            "sample": f"FP64 state converted on the host and written with the host DIOGIN "
                                      f"writer (inmsom_bt_wdstd), one snapshot, {host_s * 1e3:.1f} ms"}}
    log(f"snapshot: {out_bytes / 1e6:.1f} MB per snapshot, call {wall_s * 1e3:.2f} ms, device pack+copy "
        f"{dev_s * 1e3:.3f} ms, host writer {host_s * 1e3:.1f} ms")
    return res


CLOSURE_PARAMS = os.path.join(ROOT, "inmsom-elfs_amd", "inmsom_bt", "data", "closure_cfg1.npz")


# Multiply-adds per pixel of one forward pass of the closure network, as
# the layers are defined (3x3 taps): 4 -> 32, 32 -> 32, 32 -> 32, 32 -> 2.
CNN_MACS_PER_PIXEL = 9 * (4 * 32 + 32 * 32 + 32 * 32 + 32 * 2)
# Algorithmic HBM bytes per pixel of one training step, layer by layer with
# the stored precisions (BF16 activations and gradients of 32 channels,
# 64 B; the input's 4 channels 8 B; the two outputs and targets FP32, 8 B
# each; the mask 1 B; dL/dpred of the two outputs in BF16, 4 B):
#   forward        8 + 64 | 64 + 64 | 64 + 64 | 64 + 8            = 400
#   loss           8 + 8 + 1 read, 4 written                       =  21
#   weight grads   layer 3: 4 + 64, layers 2, 1: 64 + 64, layer 0: 64 + 8
#   data grads     layer 3: 4 + 64 (ReLU mask) + 64, layers 2, 1: 64 + 64 + 64
#                                                                  = 912
CNN_TRAIN_BYTES_PER_PIXEL = 400 + 21 + (68 + 128 + 128 + 72) + (132 + 192 + 192)
PEAK_BF16_TFLOPS = 2500.0


def cnn_train_leg(steps=20, batch=32, ny=388, nx=516):
    """Training of the closure network on the device (BF16 MFMA forward,
    data and weight gradients, FP32 Adam) on batches of the learned
    closure's shape (32 coarse 512 x 384 basins with halo), synthetic
    inputs.  Algorithmic work of a step: forward + data gradient + weight
    gradient = 3 x the forward's 2 x CNN_MACS_PER_PIXEL flops per pixel
    (the first layer's data gradient, never formed, counted anyway)."""
    from inmsom_bt.closure import ClosureNet
    from inmsom_bt.closure_experiment import load_params

    params, nm, meta = load_params(CLOSURE_PARAMS)
    rng = np.random.default_rng(3)
    nsamp = batch + 4
    with ClosureNet(batch, ny, nx) as net:
        net.set_output(meta.get("output", "uv"))
        net.set_params(params)
        net.dataset_alloc(nsamp)
        for i in range(nsamp):
            x = np.zeros((ny, nx, 8), dtype=np.float32)
            x[..., :4] = rng.standard_normal((ny, nx, 4))
            net.dataset_set(i, x, 1e-2 * rng.standard_normal((ny, nx, 2)), np.full((ny, nx), 3, dtype=np.uint8))
        idx = [list(rng.choice(nsamp, size=batch, replace=False)) for _ in range(steps + 2)]
        t0 = time.perf_counter()
        k = 0
        while k < 2 or time.perf_counter() - t0 < LEG_WARMUP_SECONDS:
            net.train_step(idx[k % 2])
            k += 1
        t0 = time.perf_counter()
        for k in range(steps):
            net.train_step(idx[2 + k])
        dt = (time.perf_counter() - t0) / steps
    npx = batch * ny * nx
    flops = 3 * 2 * CNN_MACS_PER_PIXEL * npx
    tf = flops / dt / 1e12
    nbytes = CNN_TRAIN_BYTES_PER_PIXEL * npx
    gbs = nbytes / dt / 1e9
    # the step's floor: its bytes at the HBM peak take longer than its flops
    # at the dense BF16 MFMA peak (144 flops per byte against 312)
    log(f"cnn training: {dt * 1e3:.2f} ms/step (batch {batch} x {ny} x {nx}), {gbs:.0f} GB/s, {tf:.0f} TFLOP/s")
    return {"ms_per_step": dt * 1e3, "batch": batch, "ny": ny, "nx": nx, "steps": steps,
            "value": batch / dt, "unit": "training samples/s (512x384 coarse basins with halo)",
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": gbs / PEAK_HBM_GBS, "bytes_per_step": nbytes,
                         "bytes_per_pixel": CNN_TRAIN_BYTES_PER_PIXEL},
            "mfma": {"achieved": tf, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s", "frac": tf / PEAK_BF16_TFLOPS,
                     "flops_per_step": flops}}


# The closure runs after every CLOSURE_INTERVAL-th coarse step with that
# many steps' worth of its tendency.  Measured on MI355X over 10^4 held-out
# steps (SSH / u / v relative L2 against the coarsened 1 km run at the end):
# interval 1, 2 and 4 agree to the fifth digit (0.00739 / 0.0782 / 0.0807),
# at 78.9 / 49.7 / 34.9 us per FP64 coarse step (72.8 / 45.5 FP32).
CLOSURE_INTERVAL = 2
# strengths (x the trained alpha) of the 128-member ensemble under the 3x rule
CLOSURE_ENSEMBLE_SCALES = (1.0, 0.5, 0.25)


def learned_leg(fine, rollout_steps, members, member_steps, ratio=3.0, rollout_ratio=3.0,
                interval=CLOSURE_INTERVAL, score_every=1000):
    """cfg3/cfg4: the trained closure (parameters trained on the device by
    inmsom_bt.closure_experiment, committed) on the 4x-coarsened basin:
    throughput in simulated time vs the 1 km FP64 run, SSH/u/v errors vs
    the coarsened 1 km run with and without the closure over a held-out
    period (scored every `score_every` coarse steps); then the closure
    fused into a 128-member FP32 coarse ensemble under the 3x rule, at the
    trained strength and scaled down, with the fraction of member-steps on
    which it was active."""
    from inmsom_bt.closure import ClosureNet
    from inmsom_bt.closure_experiment import load_params
    from inmsom_bt.ensemble import Ensemble
    from inmsom_bt.learned import FineTruth, rel, rollout, snapshots
    from inmsom_bt.workloads import coarse_ensemble

    if not os.path.exists(CLOSURE_PARAMS):
        log(f"learned leg skipped: {CLOSURE_PARAMS} missing")
        return None
    params, nm, meta = load_params(CLOSURE_PARAMS)
    truth = FineTruth(fine)
    cb = truth.coarse
    net = ClosureNet(1, cb.ny, cb.nx)
    net.set_output(meta.get("output", "uv"))
    net.set_params(params)
    net.set_norm(nm.shift, nm.scale, nm.out_scale)
    out = {"params": os.path.relpath(CLOSURE_PARAMS, ROOT), "trained": meta, "closure_interval": interval}
    # held-out start: after the training snapshots
    t_start = int(meta["t_start"])
    t0 = time.perf_counter()
    start = snapshots(truth, [t_start])[0]
    marks = sorted(set(list(range(score_every, rollout_steps + 1, score_every)) + [rollout_steps]))
    tr = {}
    for k in marks:
        truth.advance_to(t_start + truth.per_step * k)
        tr[k] = truth.coarse_state()
    log(f"learned: fine truth to step {truth.step} in {time.perf_counter() - t0:.1f} s")
    runs = {}
    for name, mode, use in (("plain_fp64", "fp64", False), ("learned_fp64", "fp64", True),
                            ("plain_fp32", "fp32", False), ("learned_fp32", "fp32", True)):
        r = rollout(cb, start, rollout_steps, tr, net if use else None, rollout_ratio, mode, interval=interval)
        runs[name] = {"ssh_rel_l2": r["ssh"][-1], "u_rel_l2": r["u"][-1], "v_rel_l2": r["v"][-1],
                      "scored_steps": r["steps"], "ssh_by_step": r["ssh"], "u_by_step": r["u"],
                      "v_by_step": r["v"], "range_check_bad_cells": r["bad"]}
        if use:
            runs[name]["fallback_steps"] = r["fallback_steps"]
            runs[name]["active_fraction"] = 1.0 - r["fallback_steps"] / max(rollout_steps // interval, 1)
            runs[name]["last_norm_ratio"] = r["norms"][0] / max(r["norms"][1], 1e-300)
    gain = {}
    for m in ("fp64", "fp32"):
        p, q = runs[f"plain_{m}"], runs[f"learned_{m}"]
        gain[m] = {k: 1.0 - q[f"{k}_rel_l2"] / p[f"{k}_rel_l2"] for k in ("ssh", "u", "v")}
        gain[m]["ssh_mean_over_scored_steps"] = 1.0 - float(np.mean(q["ssh_by_step"]) / np.mean(p["ssh_by_step"]))
    tm = {}
    with truth.model as m:
        warm_up(lambda k: m.run(k, "fp64", nstep=NSTEP_BAROTROP), m.synchronize, 20, LEG_WARMUP_SECONDS)
        ms = m.run_timed(1000, "fp64", nstep=NSTEP_BAROTROP)
    fine_sim = fine.dt * 1000 / (ms * 1e-3)
    tm["fine_fp64"] = {"us_per_step": ms, "sim_s_per_s": fine_sim}
    for name, mode, use, iv in (("plain_fp32", "fp32", False, 1), ("learned_fp32", "fp32", True, interval),
                                ("learned_fp64", "fp64", True, interval),
                                ("learned_fp32_every_step", "fp32", True, 1),
                                ("learned_fp64_every_step", "fp64", True, 1)):
        r = rollout(cb, start, 1000, None, net if use else None, rollout_ratio, mode, interval=iv, timed=True)
        sim = cb.dt * 1000 / (r["ms"] * 1e-3)
        tm[name] = {"us_per_step": r["ms"], "sim_s_per_s": sim, "speedup_vs_fine_fp64": sim / fine_sim,
                    "wet_cell_steps_per_s": cb.n_wet * 1000 * NSTEP_BAROTROP / (r["ms"] * 1e-3),
                    "closure_interval": iv if use else None}
    out.update({"workload": f"{cb.nx - 4}x{cb.ny - 4} at 4 km (4x block average of the 1 km basin), "
                            f"dt=40 s, nstep_barotrop=2; {rollout_steps} coarse steps from fine step "
                            f"{t_start} scored every {score_every} steps against the coarsened FP64 1 km run; "
                            f"the closure after every {interval}-th step, under the {rollout_ratio:g}x rule",
                "error_vs_coarsened_truth": runs, "gain_vs_plain": gain, "throughput": tm})
    for k, v in runs.items():
        log(f"learned {k}: ssh {v['ssh_rel_l2']:.5f} u {v['u_rel_l2']:.5f} v {v['v_rel_l2']:.5f}")
    log(f"learned fp32: {tm['learned_fp32']['us_per_step']:.1f} us/step, "
        f"{tm['learned_fp32']['speedup_vs_fine_fp64']:.1f}x the 1 km FP64 run in simulated time; "
        f"fp64 {tm['learned_fp64']['speedup_vs_fine_fp64']:.1f}x")
    truth.close()
    if members > 0:
        mem = coarse_ensemble(members)
        ens_out = {}
        with Ensemble(mem, mode="fp64") as ref:
            ref.run(member_steps, NSTEP_BAROTROP)
            ref_ssh = [ref.get_state(k)["ssh"] for k in range(len(mem))]
        # the fallback rule as stated (ratio x the plain viscous tendency) at the
        # trained strength and scaled down, then the closure applied at every step
        cases = [(f"fallback_rule_x{sc:g}" if sc != 1.0 else "fallback_rule", ratio, sc)
                 for sc in CLOSURE_ENSEMBLE_SCALES] + [("always_on", 1e300, 1.0)]
        for label, rt, sc in cases:
            net.set_norm(nm.shift, nm.scale, (nm.out_scale * np.float32(sc)).astype(np.float32))
            with Ensemble(mem, mode="fp32") as ens:
                nmem, wet, cells = ens.size()
                ens.set_closure(net, rt)
                ens.run(5, NSTEP_BAROTROP)
                ens.reset()
                ms = ens.run_timed(member_steps, NSTEP_BAROTROP)
                nf, nr = ens.closure_stats()
                errs = [rel(ens.get_state(k)["ssh"], ref_ssh[k], b.wet) for k, b in enumerate(mem)]
                bad = ens.range_check()
            total = nmem * (member_steps + 5)
            ens_out[label] = {
                "ratio": rt, "strength_scale": sc, "ms_per_step": ms / member_steps,
                "value": wet * member_steps * NSTEP_BAROTROP / (ms * 1e-3), "unit": "wet-cell steps/s",
                "ssh_rel_l2_vs_plain_fp64": {"median": float(np.median(errs)), "max": float(np.max(errs))},
                "fallback": {"members": int((nf > 0).sum()), "member_steps": int(nf.sum())},
                "active_fraction": 1.0 - float(nf.sum()) / total,
                "members_always_active": int((nf == 0).sum()),
                "last_step_norm_ratio_median": float(np.median(nr[:, 0] / np.maximum(nr[:, 1], 1e-300))),
                "range_check_bad_cells": bad}
            log(f"closure ensemble ({label}): {nmem} members, {ms / member_steps:.2f} ms/step, "
                f"{ens_out[label]['value']:.3e} wet-cell steps/s, active on "
                f"{100 * ens_out[label]['active_fraction']:.1f} % of member-steps")
        net.set_norm(nm.shift, nm.scale, nm.out_scale)
        out["ensemble_fp32_closure"] = {"members": nmem, "wet_cells": wet, "array_cells": cells,
                                        "steps": member_steps, **ens_out}
    net.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000, help="timed outer steps (10 s each)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="cfg1", choices=["cfg0", "cfg1"])
    ap.add_argument("--cpu-steps", type=int, default=150,
                    help="outer steps timed on the CPU oracle (about 15 s on 16 threads)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--modes", default="fp64,fp32,fp64_exact")
    ap.add_argument("--ensemble-steps", type=int, default=5000,
                    help="outer steps of the 64-member FP32 ensemble (0 = skip)")
    ap.add_argument("--ensemble-members", type=int, default=64)
    ap.add_argument("--learned-steps", type=int, default=10000,
                    help="coarse 40 s steps of the learned-closure rollout (0 = skip)")
    ap.add_argument("--closure-members", type=int, default=128)
    ap.add_argument("--no-graphs", action="store_true",
                    help="ensembles launch every step directly instead of replaying HIP graphs (for profilers: "
                         "rocprofv3's kernel trace of the closure ensembles' graph replays ends the process)")
    ap.add_argument("--cnn-train-steps", type=int, default=20,
                    help="timed training steps of the closure network (batch 32; 0 = skip)")
    ap.add_argument("--closure-member-steps", type=int, default=1000)
    ap.add_argument("--coupling-levels", type=int, default=COUPLING_NZ,
                    help="levels of the 2-D -> 3-D coupling leg (0: skip it)")
    ap.add_argument("--coupling-reps", type=int, default=50)
    ap.add_argument("--forcing-levels", type=int, default=COUPLING_NZ,
                    help="levels of the forcing leg (3-D model -> barotropic RHS; 0: skip it)")
    ap.add_argument("--forcing-reps", type=int, default=20)
    ap.add_argument("--baroclinic-levels", type=int, default=COUPLING_NZ,
                    help="levels of the baroclinic_dynamics leg (0: skip it)")
    ap.add_argument("--baroclinic-reps", type=int, default=20)
    ap.add_argument("--reference-steps", type=int, default=200,
                    help="outer steps of the reference Fortran timing at 289x163 (0: skip)")
    ap.add_argument("--snapshots", type=int, default=3,
                    help="snapshots of the state written in the reference's output formats (0: skip)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH", "WRITE"),
                    help="write the per-launch HBM traffic of the fused kernels from two rocprofv3 "
                         "--pmc outputs (counter_collection.csv or results.db; FETCH_SIZE, WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.no_graphs:
        from inmsom_bt.ensemble import Ensemble
        Ensemble.USE_GRAPHS = False
    if args.traffic_from:
        print(json.dumps(write_traffic(*args.traffic_from), indent=1))
        return

    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(the step is not sharded across GPUs).", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    b = single_basin(args.workload)
    ncell = b.nx * b.ny
    log(f"workload {args.workload}: {b.nx - 4}x{b.ny - 4} interior, {b.n_wet} wet cells, "
        f"generated in {time.perf_counter() - t0:.1f} s")
    modes = [x for x in args.modes.split(",") if x]
    results, states = {}, {}
    with BarotropicModel(b, device=0) as m:
        for mode in modes:
            results[mode], states[mode] = timed_mode(m, b, mode, args.steps, args.warmup, ncell)
        coupling = None
        if args.coupling_levels > 0:
            coupling, coupling_inp = coupling_leg(m, b, args.coupling_levels, args.coupling_reps)
        forcing = None
        if args.forcing_levels > 0:
            forcing, forcing_inp = forcing_leg(m, b, args.forcing_levels, args.forcing_reps)
        baroclinic = None
        if args.baroclinic_levels > 0:
            baroclinic, baroclinic_inp = baroclinic_leg(m, b, args.baroclinic_levels, args.baroclinic_reps)
        vertical = None
        if args.baroclinic_levels > 0:
            vertical_inp, vertical = vertical_leg(m, b, args.baroclinic_levels, args.baroclinic_reps)
        pressure = density = smag = None
        if args.baroclinic_levels > 0:
            pressure_inp, pressure = pressure_leg(m, b, args.baroclinic_levels, args.baroclinic_reps)
            density_inp, density = density_leg(m, b, args.baroclinic_levels, args.baroclinic_reps)
            smag_inp, smag = smagorinsky_leg(m, b, args.baroclinic_levels, args.baroclinic_reps)
        snapshot = snapshot_leg(m, b, args.snapshots) if args.snapshots > 0 else None
        dropin = dropin_leg(m, b) if "fp64" in modes else None
        dropin_dev = dropin_device_leg(m, b) if "fp64" in modes else None

    errs = {}
    if "fp64" in states:
        ref = states["fp64"]
        for mode in ("fp32", "fp64_exact"):
            if mode in states:
                errs[mode] = {k: {"rel_l2": rel_l2(states[mode][k], ref[k]),
                                  "max_abs": float(np.max(np.abs(states[mode][k] - ref[k])))}
                              for k in ("ssh", "u", "v")}
                results[mode]["err_vs_fp64"] = errs[mode]

    head = results.get("fp64") or results[modes[0]]
    head_mode = "fp64" if "fp64" in results else modes[0]
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "wet-cell steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp64_exact": "f64", "fp32": "f32+f64"}[head_mode],
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {b.nx - 4}x{b.ny - 4} closed basin, dx=1 km, "
                               "dt=10 s outer step, nstep_barotrop=2 (5 s leap-frog substeps), "
                               f"{head_mode} fused step",
                   "wet_cells": b.n_wet, "array_cells": ncell,
                   "step": "one 10 s outer step = 2 state-advancing leap-frog substeps; value "
                           "counts substeps x wet cells"},
        "roofline": head["roofline"],
    }
    for mode in modes:
        if mode != head_mode:
            line[mode] = results[mode]
    line["fp64_details"] = {k: head[k] for k in ("avg_launch_us", "range_check_bad_cells",
                                                 "volume_rel_drift", "wall_s")}
    if coupling is not None:
        if not args.no_cpu_baseline:
            coupling["cpu_baseline"] = coupling_cpu(b, coupling_inp)
        del coupling_inp
        line["coupling"] = coupling
    if forcing is not None:
        if not args.no_cpu_baseline:
            forcing["cpu_baseline"] = forcing_cpu(b, forcing_inp)
        del forcing_inp
        line["forcing"] = forcing
    if baroclinic is not None:
        if not args.no_cpu_baseline:
            baroclinic["cpu_baseline"] = baroclinic_cpu(b, baroclinic_inp)
        del baroclinic_inp
        line["baroclinic"] = baroclinic
    if vertical is not None:
        if not args.no_cpu_baseline:
            vertical["cpu_baseline"] = vertical_cpu(b, vertical_inp)
        del vertical_inp
        line["vertical_velocity"] = vertical
    if pressure is not None:
        if not args.no_cpu_baseline:
            pressure["cpu_baseline"] = pressure_cpu(b, pressure_inp)
        del pressure_inp
        line["pressure_gradients"] = pressure
    if density is not None:
        if not args.no_cpu_baseline:
            density["cpu_baseline"] = density_cpu(b, density_inp)
        del density_inp
        line["density"] = density
    if smag is not None:
        if not args.no_cpu_baseline:
            smag["cpu_baseline"] = smagorinsky_cpu(b, smag_inp)
        del smag_inp
        line["smagorinsky"] = smag
    if dropin is not None:
        line["dropin_host_arrays"] = dropin
    if dropin_dev is not None:
        line["dropin_device"] = dropin_dev
    if snapshot is not None:
        line["snapshot"] = snapshot
    if args.ensemble_steps > 0:
        line["ensemble_fp32"] = ensemble_leg(args.ensemble_steps, min(args.warmup, 20),
                                             args.ensemble_members)
    if args.learned_steps > 0 and args.workload == "cfg1":
        line["learned"] = learned_leg(b, args.learned_steps, args.closure_members, args.closure_member_steps)
    if args.cnn_train_steps > 0 and os.path.exists(CLOSURE_PARAMS):
        line["cnn_training"] = cnn_train_leg(args.cnn_train_steps)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline(b, args.cpu_steps)
        log(f"cpu baseline: {line['cpu_baseline']['value']:.3e} wet-cell steps/s on "
            f"{line['cpu_baseline']['cores']} threads")
        ref = cpu_reference_fortran(args.reference_steps)
        if ref is not None:
            line["cpu_reference_fortran"] = ref
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
