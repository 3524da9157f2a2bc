"""The measured workloads themselves, at their full sizes, on the GPU.

* 256 x 256 closed basin, 1000 outer steps of 10 s (nstep_barotrop = 2):
  the stage-by-stage path and the fused FP64_EXACT step equal the oracle
  bit for bit; the fused FP64 step is within relative L2 1e-12 of it and
  the FP32 step within 1e-5.
* 2048 x 1536 closed basin: FP64_EXACT equals the oracle bit for bit after
  20 outer steps; after 10^4 outer steps FP64 is within 1e-12 of
  FP64_EXACT and FP32 within 1e-5 of FP64 (maximum errors printed).
* 64 uneven basins (256..1024 x 256..768) advanced together in FP32 for
  5000 outer steps: sampled members agree with their single-basin runs
  to relative L2 1e-5 (the ensemble's kernel and the single-basin kernel
  may contract multiply-adds differently), no cell fails the range check,
  volumes drift by at most 1e-5.
* 128 coarse 4 km basins (128..512 x 128..384) with the committed closure
  network under the 3x fallback rule, 1000 outer steps of 40 s: members
  that fell back at every step equal the plain FP32 run bit for bit, every
  member stays within 1e-2 (SSH relative L2) of the plain FP64 run, and
  the fallback counts lie in 0..steps.
"""
import json
import os

import numpy as np
import pytest

import oracle
from _util import STATE, assert_bitwise, rel_l2
from inmsom_bt import BarotropicModel
from inmsom_bt.ensemble import Ensemble
from inmsom_bt.workloads import NSTEP_BAROTROP, coarse_ensemble, ensemble, single_basin

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "gpurun_out")


def _record(name, data):
    """Errors measured by a test, kept beside its printed line."""
    os.makedirs(OUT, exist_ok=True)
    with open(os.path.join(OUT, f"config_errors_{name}.json"), "w") as f:
        json.dump(data, f, indent=1)
    print(f"{name}: {json.dumps(data)}", flush=True)


def _errs(got, want, keys=("ssh", "u", "v")):
    return {k: {"rel_l2": rel_l2(got[k], want[k]), "max_abs": float(np.max(np.abs(got[k] - want[k])))}
            for k in keys}


def test_basin_256_1000_steps_against_oracle():
    b = single_basin("cfg0")
    steps = 1000
    print(f"256x256: {b.n_wet} wet cells, {steps} outer steps, nstep={NSTEP_BAROTROP}", flush=True)
    want = oracle.run(b, steps, NSTEP_BAROTROP, out_steps=[steps])[steps]
    res = {}
    with BarotropicModel(b) as m:
        m.model_steps(steps, nstep=NSTEP_BAROTROP)
        assert_bitwise(m.get_state(), want, STATE, "stage-by-stage path, 1000 steps")
        m.set_state(b.u, b.up, b.v, b.vp, b.ssh, b.sshp)
        m.run(steps, "fp64_exact", nstep=NSTEP_BAROTROP)
        assert_bitwise(m.get_state(), want, STATE, "fused FP64_EXACT, 1000 steps")
        for mode, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
            m.set_state(b.u, b.up, b.v, b.vp, b.ssh, b.sshp)
            m.run(steps, mode, nstep=NSTEP_BAROTROP)
            assert m.range_check() == 0
            res[mode] = _errs(m.get_state(), want)
    _record("basin256_1000", res)
    for mode, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
        for k, e in res[mode].items():
            assert e["rel_l2"] <= tol, (mode, k, e)


def test_basin_2048x1536_exact_20_steps_against_oracle():
    b = single_basin("cfg1")
    steps = 20
    print(f"2048x1536: {b.n_wet} wet cells, {steps} outer steps against the oracle", flush=True)
    want = oracle.run(b, steps, NSTEP_BAROTROP, out_steps=[steps])[steps]
    with BarotropicModel(b) as m:
        m.run(steps, "fp64_exact", nstep=NSTEP_BAROTROP)
        assert_bitwise(m.get_state(), want, STATE, "fused FP64_EXACT at 2048x1536")


def test_basin_2048x1536_10000_steps_modes():
    b = single_basin("cfg1")
    steps = 10000
    print(f"2048x1536: {steps} outer steps in FP64_EXACT, FP64 and FP32", flush=True)
    st = {}
    with BarotropicModel(b) as m:
        for mode in ("fp64_exact", "fp64", "fp32"):
            m.set_state(b.u, b.up, b.v, b.vp, b.ssh, b.sshp)
            v0 = m.volume()
            m.run(steps, mode, nstep=NSTEP_BAROTROP)
            assert m.range_check() == 0, mode
            st[mode] = m.get_state()
            st[mode]["volume_rel_drift"] = abs(m.volume() - v0) / abs(v0)
            print(f"{mode}: done, volume drift {st[mode]['volume_rel_drift']:.3e}", flush=True)
    res = {"fp64_vs_fp64_exact": _errs(st["fp64"], st["fp64_exact"]),
           "fp32_vs_fp64": _errs(st["fp32"], st["fp64"]),
           "volume_rel_drift": {k: st[k]["volume_rel_drift"] for k in st}}
    _record("basin2048x1536_10000", res)
    for k, e in res["fp64_vs_fp64_exact"].items():
        assert e["rel_l2"] <= 1e-12, ("fp64", k, e)
    for k, e in res["fp32_vs_fp64"].items():
        assert e["rel_l2"] <= 1e-5, ("fp32", k, e)


def test_ensemble_64_members_5000_steps():
    basins = ensemble()
    steps = 5000
    sample = sorted({0, int(np.argmin([b.nx * b.ny for b in basins])),
                     int(np.argmax([b.nx * b.ny for b in basins])), len(basins) // 2})
    print(f"64-member FP32 ensemble, {steps} outer steps, sampled members {sample}", flush=True)
    with Ensemble(basins, mode="fp32") as ens:
        v0 = [ens.volume(k) for k in range(len(basins))]
        ens.run(steps, nstep=NSTEP_BAROTROP)
        assert ens.range_check() == 0
        drift = [abs(ens.volume(k) - v0[k]) / abs(v0[k]) for k in range(len(basins))]
        got = {k: ens.get_state(k) for k in sample}
        stages = ens.pipeline()
    _record("ensemble64_5000", {"max_volume_rel_drift": max(drift), "sampled": sample, "pipeline": stages})
    assert max(drift) <= 1e-5, max(drift)
    for k in sample:
        b = basins[k]
        # the single basin runs the same kernel body (the two-wave pipeline
        # when the ensemble uses it); the two kernels may contract
        # multiply-adds differently, so the members agree to FP32 rounding
        # grown over 5000 steps, not bit for bit
        print(f"member {k}: {b.nx}x{b.ny} single-basin run, pipeline {stages}", flush=True)
        with BarotropicModel(b) as m:
            m.set_pipeline(stages if stages else 1)
            m.run(steps, "fp32", nstep=NSTEP_BAROTROP)
            want = m.get_state()
        for f in STATE:
            e = rel_l2(got[k][f], want[f])
            print(f"  {f}: rel L2 {e:.2e}")
            assert e <= 1e-5, (k, f, e)


CLOSURE_PARAMS = os.path.join(ROOT, "inmsom-elfs_amd", "inmsom_bt", "data", "closure_cfg1.npz")


def test_coarse_ensemble_128_members_with_closure():
    from inmsom_bt.closure import ClosureNet
    from inmsom_bt.closure_experiment import load_params
    from inmsom_bt.learned import rel

    members = coarse_ensemble()
    steps = 1000
    params, nm, meta = load_params(CLOSURE_PARAMS)
    print(f"128 coarse members with the committed closure, {steps} outer steps", flush=True)
    with Ensemble(members, mode="fp64") as ref:
        ref.run(steps, NSTEP_BAROTROP)
        ref64 = [ref.get_state(k) for k in range(len(members))]
    with Ensemble(members, mode="fp32") as plain:
        plain.run(steps, NSTEP_BAROTROP)
        ref32 = [plain.get_state(k) for k in range(len(members))]
    with ClosureNet(1, 8, 32) as net:
        net.set_output(meta.get("output", "uv"))
        net.set_params(params)
        net.set_norm(nm.shift, nm.scale, nm.out_scale)
        with Ensemble(members, mode="fp32") as ens:
            ens.set_closure(net, 3.0)
            ens.run(steps, NSTEP_BAROTROP)
            nfall, norms = ens.closure_stats()
            assert ens.range_check() == 0
            got = [ens.get_state(k) for k in range(len(members))]
            ens.set_closure(None)
    errs = [rel(got[k]["ssh"], ref64[k]["ssh"], b.wet) for k, b in enumerate(members)]
    always = [k for k in range(len(members)) if nfall[k] == steps]
    res = {"steps": steps, "members": len(members),
           "active_member_steps_frac": float(1.0 - nfall.sum() / (steps * len(members))),
           "members_never_falling_back": int((nfall == 0).sum()),
           "members_always_falling_back": len(always),
           "ssh_rel_l2_vs_plain_fp64": {"median": float(np.median(errs)), "max": float(np.max(errs))}}
    _record("coarse_ensemble128_closure", res)
    assert np.all((nfall >= 0) & (nfall <= steps))
    for k in always:
        assert_bitwise(got[k], ref32[k], STATE, f"member {k} (fell back at every step) vs plain FP32")
    assert max(errs) <= 1e-2, max(errs)
