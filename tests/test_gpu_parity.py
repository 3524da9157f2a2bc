"""GPU parity of the HIP path against the oracle (C restatement pinned to
the compiled reference).  Integer-exact comparisons (bit patterns) for the
reference-order paths, relative L2 for the reassociated FP64 and the FP32
paths, with the tolerances written in each test."""
import numpy as np
import pytest

import oracle
from _util import STATE, assert_bitwise, rel_l2
from inmsom_bt import BarotropicModel, make_basin, randomize_state

pytestmark = pytest.mark.gpu

ACC = ("ubrtr_i", "vbrtr_i", "ssh4gradx", "ssh4grady")
DEPTHS = ("hhq_e", "hhqp_e", "hhqn_e", "hhu_e", "hhup_e", "hhun_e",
          "hhv_e", "hhvp_e", "hhvn_e", "hhh_e", "hhhp_e", "hhhn_e")


def _case(kind):
    if kind == "uniform":
        return make_basin(150, 110, seed=21, sigma=6.0), 1, 0
    if kind == "stretched_ksw4_nstep3":
        b = make_basin(130, 97, seed=22, sigma=5.0, metrics="stretched", lvisc_4=1e-3)
        return randomize_state(b, 7, wflux_amp=1e-5), 3, 1
    if kind == "ragged_small":
        return randomize_state(make_basin(7, 5, seed=23, sigma=1.0, hland=60.0), 8), 2, 0
    raise KeyError(kind)


@pytest.mark.parametrize("kind", ["uniform", "stretched_ksw4_nstep3", "ragged_small"])
def test_model_steps_bitwise(kind):
    b, nstep, ksw4 = _case(kind)
    steps = 12
    print(f"case {kind}: {b.nx}x{b.ny} nstep={nstep} ksw4={ksw4}", flush=True)
    want = oracle.run(b, steps, nstep, ksw4, out_steps=[steps])[steps]
    with BarotropicModel(b) as m:
        m.model_steps(steps, nstep=nstep, ksw4=ksw4)
        got = m.get_state()
        got.update(m.means())
        got.update(m.depths())
    assert_bitwise(got, want, STATE + ACC + DEPTHS, f"model_steps[{kind}]")


def test_barotropic_dynamics_dropin_bitwise():
    """The drop-in entry point with caller-supplied RHS, wflux, mu, mu4."""
    b, nstep, ksw4 = _case("stretched_ksw4_nstep3")
    og = oracle.OracleGrid(b)
    rdis = og.rayleigh()
    rng = np.random.default_rng(5)
    rhsx = rng.standard_normal((b.ny, b.nx)) * 1e3
    rhsy = rng.standard_normal((b.ny, b.nx)) * 1e3
    args_o = {k: np.ascontiguousarray(getattr(b, k)).copy() for k in ("u", "up", "v", "vp", "ssh", "sshp")}
    args_g = {k: v.copy() for k, v in args_o.items()}
    acc_o = {k: np.zeros((b.ny, b.nx)) for k in ACC}
    acc_g = {k: np.zeros((b.ny, b.nx)) for k in ACC}
    want = oracle.barotropic_dynamics(og, b.dt, nstep, ksw4, args_o, acc_o, rhsx, rhsy,
                                      b.wflux, b.mu, b.mu4, rdis)
    with BarotropicModel(b, rdis=rdis) as m:
        m.barotropic_dynamics(b.dt, nstep, ksw4, args_g["u"], args_g["up"], args_g["v"],
                              args_g["vp"], args_g["ssh"], args_g["sshp"], acc_g["ubrtr_i"],
                              acc_g["vbrtr_i"], acc_g["ssh4gradx"], acc_g["ssh4grady"], rhsx, rhsy,
                              np.ascontiguousarray(b.wflux), np.ascontiguousarray(b.mu),
                              np.ascontiguousarray(b.mu4), rdis)
        dep = m.depths()
    got = dict(args_g)
    got.update(acc_g)
    got.update(dep)
    want_all = dict(args_o)
    want_all.update(acc_o)
    want_all.update(want)
    assert_bitwise(got, want_all, STATE + ACC + DEPTHS, "barotropic_dynamics")


@pytest.mark.parametrize("shape,nstep", [((150, 110), 1), ((150, 110), 2), ((61, 7), 3),
                                         ((300, 64), 2)])
def test_fused_exact_bitwise(shape, nstep):
    b = make_basin(*shape, seed=31, sigma=5.0)
    steps = 20
    print(f"fused exact {shape} nstep={nstep}", flush=True)
    want = oracle.run(b, steps, nstep, out_steps=[steps])[steps]
    with BarotropicModel(b) as m:
        m.run(steps, "fp64_exact", nstep=nstep)
        got = m.get_state()
    assert_bitwise(got, want, STATE, f"fused fp64_exact {shape} nstep={nstep}")


@pytest.mark.parametrize("mode", ["fp64_exact", "fp64", "fp32"])
@pytest.mark.parametrize("shape,nstep,chunk", [((150, 110), 2, 0), ((150, 110), 3, 5), ((61, 7), 4, 0),
                                               ((300, 64), 2, 3), ((57, 40), 2, 1), ((116, 91), 4, 7)])
def test_two_substeps_per_launch_bitwise(mode, shape, nstep, chunk):
    """Two substeps in one march give what two launches give, bit for bit,
    in every mode, over strip edges, chunk edges and odd nstep."""
    b = randomize_state(make_basin(*shape, seed=33, sigma=5.0), 9)
    steps = 9
    print(f"two substeps/launch {mode} {shape} nstep={nstep} chunk={chunk}", flush=True)
    out = {}
    for k in (1, 2):
        with BarotropicModel(b) as m:
            m.set_launch(chunk)
            m.set_substeps_per_launch(k)
            m.run(steps, mode, nstep=nstep)
            out[k] = m.get_state()
    assert_bitwise(out[2], out[1], STATE, f"two vs one substep per launch, {mode} {shape} nstep={nstep}")


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
@pytest.mark.parametrize("stages", [2, 3, 4])
@pytest.mark.parametrize("shape,nstep,steps,chunk", [((150, 110), 2, 9, 0), ((150, 110), 3, 7, 5),
                                                     ((61, 7), 2, 5, 0), ((300, 64), 1, 6, 3),
                                                     ((116, 91), 4, 3, 7)])
def test_pipeline_matches_one_launch_per_substep(mode, stages, shape, nstep, steps, chunk):
    """K substeps per launch as a K-wave pipeline (across outer steps, with
    one- and two-substep launches for the remainder) give what one launch
    per substep gives, over strip and chunk edges, odd nstep and a ragged
    last chunk: to rounding (FP64 relative L2 <= 1e-14, FP32 <= 1e-6; the
    modes contract multiply-adds as the compiler finds them, which the
    pipeline's code may do differently)."""
    b = randomize_state(make_basin(*shape, seed=34, sigma=5.0), 10)
    print(f"pipeline {stages} {mode} {shape} nstep={nstep} steps={steps} chunk={chunk}", flush=True)
    out = {}
    for k in (1, stages):
        with BarotropicModel(b) as m:
            m.set_launch(chunk)
            m.set_substeps_per_launch(1)
            m.set_pipeline(k)
            m.run(steps, mode, nstep=nstep)
            out[k] = m.get_state()
    tol = 1e-14 if mode == "fp64" else 1e-6
    for k in STATE:
        e = rel_l2(out[stages][k], out[1][k])
        print(f"  {k}: rel L2 {e:.2e}")
        assert e <= tol, (k, e)


def test_fused_fp64_within_1e12():
    """Reassociated FP64 step: relative L2 <= 1e-12 after 200 outer steps."""
    b = make_basin(200, 150, seed=41, sigma=8.0)
    steps, nstep = 200, 2
    want = oracle.run(b, steps, nstep, out_steps=[steps])[steps]
    with BarotropicModel(b) as m:
        m.run(steps, "fp64", nstep=nstep)
        got = m.get_state()
    for k in ("ssh", "u", "v", "sshp", "up", "vp"):
        e = rel_l2(got[k], want[k])
        print(f"fp64 {k}: rel L2 {e:.3e}")
        assert e <= 1e-12, (k, e)


def test_fused_fp32_tolerance():
    """FP32 stencils, FP64 filtered level: rel L2 <= 1e-5 after 200 steps."""
    b = make_basin(200, 150, seed=42, sigma=8.0)
    steps, nstep = 200, 2
    want = oracle.run(b, steps, nstep, out_steps=[steps])[steps]
    with BarotropicModel(b) as m:
        m.run(steps, "fp32", nstep=nstep)
        got = m.get_state()
    for k in ("ssh", "u", "v"):
        e = rel_l2(got[k], want[k])
        print(f"fp32 {k}: rel L2 {e:.3e}")
        assert e <= 1e-5, (k, e)


# ---------------------------------------------------------------- golden vectors
import hashlib  # noqa: E402
import json  # noqa: E402
import os  # noqa: E402
import sys  # noqa: E402

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(_HERE, "golden"))
import make_golden  # noqa: E402

with open(os.path.join(_HERE, "golden", "reference_vectors.json")) as _f:
    _GOLD = json.load(_f)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("name", [k for k, v in sorted(_GOLD["cases"].items())
                                  if v["run"].get("mode", 0) == 0])
@pytest.mark.parametrize("path", ["model_steps", "fused_exact"])
def test_gpu_against_reference_vectors(name, path):
    """The HIP paths against the vectors the compiled reference wrote."""
    rec = _GOLD["cases"][name]
    run = rec["run"]
    b = make_golden.build_case(name)
    if path == "fused_exact" and (run["ksw4"] or rec["basin"].get("metrics") == "stretched"):
        pytest.skip("the fused step covers uniform metrics without the biharmonic term")
    print(f"golden {name} via {path}", flush=True)
    with BarotropicModel(b) as m:
        last = 0
        for step in sorted(int(s) for s in rec["outputs"]):
            if path == "model_steps":
                m.model_steps(step - last, nstep=run["nstep"], ksw4=run["ksw4"])
            else:
                m.run(step - last, "fp64_exact", nstep=run["nstep"])
            last = step
            got = m.get_state()
            names = {"ssh": "ssh", "sshp": "sshp", "u": "u", "up": "up", "v": "v", "vp": "vp"}
            if path == "model_steps":
                got.update(m.means())
                got.update(m.depths())
            bad = [k for k, d in rec["outputs"][str(step)].items()
                   if k in got and _sha(got[k]) != d["sha256"]]
            assert not bad, f"{name} step {step} via {path}: {bad}"


@pytest.mark.parametrize("mode", ["fp64_exact", "fp32"])
def test_save_restore_state_reruns_bitwise(mode):
    """inmsom_bt_save_state / _restore_state: a run from the restored state
    equals, bit for bit, the same run from the state uploaded with
    set_state (and the exact path still equals the oracle), also when the
    state was on the fused path when saved; restoring before any save is
    refused."""
    b = make_basin(120, 90, seed=43, sigma=6.0)
    steps, nstep = 6, 2
    with BarotropicModel(b) as m:
        with pytest.raises(Exception):
            m.restore_state()
        m.run(3, mode, nstep=nstep)              # the saved state lives on the fused path
        m.save_state()
        mid = m.get_state()
        m.run(steps, mode, nstep=nstep)
        first = m.get_state()
        m.restore_state()
        assert_bitwise(m.get_state(), mid, STATE, f"{mode} restored state")
        m.run(steps, mode, nstep=nstep)
        assert_bitwise(m.get_state(), first, STATE, f"{mode} rerun from the restored state")
        m.set_state(mid["u"], mid["up"], mid["v"], mid["vp"], mid["ssh"], mid["sshp"])
        m.run(steps, mode, nstep=nstep)
        assert_bitwise(m.get_state(), first, STATE, f"{mode} rerun from set_state")
    if mode == "fp64_exact":
        want = oracle.run(b, 3 + steps, nstep, out_steps=[3 + steps])[3 + steps]
        assert_bitwise(first, want, STATE, "exact path against the oracle")
