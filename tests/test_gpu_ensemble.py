"""Batched ensembles on the GPU: every member equals the single-basin
fused step bit for bit (same arithmetic, different work mapping), the
FP64_EXACT members equal the oracle, graph replays equal direct launches,
the launch shape never changes results, and the autotuner restores the
state it times from."""
import numpy as np
import pytest

import oracle
from _util import STATE, assert_bitwise, rel_l2
from inmsom_bt import BarotropicModel
from inmsom_bt.ensemble import Ensemble
from inmsom_bt.workloads import ensemble

pytestmark = pytest.mark.gpu


def _members(n=5, seed=3):
    return ensemble(n, seed=seed, size_lo=(40, 30), size_hi=(150, 120), sigma=5.0)


@pytest.mark.parametrize("mode", ["fp64_exact", "fp64", "fp32"])
def test_members_equal_single_basin(mode):
    basins = _members()
    steps = 8
    print(f"ensemble {mode}: " + ", ".join(f"{b.nx}x{b.ny}" for b in basins), flush=True)
    with Ensemble(basins, mode=mode) as ens:
        ens.run(steps, nstep=2)
        got = [ens.get_state(k) for k in range(len(basins))]
    for k, b in enumerate(basins):
        with BarotropicModel(b) as m:
            m.run(steps, mode, nstep=2)
            want = m.get_state()
        assert_bitwise(got[k], want, STATE, f"member {k} ({b.nx}x{b.ny}) {mode}")
    if mode == "fp64_exact":
        ref = oracle.run(basins[1], steps, 2, out_steps=[steps])[steps]
        assert_bitwise(got[1], ref, STATE, "member 1 vs oracle")


def test_graph_replays_match_direct_launches():
    basins = _members(4, seed=5)
    with Ensemble(basins, mode="fp32") as g, Ensemble(basins, mode="fp32") as d:
        for rep in range(3):
            g.run(1, nstep=2, graph=True)
            d.run(1, nstep=2, graph=False)
            # the program's own eager work between replays
            assert g.range_check() == 0 and d.range_check() == 0
            g.volume(0)
            for k in range(len(basins)):
                assert_bitwise(g.get_state(k), d.get_state(k), STATE, f"replay {rep} member {k}")


def test_launch_shape_does_not_change_results():
    basins = _members(3, seed=9)
    outs = []
    for rows, waves in ((0, 0), (7, 37), (48, 512), (3, 130)):
        with Ensemble(basins, mode="fp64") as ens:
            ens.set_launch(rows, waves)
            ens.run(5, nstep=2)
            outs.append([ens.get_state(k) for k in range(len(basins))])
    for o in outs[1:]:
        for k in range(len(basins)):
            assert_bitwise(o[k], outs[0][k], STATE, f"member {k}")


@pytest.mark.parametrize("mode,nstep", [("fp32", 2), ("fp64", 3), ("fp64_exact", 2), ("fp32", 4)])
def test_two_substeps_per_launch_equal_one(mode, nstep):
    """Substep pairs in one persistent launch give the same members, bit for
    bit, as one launch per substep, with and without graphs."""
    basins = _members(4, seed=11)
    outs = []
    for k, graph in ((1, True), (2, True), (2, False)):
        with Ensemble(basins, mode=mode) as ens:
            ens.set_substeps_per_launch(k)
            ens.set_pipeline(1)
            ens.run(5, nstep=nstep, graph=graph)
            outs.append([ens.get_state(j) for j in range(len(basins))])
    for o in outs[1:]:
        for j in range(len(basins)):
            assert_bitwise(o[j], outs[0][j], STATE, f"member {j} {mode} nstep={nstep}")


@pytest.mark.parametrize("mode,nstep", [("fp32", 2), ("fp64", 3), ("fp64", 4), ("fp32", 5)])
def test_pipeline_pairs_match_one_launch_per_substep(mode, nstep):
    """Substep pairs as a two-wave pipeline agree with one launch per
    substep to rounding (FP64 relative L2 <= 1e-14, FP32 <= 1e-6: the
    modes contract multiply-adds as the compiler finds them), and graph
    replays of the pipeline equal its direct launches bit for bit."""
    basins = _members(4, seed=17)
    outs = []
    for k, pipe, graph in ((1, 1, True), (2, 2, True), (2, 2, False)):
        print(f"ensemble {mode} nstep={nstep} substeps/launch={k} pipeline={pipe} graph={graph}", flush=True)
        with Ensemble(basins, mode=mode) as ens:
            ens.set_substeps_per_launch(k)
            ens.set_pipeline(pipe)
            assert ens.pipeline() == (2 if pipe == 2 else 0)
            ens.run(6, nstep=nstep, graph=graph)
            outs.append([ens.get_state(j) for j in range(len(basins))])
    tol = 1e-14 if mode == "fp64" else 1e-6
    for j in range(len(basins)):
        assert_bitwise(outs[2][j], outs[1][j], STATE, f"member {j}: graph vs direct")
        for f in STATE:
            e = rel_l2(outs[1][j][f], outs[0][j][f])
            assert e <= tol, (j, f, e)


def test_autotune_restores_state_and_conserves_volume(tmp_path):
    basins = _members(4, seed=11)
    with Ensemble(basins, mode="fp32") as ens:
        v0 = [ens.volume(k) for k in range(len(basins))]
        res = ens.autotune(nstep=2, trial_steps=3, rows=(12, 24), waves=(128, 384),
                           cache=str(tmp_path / "tune.json"))
        assert (res["rows_per_wave"], res["waves_per_xcd"]) in {(12, 128), (12, 384), (24, 128),
                                                                (24, 384)}
        assert len(res["trials"]) == 4
        for k, b in enumerate(basins):  # state is the initial one again
            st = ens.get_state(k)
            assert np.array_equal(st["ssh"], b.ssh.astype(np.float32).astype(np.float64))
        ens.run(30, nstep=2)
        assert ens.range_check() == 0
        for k in range(len(basins)):
            assert abs(ens.volume(k) - v0[k]) <= 1e-5 * abs(v0[k])
        again = ens.autotune(nstep=2, trial_steps=3, rows=(12, 24), waves=(128, 384),
                             cache=str(tmp_path / "tune.json"))
        assert again["cached"]


def test_graphs_follow_table_rebuilds():
    """Graph run at launch shape A, a direct run at shape B (which rebuilds
    the member tables and work list the graphs point at), then a graph run
    at A again: every stage equals direct launches bit for bit (graphs
    captured under replaced tables are captured again, not replayed)."""
    basins = _members(4, seed=13)
    with Ensemble(basins, mode="fp64") as g, Ensemble(basins, mode="fp64") as d:
        for stage, (rows, waves, graph) in enumerate(((0, 0, True), (5, 61, False), (0, 0, True),
                                                      (5, 61, True), (0, 0, True))):
            print(f"graph rebuild stage {stage}: rows {rows} waves {waves} graph {graph}", flush=True)
            g.set_launch(rows, waves)
            d.set_launch(rows, waves)
            g.run(GRAPH_STEPS + 1, nstep=2, graph=graph)
            d.run(GRAPH_STEPS + 1, nstep=2, graph=False)
            for k in range(len(basins)):
                assert_bitwise(g.get_state(k), d.get_state(k), STATE, f"stage {stage} member {k}")


GRAPH_STEPS = 8  # outer steps per replay of the ensemble's long graph
