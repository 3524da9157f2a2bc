"""The learned-closure workflow's device pieces: coarsening of the fine
state, the one-step targets, and the snapshot pipeline end to end on a
small basin."""
import numpy as np
import pytest

from inmsom_bt import BarotropicModel
from inmsom_bt.cases import make_basin, randomize_state
from inmsom_bt.learned import FineTruth, coarse_steps, mask_coarse, samples, snapshots, uv_masks

pytestmark = pytest.mark.gpu


def _coarsen_np(st, b, f):
    """Block averages restated in numpy: ssh over wet cells, u on the
    block's east faces, v on its north faces."""
    nyi, nxi = b.ny - 4, b.nx - 4
    cy, cx = nyi // f, nxi // f
    wet = b.wet[2:-2, 2:-2].astype(np.float64)
    s = st["ssh"][2:-2, 2:-2]
    num = (s * wet).reshape(cy, f, cx, f).sum(axis=(1, 3))
    den = wet.reshape(cy, f, cx, f).sum(axis=(1, 3))
    out = {k: np.zeros((cy + 4, cx + 4)) for k in ("u", "v", "ssh")}
    out["ssh"][2:-2, 2:-2] = np.where(den > 0, num / np.maximum(den, 1), 0.0)
    u = st["u"][2:-2, 2:-2]
    v = st["v"][2:-2, 2:-2]
    out["u"][2:-2, 2:-2] = u[:, f - 1::f].reshape(cy, f, cx).mean(axis=1)
    out["v"][2:-2, 2:-2] = v[f - 1::f, :].reshape(cy, cx, f).mean(axis=2)
    return out


@pytest.mark.parametrize("mode", ["fp64", "fp32", "literal"])
def test_coarsen_matches_block_average(mode):
    b = randomize_state(make_basin(96, 64, seed=4, sigma=5.0), 9)
    print(f"coarsen {mode}", flush=True)
    with BarotropicModel(b) as m:
        if mode == "literal":
            m.model_steps(3, nstep=2)
        else:
            m.run(3, mode, nstep=2)
        got = m.coarsen(4)
        st = m.get_state()
    want = _coarsen_np(st, b, 4)
    for k in ("u", "v", "ssh"):
        tol = 1e-14 if mode != "fp32" else 1e-6
        d = np.abs(got[k] - want[k]).max()
        assert d <= tol * max(np.abs(want[k]).max(), 1e-30), (k, d)


def test_coarse_steps_equal_single_basin_steps():
    """The batched coarse steps behind the targets equal one FP64 fused
    step of the coarse basin from the same levels, bit for bit."""
    fine = make_basin(128, 96, seed=5, sigma=8.0)
    truth = FineTruth(fine)
    try:
        snaps = snapshots(truth, [4, 12, 30])
    finally:
        truth.close()
    cb = truth.coarse
    plain = coarse_steps(cb, snaps, width=2)
    for s, p in zip(snaps, plain):
        with BarotropicModel(cb) as m:
            m.set_state(s.now["u"], s.prev["u"], s.now["v"], s.prev["v"], s.now["ssh"], s.prev["ssh"])
            m.run(1, "fp64", nstep=2)
            want = m.get_state()
        for k in want:
            assert np.array_equal(p[k], want[k]), k
    raw, tgt, mask = samples(cb, snaps, plain)
    lcu, lcv = uv_masks(cb)
    for k, (s, p) in enumerate(zip(snaps, plain)):
        np.testing.assert_array_equal(tgt[k, ..., 0], ((s.nxt["u"] - p["u"]) / cb.dt * lcu).astype(np.float32))
    assert np.array_equal(mask, lcu.astype(np.uint8) | (lcv.astype(np.uint8) << 1))


def test_snapshots_are_consistent():
    """A snapshot's levels are the coarsened fine states 2, 0 and -4 fine
    steps before t (coarse substep 20 s, coarse step 40 s, fine step 10 s)."""
    fine = make_basin(128, 96, seed=6, sigma=8.0)
    truth = FineTruth(fine)
    assert (truth.per_sub, truth.per_step) == (2, 4)
    try:
        snaps = snapshots(truth, [4, 20])
        with BarotropicModel(fine) as m:
            m.run(20, "fp64", nstep=2)
            direct = mask_coarse(m.coarsen(4), truth.coarse)
    finally:
        truth.close()
    for k in ("u", "v", "ssh"):
        assert np.array_equal(snaps[1].now[k], direct[k])
    lcu, lcv = uv_masks(truth.coarse)
    assert not snaps[0].now["u"][~lcu].any() and not snaps[0].now["v"][~lcv].any()
