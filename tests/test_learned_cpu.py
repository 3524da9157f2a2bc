"""Host-side pieces of the learned-closure workflow (no GPU): network
input features, target assembly, normalisation, parameter files."""
import json

import numpy as np

from inmsom_bt.cases import coarsen_basin, make_basin
from inmsom_bt.closure import features
from inmsom_bt.closure_experiment import init_params, load_params, save_params
from inmsom_bt.learned import Norms, Snapshot, fit_norms, samples, uv_masks
import cnn_oracle as co


def test_features_are_float32_affine():
    rng = np.random.default_rng(0)
    u, v, s, h = (rng.standard_normal((5, 7)) for _ in range(4))
    x = features(u, v, s, h, [0.1, 0.2, 0.3, 100.0], [2.0, 3.0, 4.0, 0.01])
    assert x.shape == (5, 7, 8) and x.dtype == np.float32
    np.testing.assert_array_equal(x[..., 0], (u.astype(np.float32) - np.float32(0.1)) * np.float32(2.0))
    np.testing.assert_array_equal(x[..., 3], (h.astype(np.float32) - np.float32(100.0)) * np.float32(0.01))
    assert not x[..., 4:].any()


def test_samples_targets_are_masked_residual_tendencies():
    fine = make_basin(64, 48, seed=1, sigma=4.0)
    cb = coarsen_basin(fine, 4)
    lcu, lcv = uv_masks(cb)
    rng = np.random.default_rng(2)

    def st():
        return {"u": rng.standard_normal((cb.ny, cb.nx)) * lcu, "v": rng.standard_normal((cb.ny, cb.nx)) * lcv,
                "ssh": rng.standard_normal((cb.ny, cb.nx)) * cb.wet}
    snaps = [Snapshot(10, st(), st(), st()) for _ in range(3)]
    plain = [dict(st(), up=0, vp=0) for _ in range(3)]
    raw, tgt, mask = samples(cb, snaps, plain)
    assert raw.shape == (3, 4, cb.ny, cb.nx) and tgt.shape == (3, cb.ny, cb.nx, 2)
    for k in range(3):
        np.testing.assert_allclose(tgt[k, ..., 0], (snaps[k].nxt["u"] - plain[k]["u"]) / cb.dt * lcu, rtol=1e-6,
                                   atol=1e-12)
        np.testing.assert_array_equal(raw[k, 3], cb.hhq_rest)
    assert np.array_equal(mask & 1, lcu.astype(np.uint8)) and np.array_equal(mask >> 1, lcv.astype(np.uint8))
    nm = fit_norms(cb, raw, tgt)
    # normalised targets have unit RMS on their points
    assert abs(np.sqrt(((tgt[..., 0] / nm.out_scale[0])[:, lcu] ** 2).mean()) - 1.0) < 1e-5
    assert nm.shift[3] > 0 and np.all(nm.scale > 0)


def test_parameter_file_round_trip(tmp_path):
    p = init_params(3, co.num_params(), out_scale=0.1)
    # same layout and distribution family as the oracle's initialiser
    assert p.size == co.num_params()
    for l, (woff, boff, kp) in enumerate(co.layout()):
        w = p[woff:woff + 32 * kp].reshape(32, kp)
        assert not w[:, 9 * co.CIN[l]:].any() and not p[boff:boff + 32].any()

    class FakeNet:
        def params(self):
            return p
    nm = Norms(np.arange(4, dtype=np.float32), np.ones(4, dtype=np.float32), np.array([2.0, 3.0], np.float32))
    path = tmp_path / "c.npz"
    save_params(str(path), FakeNet(), nm, {"t_start": 123, "workload": "cfg1"})
    q, nm2, meta = load_params(str(path))
    np.testing.assert_array_equal(q, p)
    np.testing.assert_array_equal(nm2.out_scale, nm.out_scale)
    assert meta == {"t_start": 123, "workload": "cfg1"}
    json.dumps(meta)
