"""The closure network's GPU kernels against the CPU oracle.

Layer by layer (test_layer_*): the oracle is fed the device's own BF16
inputs of each layer (its activations, its back-propagated dL/dz), so only
the FP32 accumulation order differs from the float64 sums.  Every element
must lie within half a BF16 ulp of the float64 value plus the FP32
accumulation bound 2^-24 K S (K terms, S the sum of their magnitudes: a
relative 1.7e-5 at K = 288), at least 99 % must equal the oracle's own
BF16 rounding, and the FP32 outputs and weight gradients agree to relative
L2 2e-5; shapes include widths that are not multiples of 32.

End to end (the network run by each side alone): single BF16 roundings
can flip and propagate, so forward outputs agree to relative L2 <= 1e-2,
each layer's gradient to <= 2e-2, the Adam update to float32 rounding."""
import numpy as np
import pytest

import cnn_oracle as co
from _util import rel_l2
from inmsom_bt.closure import ClosureNet, features
from inmsom_bt.ensemble import Ensemble
from inmsom_bt.cases import randomize_state
from inmsom_bt.workloads import ensemble

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-2
GRAD_TOL = 2e-2


def _batch(seed, b, h, w):
    rng = np.random.default_rng(seed)
    x = np.zeros((b, h, w, 8), dtype=np.float32)
    x[..., :4] = rng.standard_normal((b, h, w, 4))
    tgt = rng.standard_normal((b, h, w, 2)).astype(np.float32)
    mask = rng.integers(0, 4, size=(b, h, w)).astype(np.uint8)
    return x, tgt, mask


def _params(seed):
    p = co.init_params(seed, out_scale=0.5)
    rng = np.random.default_rng(seed + 100)
    for woff, boff, kp in co.layout():
        p[boff:boff + 32] = rng.standard_normal(32).astype(np.float32) * 0.1
    return p


# shapes: one tile, ragged tiles (width not a multiple of 32), several images
SHAPES = [(1, 8, 32), (2, 13, 45), (3, 20, 70)]


# one tile, ragged tiles (widths 33, 45, 70, 97: not multiples of 32), several images
LAYER_SHAPES = [(1, 8, 32), (2, 13, 45), (1, 7, 33), (3, 20, 70), (1, 5, 97)]
EPS32 = 2.0 ** -24


def _bf16_ulp(a):
    """Spacing of BF16 numbers at |a| (float64)."""
    m = np.maximum(np.abs(a), 2.0 ** -126)
    return 2.0 ** (np.floor(np.log2(m)) - 7)


def _check_bf16(got, exact, err, label):
    """got: the device's BF16 values; exact: the float64 value it rounded
    (to within err, the FP32 accumulation bound)."""
    got = got.astype(np.float64)
    bound = 0.5 * _bf16_ulp(np.abs(exact) + err) + err
    bad = np.abs(got - exact) > bound
    assert not bad.any(), (label, int(bad.sum()), np.argwhere(bad)[:5].tolist())
    same = got == co.bf16(exact.astype(np.float32)).astype(np.float64)
    assert same.mean() >= 0.99, (label, same.mean())


def _conv_abs(x, w, b):
    """Sum of the magnitudes of the terms of each output of conv(x, w, b)."""
    return co.conv(np.abs(x), np.abs(w), np.abs(b))


@pytest.mark.parametrize("b,h,w", LAYER_SHAPES)
def test_layer_forward_on_device_inputs(b, h, w):
    print(f"cnn layer forward {b}x{h}x{w}", flush=True)
    p = _params(11)
    x, _, _ = _batch(12, b, h, w)
    with ClosureNet(b, h, w) as net:
        net.set_params(p)
        net.set_batch(x)
        pred = net.forward()
        acts = [co.bf16(x)] + [net.activation(l) for l in range(3)]
    for l in range(4):
        wl, bl = co.weights(p, l)
        wq = co.bf16(wl)
        z = co.conv(acts[l], wq, bl)
        err = EPS32 * (9 * co.CIN[l] + 1) * _conv_abs(acts[l], wq, bl)
        if l < 3:
            _check_bf16(acts[l + 1], np.maximum(z, 0.0), err, f"activation {l}")
        else:
            exact = z[..., :co.NOUT]
            dev = np.abs(pred - exact)
            assert (dev <= err[..., :co.NOUT] + EPS32 * np.abs(exact)).all(), ("output", dev.max())
            assert rel_l2(pred, exact) <= 2e-5, rel_l2(pred, exact)


@pytest.mark.parametrize("b,h,w", LAYER_SHAPES)
def test_layer_backward_on_device_inputs(b, h, w):
    print(f"cnn layer backward {b}x{h}x{w}", flush=True)
    p = _params(13)
    x, tgt, mask = _batch(14, b, h, w)
    with ClosureNet(b, h, w) as net:
        net.set_params(p)
        net.keep_backprop(True)
        net.set_batch(x, tgt, mask)
        net.loss_grad()
        pred = net.forward()
        acts = [co.bf16(x)] + [net.activation(l) for l in range(3)]
        gz = [net.backprop(l) for l in range(4)]
        g = net.grad()
    # dL/dz_3 from the device's outputs, in the device's FP32 operations
    m = np.stack([(mask & 1) != 0, (mask & 2) != 0], axis=-1).astype(np.float32)
    count = int(m.sum())
    d = (pred.astype(np.float32) - tgt) * m
    want3 = co.bf16(np.float32(2.0) * d * (np.float32(1.0) / np.float32(count)))
    assert np.array_equal(gz[3][..., :co.NOUT], want3), "dL/dz_3"
    assert not gz[3][..., co.NOUT:].any()
    for l in range(3, -1, -1):
        woff, boff, kp = co.layout()[l]
        gz64 = gz[l].astype(np.float64)
        cols = co._cols(acts[l].astype(np.float64))
        dw = np.einsum("bhwo,bhwtc->otc", gz64, cols).reshape(co.CH, -1)
        gw = g[woff:woff + co.CH * kp].reshape(co.CH, kp)
        assert rel_l2(gw[:, :9 * co.CIN[l]], dw) <= 2e-5, (l, "W", rel_l2(gw[:, :9 * co.CIN[l]], dw))
        assert not gw[:, 9 * co.CIN[l]:].any()
        db = gz64.sum(axis=(0, 1, 2))
        assert rel_l2(g[boff:boff + co.CH], db) <= 2e-5, (l, "b")
        if l > 0:
            wq = co.bf16(co.weights(p, l)[0]).astype(np.float64)
            dx = co._col2im(np.einsum("bhwo,otc->bhwtc", gz64, wq))
            sx = co._col2im(np.einsum("bhwo,otc->bhwtc", np.abs(gz64), np.abs(wq)))
            live = acts[l] > 0
            exact = dx * live
            err = EPS32 * 9 * co.CH * sx * live
            _check_bf16(gz[l - 1], exact, err, f"dL/dz_{l - 1}")


@pytest.mark.parametrize("b,h,w", SHAPES)
def test_forward_matches_oracle(b, h, w):
    print(f"cnn forward {b}x{h}x{w}", flush=True)
    p = _params(1)
    x, _, _ = _batch(2, b, h, w)
    with ClosureNet(b, h, w) as net:
        net.set_params(p)
        net.set_batch(x)
        got = net.forward()
        acts = [net.activation(l) for l in range(3)]
    want, wacts = co.forward(p, x, quant=True)
    assert rel_l2(got, want) <= FWD_TOL, rel_l2(got, want)
    for l in range(3):
        assert rel_l2(acts[l], wacts[l + 1]) <= FWD_TOL, (l, rel_l2(acts[l], wacts[l + 1]))


@pytest.mark.parametrize("b,h,w", SHAPES)
def test_gradient_matches_oracle(b, h, w):
    print(f"cnn gradient {b}x{h}x{w}", flush=True)
    p = _params(3)
    x, tgt, mask = _batch(4, b, h, w)
    with ClosureNet(b, h, w) as net:
        net.set_params(p)
        net.set_batch(x, tgt, mask)
        loss = net.loss_grad()
        g = net.grad()
    wl, wg = co.backward(p, x, tgt, mask, quant=True)
    assert abs(loss - wl) <= 1e-2 * abs(wl), (loss, wl)
    for l, (woff, boff, kp) in enumerate(co.layout()):
        gw, ww = g[woff:woff + 32 * kp], wg[woff:woff + 32 * kp]
        gb, wb = g[boff:boff + 32], wg[boff:boff + 32]
        assert rel_l2(gw, ww) <= GRAD_TOL, (l, "W", rel_l2(gw, ww))
        assert rel_l2(gb, wb) <= GRAD_TOL, (l, "b", rel_l2(gb, wb))
        # padding and unused outputs get exactly zero
        full = gw.reshape(32, kp)
        assert not full[:, 9 * co.CIN[l]:].any()
        if l == 3:
            assert not full[2:].any() and not gb[2:].any()


def test_adam_step_matches_oracle():
    b, h, w = 2, 16, 40
    p = _params(5)
    x, tgt, mask = _batch(6, 3, h, w)
    with ClosureNet(b, h, w) as net:
        net.set_params(p)
        net.dataset_alloc(3)
        for i in range(3):
            net.dataset_set(i, x[i], tgt[i], mask[i])
        pk = p.copy()
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        for t, idx in enumerate(([0, 2], [1, 0], [2, 1]), start=1):
            loss = net.train_step(idx, lr=1e-3)
            g = net.grad()
            got = net.params()
            want, m, v = co.adam(pk, g, m, v, t, lr=1e-3)
            np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-9)
            wl, wg = co.backward(pk, x[idx], tgt[idx], mask[idx], quant=True)
            assert abs(loss - wl) <= 1e-2 * abs(wl)
            assert rel_l2(g, wg) <= GRAD_TOL
            pk = got


def test_training_follows_oracle_training():
    """Device training (forward, backward, Adam on the device-resident set)
    follows the oracle's own Adam run on the same batches: the loss curves
    agree within 5 % at every step, and both fall by at least a third in 60
    steps of learning a teacher network."""
    b, h, w, n, steps = 4, 24, 40, 16, 60
    rng = np.random.default_rng(8)
    x = np.zeros((n, h, w, 8), dtype=np.float32)
    x[..., :4] = rng.standard_normal((n, h, w, 4))
    teacher = co.init_params(9, out_scale=1.0)
    tgt = co.forward(teacher, x, quant=False)[0].astype(np.float32)
    mask = np.full((n, h, w), 3, dtype=np.uint8)
    batches = [rng.choice(n, size=b, replace=False) for _ in range(steps)]
    p0 = co.init_params(10, out_scale=0.1)
    with ClosureNet(b, h, w) as net:
        net.set_params(p0)
        net.dataset_alloc(n)
        for i in range(n):
            net.dataset_set(i, x[i], tgt[i], mask[i])
        gpu = [net.train_step(idx, lr=1e-3) for idx in batches]
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    ref = []
    for t, idx in enumerate(batches, start=1):
        loss, g = co.backward(p, x[idx], tgt[idx], mask[idx], quant=True)
        ref.append(loss)
        p, m, v = co.adam(p, g.astype(np.float32), m, v, t, lr=1e-3)
    gpu, ref = np.array(gpu), np.array(ref)
    print(f"loss gpu {gpu[:3].mean():.4f} -> {gpu[-3:].mean():.4f}, oracle {ref[:3].mean():.4f} -> "
          f"{ref[-3:].mean():.4f}, max rel gap {np.max(np.abs(gpu - ref) / ref):.3e}", flush=True)
    assert np.all(np.abs(gpu - ref) <= 0.05 * ref)
    assert gpu[-3:].mean() * 1.5 <= gpu[:3].mean()


def _lcu_lcv(bn):
    wet = bn.wet
    lcu = np.zeros_like(wet)
    lcv = np.zeros_like(wet)
    lcu[1:-2, 1:-2] = wet[1:-2, 1:-2] & wet[1:-2, 2:-1]
    lcv[1:-2, 1:-2] = wet[1:-2, 1:-2] & wet[2:-1, 1:-2]
    return lcu, lcv


def _closure_members():
    basins = ensemble(3, seed=11, size_lo=(40, 30), size_hi=(90, 70), sigma=5.0)
    return [randomize_state(bn, 20 + k) for k, bn in enumerate(basins)]


NORM = dict(shift=[0.0, 0.0, 0.0, 100.0], scale=[5.0, 5.0, 3.0, 0.01], out_scale=[1e-5, 1e-5])


def test_ensemble_closure_matches_oracle():
    """One outer step with the closure equals the plain step plus dt times
    the oracle network's tendency on U/V points; ssh is untouched."""
    basins = _closure_members()
    p = co.init_params(12, out_scale=0.3)
    dt = basins[0].dt
    with Ensemble(basins, mode="fp64") as plain, Ensemble(basins, mode="fp64") as cl, \
            ClosureNet(1, 8, 32) as net:
        net.set_params(p)
        net.set_norm(**NORM)
        cl.set_closure(net, ratio=1e30)
        plain.run(1, nstep=2)
        cl.run(1, nstep=2)
        nfall, norms = cl.closure_stats()
        for k, bn in enumerate(basins):
            a, b = cl.get_state(k), plain.get_state(k)
            for f in ("ssh", "sshp", "up", "vp"):
                assert np.array_equal(a[f], b[f]), f
            x = features(b["u"], b["v"], b["ssh"], bn.hhq_rest, NORM["shift"], NORM["scale"])
            pred = co.forward(p, x[None], quant=True)[0][0]
            lcu, lcv = _lcu_lcv(bn)
            du = dt * NORM["out_scale"][0] * pred[..., 0] * lcu
            dv = dt * NORM["out_scale"][1] * pred[..., 1] * lcv
            assert rel_l2(a["u"] - b["u"], du) <= FWD_TOL, rel_l2(a["u"] - b["u"], du)
            assert rel_l2(a["v"] - b["v"], dv) <= FWD_TOL, rel_l2(a["v"] - b["v"], dv)
            assert not (a["u"] - b["u"])[~lcu].any() and not (a["v"] - b["v"])[~lcv].any()
            # reported norms: closure tendency and nu * Laplacian of the plain step
            cn = np.sqrt((du[lcu] ** 2).sum() + (dv[lcv] ** 2).sum()) / dt
            nu = float(bn.mu.max())

            def lap(f):
                out = np.zeros_like(f)
                out[1:-1, 1:-1] = (f[1:-1, 2:] + f[1:-1, :-2] + f[2:, 1:-1] + f[:-2, 1:-1]
                                   - 4 * f[1:-1, 1:-1])
                return nu * out / (bn.dx[2, 2] ** 2)
            vn = np.sqrt((lap(b["u"])[lcu] ** 2).sum() + (lap(b["v"])[lcv] ** 2).sum())
            assert abs(norms[k, 0] - cn) <= FWD_TOL * cn
            assert abs(norms[k, 1] - vn) <= 1e-10 * vn
        assert not nfall.any()


def test_ensemble_closure_fallback_and_graph_replays():
    """ratio ~ 0 makes every member fall back (bitwise the plain step); graph
    replays with the closure equal direct launches, replay after replay."""
    basins = _closure_members()
    p = co.init_params(13, out_scale=0.3)
    with Ensemble(basins, mode="fp32") as plain, Ensemble(basins, mode="fp32") as cl, \
            ClosureNet(1, 8, 32) as net:
        net.set_params(p)
        net.set_norm(**NORM)
        cl.set_closure(net, ratio=1e-30)
        plain.run(3, nstep=2)
        cl.run(3, nstep=2)
        nfall, _ = cl.closure_stats()
        assert (nfall == 3).all()
        for k in range(len(basins)):
            a, b = cl.get_state(k), plain.get_state(k)
            for f in a:
                assert np.array_equal(a[f], b[f]), (k, f)
    with Ensemble(basins, mode="fp32") as g, Ensemble(basins, mode="fp32") as d, \
            ClosureNet(1, 8, 32) as net:
        net.set_params(p)
        net.set_norm(**NORM)
        g.set_closure(net, ratio=1e30)
        d.set_closure(net, ratio=1e30)
        for rep in range(3):
            g.run(1, nstep=2, graph=True)
            d.run(1, nstep=2, graph=False)
            for k in range(len(basins)):
                a, b = g.get_state(k), d.get_state(k)
                for f in a:
                    assert np.array_equal(a[f], b[f]), (rep, k, f)


@pytest.mark.parametrize("interval", [2, 4])
def test_closure_interval(interval):
    """The closure after every interval-th step: graph replays equal direct
    launches bit for bit, a run split at any step (3 + 9 + 4) equals one
    run of 16 (the count of steps since the closure carries over), the
    closure runs 16 / interval times, and its tendency is applied with
    interval steps' worth of time (16 steps at interval k, tendency x k
    every k steps, against interval 1: the same order of magnitude of
    change from the plain run, not equal)."""
    basins = _closure_members()
    p = co.init_params(15, out_scale=0.3)
    states = {}
    with ClosureNet(1, 8, 32) as net:
        net.set_params(p)
        net.set_norm(**NORM)
        for label, graph, parts, iv in (("graph", True, (16,), interval), ("direct", False, (16,), interval),
                                        ("split", True, (3, 9, 4), interval), ("every", True, (16,), 1)):
            print(f"closure interval {iv}: {label} {parts}", flush=True)
            with Ensemble(basins, mode="fp32") as ens:
                ens.set_closure(net, ratio=1e30, interval=iv)
                for n in parts:
                    ens.run(n, nstep=2, graph=graph)
                nfall, norms = ens.closure_stats()
                assert (nfall == 0).all()
                states[label] = [ens.get_state(k) for k in range(len(basins))]
        with Ensemble(basins, mode="fp32") as ens:
            ens.run(16, nstep=2)
            states["plain"] = [ens.get_state(k) for k in range(len(basins))]
    for k in range(len(basins)):
        for f in states["graph"][k]:
            assert np.array_equal(states["graph"][k][f], states["direct"][k][f]), ("direct", k, f)
            assert np.array_equal(states["graph"][k][f], states["split"][k][f]), ("split", k, f)
        du = np.abs(states["graph"][k]["u"] - states["plain"][k]["u"]).max()
        de = np.abs(states["every"][k]["u"] - states["plain"][k]["u"]).max()
        assert 0.3 * de < du < 3.0 * de, (k, du, de)


@pytest.mark.parametrize("b,h,w", [(1, 8, 32), (2, 13, 45)])
def test_gradient_psi_output_matches_oracle(b, h, w):
    """Stream-function outputs: loss on the differenced tendencies, gradient
    through their adjoint."""
    p = _params(21)
    x, tgt, mask = _batch(22, b, h, w)
    with ClosureNet(b, h, w) as net:
        net.set_output("psi")
        net.set_params(p)
        net.set_batch(x, tgt, mask)
        loss = net.loss_grad()
        g = net.grad()
    wl, wg = co.backward(p, x, tgt, mask, quant=True, kind="psi")
    assert abs(loss - wl) <= 1e-2 * abs(wl), (loss, wl)
    for l, (woff, boff, kp) in enumerate(co.layout()):
        assert rel_l2(g[woff:woff + 32 * kp], wg[woff:woff + 32 * kp]) <= GRAD_TOL, l
        assert rel_l2(g[boff:boff + 32], wg[boff:boff + 32]) <= GRAD_TOL, l


def test_ensemble_closure_psi_matches_oracle():
    """Stream-function closure in the step: u, v change by dt times the
    differenced psi on U/V points."""
    basins = _closure_members()
    p = co.init_params(14, out_scale=0.3)
    dt = basins[0].dt
    with Ensemble(basins, mode="fp64") as plain, Ensemble(basins, mode="fp64") as cl, \
            ClosureNet(1, 8, 32) as net:
        net.set_output("psi")
        net.set_params(p)
        net.set_norm(**NORM)
        cl.set_closure(net, ratio=1e30)
        plain.run(1, nstep=2)
        cl.run(1, nstep=2)
        for k, bn in enumerate(basins):
            a, b = cl.get_state(k), plain.get_state(k)
            x = features(b["u"], b["v"], b["ssh"], bn.hhq_rest, NORM["shift"], NORM["scale"])
            t = co.tendencies(co.forward(p, x[None], quant=True)[0], "psi")[0]
            lcu, lcv = _lcu_lcv(bn)
            du = dt * NORM["out_scale"][0] * t[..., 0] * lcu
            dv = dt * NORM["out_scale"][0] * t[..., 1] * lcv
            assert rel_l2(a["u"] - b["u"], du) <= FWD_TOL, rel_l2(a["u"] - b["u"], du)
            assert rel_l2(a["v"] - b["v"], dv) <= FWD_TOL, rel_l2(a["v"] - b["v"], dv)


def test_closure_graphs_follow_norm_changes():
    """The closure's normalisation is captured into the ensemble's graphs by
    value: after set_norm (and set_output) on an attached network, graph
    replays equal direct launches bit for bit.  Destroying an attached
    network detaches it: the ensemble's next run fails with a clear error
    (no replay of freed memory) until a closure is set or cleared."""
    from inmsom_bt._capi import BtError

    basins = _closure_members()
    p = co.init_params(12, out_scale=0.3)
    other = dict(shift=[0.01, -0.02, 0.0, 90.0], scale=[4.0, 6.0, 2.0, 0.02], out_scale=[2e-5, 0.5e-5])
    with Ensemble(basins, mode="fp64") as g, Ensemble(basins, mode="fp64") as d:
        net = ClosureNet(1, 8, 32)
        net.set_params(p)
        net.set_norm(**NORM)
        g.set_closure(net, ratio=1e30)
        d.set_closure(net, ratio=1e30)
        for stage, change in enumerate((None, "norm", "output", None)):
            print(f"closure graph stage {stage}: {change}", flush=True)
            if change == "norm":
                net.set_norm(**other)
            elif change == "output":
                net.set_output("psi")
            g.run(9, nstep=2, graph=True)
            d.run(9, nstep=2, graph=False)
            for k in range(len(basins)):
                a, b = g.get_state(k), d.get_state(k)
                for f in ("ssh", "sshp", "u", "up", "v", "vp"):
                    assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), (stage, k, f)
        net.close()
        for ens, graph in ((g, True), (d, False)):
            with pytest.raises(BtError, match="was destroyed"):
                ens.run(1, nstep=2, graph=graph)
            ens.set_closure(None)
            ens.run(1, nstep=2, graph=graph)


@pytest.mark.parametrize("mode", ["fp32", "fp64"])
def test_fused_inference_equals_layer_by_layer(mode):
    """The closure's four layers in one pass per tile (k_closure4) give the
    tendencies of one pass per layer bit for bit: 12 outer steps of a
    ragged 4-member ensemble (sizes that are not multiples of the 8 x 32
    tiles) with the closure always applied and under the 3x rule, graph
    replays and direct launches, the same states and closure statistics."""
    basins = _closure_members() + [randomize_state(ensemble(1, seed=5, size_lo=(33, 9), size_hi=(33, 9),
                                                            sigma=3.0)[0], 9)]
    p = co.init_params(17, out_scale=0.3)
    out = {}
    with ClosureNet(1, 8, 32) as net:
        net.set_params(p)
        net.set_norm(**NORM)
        for fused in ("always", "never"):
            net.set_fused(fused)
            for ratio in (1e30, 3.0):
                for graph in (True, False):
                    print(f"{mode} closure inference {fused}, ratio {ratio:g}, graph {graph}", flush=True)
                    with Ensemble(basins, mode=mode) as ens:
                        ens.set_closure(net, ratio=ratio)
                        ens.run(12, nstep=2, graph=graph)
                        nfall, norms = ens.closure_stats()
                        out[(fused, ratio, graph)] = ([ens.get_state(k) for k in range(len(basins))],
                                                      np.array(nfall), np.array(norms))
        net.set_fused("auto")
    for (fused, ratio, graph), (st, nf, nr) in out.items():
        ref_st, ref_nf, ref_nr = out[("never", ratio, True)]
        assert np.array_equal(nf, ref_nf) and np.array_equal(nr.view(np.uint64), ref_nr.view(np.uint64)), \
            (fused, ratio, graph)
        for k in range(len(basins)):
            for f in ref_st[k]:
                assert np.array_equal(st[k][f].view(np.uint64), ref_st[k][f].view(np.uint64)), (fused, ratio, graph,
                                                                                               k, f)
    assert (out[("never", 3.0, True)][2][:, 0] > 0).all()              # every member had a closure tendency
