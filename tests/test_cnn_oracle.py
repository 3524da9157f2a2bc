"""Pins the CPU CNN oracle (oracle/cnn_oracle.py) by checks that do not
restate it: the convolution against scipy's 2-D correlation, the backward
pass against central finite differences, the BF16 rounding against torch's
bfloat16 cast, and Adam's first step against its closed form."""
import numpy as np
import pytest
import scipy.signal
import torch

import cnn_oracle as co


def _case(seed=0, b=2, h=6, w=7):
    rng = np.random.default_rng(seed)
    x = np.zeros((b, h, w, 8), dtype=np.float32)
    x[..., :4] = rng.standard_normal((b, h, w, 4))
    tgt = rng.standard_normal((b, h, w, 2)).astype(np.float32)
    mask = rng.integers(0, 4, size=(b, h, w)).astype(np.uint8)
    p = co.init_params(seed + 1, out_scale=0.5)
    # non-zero biases so that their gradients and the ReLU masks are exercised
    for l, (woff, boff, kp) in enumerate(co.layout()):
        p[boff:boff + 32] = rng.standard_normal(32).astype(np.float32) * 0.1
    return p, x, tgt, mask


def test_layout_sizes():
    lay = co.layout()
    assert [kp for _, _, kp in lay] == [80, 288, 288, 288]
    assert co.num_params() == 32 * 80 + 32 + 3 * (32 * 288 + 32)


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1, 9, 11, 3))
    w = rng.standard_normal((4, 9, 3))
    b = rng.standard_normal(4)
    got = co.conv(x, w, b)
    for o in range(4):
        want = np.full((9, 11), b[o])
        for c in range(3):
            k = w[o, :, c].reshape(3, 3)  # taps t = 3 (dy + 1) + (dx + 1)
            want += scipy.signal.correlate2d(x[0, :, :, c], k, mode="same", boundary="fill")
        np.testing.assert_allclose(got[0, :, :, o], want, rtol=1e-12, atol=1e-12)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 10000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 65504.0, 1e-30], dtype=np.float32)])
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(co.bf16(a), want)


@pytest.mark.parametrize("seed,kind", [(0, "uv"), (1, "uv"), (2, "psi")])
def test_backward_matches_finite_differences(seed, kind):
    p, x, tgt, mask = _case(seed)
    p64 = p.astype(np.float64)
    _, g = co.backward(p64, x.astype(np.float64), tgt, mask, quant=False, kind=kind)
    rng = np.random.default_rng(seed + 10)
    lay = co.layout()
    idx = []
    for l, (woff, boff, kp) in enumerate(lay):
        cin = co.CIN[l]
        used = 4 if l == 0 else cin
        for _ in range(12):
            o = rng.integers(0, 2 if l == 3 else 32)
            t = rng.integers(0, 9)
            c = rng.integers(0, used)
            idx.append(woff + o * kp + t * cin + c)
        idx.extend(boff + j for j in ((0, 1) if l == 3 else (0, 5, 17)))
    eps = 1e-6
    for i in idx:
        pp = p64.copy()
        pp[i] += eps
        lp = co.loss(co.tendencies(co.forward(pp, x, quant=False)[0], kind), tgt, mask)[0]
        pp[i] -= 2 * eps
        lm = co.loss(co.tendencies(co.forward(pp, x, quant=False)[0], kind), tgt, mask)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_unused_parameters_have_zero_gradient():
    p, x, tgt, mask = _case(3)
    _, g = co.backward(p, x, tgt, mask, quant=True)
    woff, boff, kp = co.layout()[3]
    w3 = g[woff:woff + 32 * kp].reshape(32, kp)
    assert not w3[2:].any() and not g[boff + 2:boff + 32].any()
    woff0, _, kp0 = co.layout()[0]
    w0 = g[woff0:woff0 + 32 * kp0].reshape(32, kp0)[:, :72].reshape(32, 9, 8)
    assert not w0[:, :, 4:].any() and not g[woff0:woff0 + 32 * kp0].reshape(32, kp0)[:, 72:].any()


def test_quantised_close_to_exact():
    p, x, tgt, mask = _case(4)
    pq, _ = co.forward(p, x, quant=True)
    pe, _ = co.forward(p, x, quant=False)
    assert np.linalg.norm(pq - pe) / np.linalg.norm(pe) < 2e-2
    _, gq = co.backward(p, x, tgt, mask, quant=True)
    _, ge = co.backward(p, x, tgt, mask, quant=False)
    # BF16 rounding of weights, activations and back-propagated tendencies
    assert np.linalg.norm(gq - ge) / np.linalg.norm(ge) < 6e-2


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(2)
    p = rng.standard_normal(100).astype(np.float32)
    g = rng.standard_normal(100).astype(np.float32)
    z = np.zeros(100, dtype=np.float32)
    p1, m, v = co.adam(p, g, z, z, 1, lr=1e-3)
    # t = 1: m_hat = g, v_hat = g^2, so the step is lr g / (|g| + eps)
    np.testing.assert_allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=1e-5)


def test_psi_tendencies_are_divergence_free():
    """Interior T cells: du(x, y) - du(x-1, y) + dv(x, y) - dv(x, y-1) = 0 for
    any psi (the C-grid divergence of the stream-function tendencies)."""
    rng = np.random.default_rng(7)
    psi = rng.standard_normal((2, 9, 11))
    du, dv = co.psi_uv(psi)
    div = du[:, 1:, 1:] - du[:, 1:, :-1] + dv[:, 1:, 1:] - dv[:, :-1, 1:]
    assert np.abs(div).max() < 1e-12
    # and psi = y (a uniform zonal flow) gives du = -1, dv = 0 inside
    yy = np.broadcast_to(np.arange(9.0)[None, :, None], (1, 9, 11))
    du, dv = co.psi_uv(yy)
    assert np.all(du[:, 1:, :] == -1.0) and np.all(dv[:, :, 1:] == 0.0)
