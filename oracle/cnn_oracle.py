"""CPU restatement of the learned-closure CNN: forward, loss, backward, Adam.

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in
inmsom-elfs_amd/csrc/bt_cnn.hip.  Only tests/ and __graft_entry__.smoke()
use it.

The network has no counterpart in the reference (whose lateral momentum
closure is the viscosity of Function/mixing.f90:1-146), so this module
restates the network's mathematical definition:

    z_l[p, co] = b_l[co] + sum_{t, ci} W_l[co][t cin_l + ci] a_{l-1}[p + o_t, ci]
    a_l = relu(z_l) (l = 0..2),  pred = z_3[:, :, :, 0:2]
    L = sum_{p, c} m_c(p) (pred[p, c] - target[p, c])^2 / count

with taps t = 3 (dy + 1) + (dx + 1), o_t = (dy, dx), and zero padding.
The parameter vector is laid out as the C-ABI states: per layer W[32][kpad]
(kpad = 9 cin rounded up to 16, cin = 8, 32, 32, 32) then b[32].

quant=True rounds to BF16 exactly where the device does (inputs, weights,
hidden activations, back-propagated tendencies); quant=False is the plain
float64 definition, used to pin the backward pass against finite
differences.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

CIN = (8, 32, 32, 32)
CH = 32
NOUT = 2


def kpad(cin: int) -> int:
    return (9 * cin + 15) // 16 * 16


def layout() -> List[Tuple[int, int, int]]:
    """(woff, boff, kpad) per layer."""
    out, off = [], 0
    for cin in CIN:
        kp = kpad(cin)
        out.append((off, off + CH * kp, kp))
        off += CH * kp + CH
    return out


def num_params() -> int:
    woff, boff, kp = layout()[-1]
    return boff + CH


def init_params(seed: int = 0, out_scale: float = 1.0) -> np.ndarray:
    """He-normal weights on the used entries, zero biases and padding; the
    last layer is scaled by out_scale."""
    rng = np.random.default_rng(seed)
    p = np.zeros(num_params(), dtype=np.float32)
    for l, (woff, boff, kp) in enumerate(layout()):
        cin_used = 4 if l == 0 else CH
        w = np.zeros((CH, 9, CIN[l]), dtype=np.float64)
        w[:, :, :cin_used] = rng.standard_normal((CH, 9, cin_used)) * np.sqrt(2.0 / (9 * cin_used))
        if l == 3:
            w *= out_scale
        full = np.zeros((CH, kp))
        full[:, :9 * CIN[l]] = w.reshape(CH, 9 * CIN[l])
        p[woff:woff + CH * kp] = full.reshape(-1).astype(np.float32)
    return p


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to BF16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = np.where(nan, a, out)
    return out


def weights(params: np.ndarray, l: int) -> Tuple[np.ndarray, np.ndarray]:
    """(W [32, 9, cin], b [32]) of layer l."""
    woff, boff, kp = layout()[l]
    w = params[woff:woff + CH * kp].reshape(CH, kp)[:, :9 * CIN[l]].reshape(CH, 9, CIN[l])
    return w, params[boff:boff + CH]


def _cols(x: np.ndarray) -> np.ndarray:
    """im2col: x [B, H, W, C] -> [B, H, W, 9, C] with zero padding."""
    b, h, w, c = x.shape
    xp = np.zeros((b, h + 2, w + 2, c), dtype=x.dtype)
    xp[:, 1:-1, 1:-1] = x
    out = np.empty((b, h, w, 9, c), dtype=x.dtype)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        out[:, :, :, t] = xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
    return out


def _col2im(dc: np.ndarray) -> np.ndarray:
    """adjoint of _cols: [B, H, W, 9, C] -> [B, H, W, C]."""
    b, h, w, _, c = dc.shape
    xp = np.zeros((b, h + 2, w + 2, c), dtype=dc.dtype)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] += dc[:, :, :, t]
    return xp[:, 1:-1, 1:-1]


def conv(x: np.ndarray, w: np.ndarray, b: np.ndarray) -> np.ndarray:
    """3x3 convolution (cross-correlation) in float64: x [B,H,W,cin], w [co,9,cin]."""
    cols = _cols(x.astype(np.float64))
    return np.einsum("bhwtc,otc->bhwo", cols, w.astype(np.float64)) + b.astype(np.float64)


def forward(params: np.ndarray, x: np.ndarray, quant: bool = True):
    """Returns (pred [B,H,W,2], activations [x, a0, a1, a2])."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    acts = [q(x)]
    for l in range(4):
        w, b = weights(params, l)
        z = conv(acts[-1], q(w), b)
        if l < 3:
            acts.append(q(np.maximum(z, 0.0).astype(np.float32) if quant else np.maximum(z, 0.0)))
        else:
            pred = z[..., :NOUT]
    return pred, acts


def psi_uv(psi: np.ndarray):
    """Stream-function output (channel 0, H points) -> u, v tendencies:
    du(x, y) = -(psi(x, y) - psi(x, y-1)), dv(x, y) = psi(x, y) - psi(x-1, y),
    zero psi outside the image; psi [B, H, W]."""
    s = np.zeros_like(psi)
    s[:, 1:, :] = psi[:, :-1, :]
    w = np.zeros_like(psi)
    w[:, :, 1:] = psi[:, :, :-1]
    return -(psi - s), psi - w


def tendencies(pred: np.ndarray, kind: str = "uv") -> np.ndarray:
    """Network outputs [B, H, W, 2] -> u, v tendencies [B, H, W, 2] (normalised)."""
    if kind == "uv":
        return pred
    du, dv = psi_uv(pred[..., 0])
    return np.stack([du, dv], axis=-1)


def loss(pred: np.ndarray, target: np.ndarray, mask: np.ndarray) -> Tuple[float, np.ndarray, int]:
    """(loss, d loss / d pred, count) of the masked mean-square error."""
    m = np.stack([(mask & 1) != 0, (mask & 2) != 0], axis=-1).astype(np.float64)
    count = int(m.sum())
    d = (pred - target) * m
    return float((d * d).sum() / max(count, 1)), 2.0 * d / max(count, 1), count


def backward(params: np.ndarray, x: np.ndarray, target: np.ndarray, mask: np.ndarray,
             quant: bool = True, kind: str = "uv"):
    """(loss, gradient vector in the parameter layout).  kind 'psi': the loss
    is on tendencies(pred, 'psi') and dL/dpsi is the adjoint of psi_uv."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    pred, acts = forward(params, x, quant)
    m = np.stack([(mask & 1) != 0, (mask & 2) != 0], axis=-1)
    count = int(m.sum())
    d = (tendencies(pred, kind) - target) * m
    lval = float((d * d).sum() / max(count, 1))
    if kind == "psi":
        # psi(x, y) enters du(x, y) (-1), du(x, y+1) (+1), dv(x, y) (+1), dv(x+1, y) (-1)
        ru, rv = d[..., 0], d[..., 1]
        ru_n = np.zeros_like(ru)
        ru_n[:, :-1, :] = ru[:, 1:, :]
        rv_e = np.zeros_like(rv)
        rv_e[:, :, :-1] = rv[:, :, 1:]
        adj = -ru + ru_n + rv - rv_e
        d = np.stack([adj, np.zeros_like(adj)], axis=-1)
    if quant:
        inv = np.float32(1.0) / np.float32(max(count, 1))
        g = bf16((np.float32(2.0) * inv * d.astype(np.float32)).astype(np.float32))
    else:
        g = 2.0 * d / max(count, 1)
    gz = np.zeros(pred.shape[:3] + (CH,))
    gz[..., :NOUT] = g
    grad = np.zeros(num_params(), dtype=np.float64)
    for l in range(3, -1, -1):
        woff, boff, kp = layout()[l]
        w, _ = weights(params, l)
        wq = q(w).astype(np.float64)
        cols = _cols(acts[l].astype(np.float64))  # [B,H,W,9,cin]
        dw = np.einsum("bhwo,bhwtc->otc", gz, cols)
        full = np.zeros((CH, kp))
        full[:, :9 * CIN[l]] = dw.reshape(CH, -1)
        grad[woff:woff + CH * kp] = full.reshape(-1)
        grad[boff:boff + CH] = gz.sum(axis=(0, 1, 2))
        if l > 0:
            dx = _col2im(np.einsum("bhwo,otc->bhwtc", gz, wq))
            dz = dx * (acts[l] > 0)
            gz = q(dz.astype(np.float32)).astype(np.float64) if quant else dz
    return lval, grad


def adam(params: np.ndarray, grad: np.ndarray, m: np.ndarray, v: np.ndarray, t: int,
         lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam step in float32 (Kingma & Ba 2015, Algorithm 1); returns new
    (params, m, v).  The constants 1 - beta and 1 - beta^t are formed in
    float64 and rounded once."""
    f = np.float32
    g = grad.astype(f)
    m = f(b1) * m.astype(f) + f(1.0 - b1) * g
    v = f(b2) * v.astype(f) + f(1.0 - b2) * g * g
    c1 = f(1.0 - b1 ** t)
    c2 = f(1.0 - b2 ** t)
    p = params.astype(f) - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return p.astype(f), m, v
