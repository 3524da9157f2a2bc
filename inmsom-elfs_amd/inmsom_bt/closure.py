"""Learned lateral-momentum closure (host side of inmsom_bt_cnn_*).

``ClosureNet`` owns one network on one device: FP32 master parameters, their
BF16 copies, Adam moments, a training batch of ``batch`` images of
``ny x nx`` grid points and an optional device-resident training set.  The
forward, backward and Adam kernels are the library's own (BF16 matrix-core
convolutions with FP32 accumulation); nothing here computes on the CPU.

Network inputs per grid point are (u, v, ssh, depth) of a basin's (ny, nx)
arrays, normalised as x_c = (field_c - shift_c) * scale_c and stored in
channels 0..3 of 8; outputs times ``out_scale`` are the u and v tendencies
[m/s^2] at the U and V points of the same array index.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _capi
from ._capi import check, lib

CIN0, CH, NIN, NOUT = 8, 32, 4, 2
OUTPUT_KINDS = {"uv": 0, "psi": 1}


def _f(a: np.ndarray):
    return a.ctypes.data_as(_capi._fp)


def features(u, v, ssh, depth, shift: Sequence[float], scale: Sequence[float]) -> np.ndarray:
    """Network inputs [ny, nx, 8] (float32) of one basin, computed as the
    device's feature kernel does: float32(field), then (x - shift) * scale in
    float32."""
    out = np.zeros(u.shape + (CIN0,), dtype=np.float32)
    for c, a in enumerate((u, v, ssh, depth)):
        out[..., c] = (np.asarray(a, dtype=np.float32) - np.float32(shift[c])) * np.float32(scale[c])
    return out


class ClosureNet:
    def __init__(self, batch: int, ny: int, nx: int, device: int = 0):
        self.batch, self.ny, self.nx = batch, ny, nx
        h = C.c_void_p()
        check(lib().inmsom_bt_cnn_create(batch, ny, nx, device, C.byref(h)), "inmsom_bt_cnn_create")
        self._h = h
        n = C.c_longlong()
        check(lib().inmsom_bt_cnn_num_params(self._h, C.byref(n)), "inmsom_bt_cnn_num_params")
        self.nparam = n.value
        self.shift = np.zeros(NIN, dtype=np.float32)
        self.scale = np.ones(NIN, dtype=np.float32)
        self.out_scale = np.ones(NOUT, dtype=np.float32)
        self.nsamples = 0
        self.kind = "uv"

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            # an ensemble holding it as its closure drops it first (its next run fails)
            check(lib().inmsom_bt_cnn_destroy(self._h), "inmsom_bt_cnn_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- setup
    def set_norm(self, shift, scale, out_scale):
        self.shift = np.ascontiguousarray(shift, dtype=np.float32)
        self.scale = np.ascontiguousarray(scale, dtype=np.float32)
        self.out_scale = np.ascontiguousarray(out_scale, dtype=np.float32)
        check(lib().inmsom_bt_cnn_set_norm(self._h, _f(self.shift), _f(self.scale), _f(self.out_scale)),
              "inmsom_bt_cnn_set_norm")

    def set_output(self, kind: str = "uv"):
        """'uv': outputs are the u, v tendencies; 'psi': output 0 is a
        stream-function tendency on H points, differenced into
        divergence-free u, v tendencies (inmsom_bt_cnn_set_output)."""
        check(lib().inmsom_bt_cnn_set_output(self._h, OUTPUT_KINDS[kind]), "inmsom_bt_cnn_set_output")
        self.kind = kind

    def set_fused(self, mode: str = "auto"):
        """The closure's inference in the step: 'always' the four layers in
        one pass per tile, 'never' one pass per layer, 'auto' by batch size
        (inmsom_bt_cnn_set_fused); the same tendencies either way."""
        modes = {"auto": 0, "never": 1, "always": 2}
        check(lib().inmsom_bt_cnn_set_fused(self._h, modes[mode]), "inmsom_bt_cnn_set_fused")

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {p.size}")
        check(lib().inmsom_bt_cnn_set_params(self._h, _f(p)), "inmsom_bt_cnn_set_params")

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparam, dtype=np.float32)
        check(lib().inmsom_bt_cnn_get_params(self._h, _f(p)), "inmsom_bt_cnn_get_params")
        return p

    def grad(self) -> np.ndarray:
        g = np.zeros(self.nparam, dtype=np.float32)
        check(lib().inmsom_bt_cnn_get_grad(self._h, _f(g)), "inmsom_bt_cnn_get_grad")
        return g

    # ---------------------------------------------------------------- batch
    def set_batch(self, x: np.ndarray, target: Optional[np.ndarray] = None,
                  mask: Optional[np.ndarray] = None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.shape != (self.batch, self.ny, self.nx, CIN0):
            raise ValueError(f"x must be {(self.batch, self.ny, self.nx, CIN0)}, got {x.shape}")
        t = None if target is None else np.ascontiguousarray(target, dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        if t is not None and t.shape != (self.batch, self.ny, self.nx, NOUT):
            raise ValueError("target shape")
        if m is not None and m.shape != (self.batch, self.ny, self.nx):
            raise ValueError("mask shape")
        self._batch_keep = (x, t, m)
        check(lib().inmsom_bt_cnn_set_batch(self._h, _f(x), None if t is None else _f(t),
                                            None if m is None else m.ctypes.data_as(_capi._u8p)),
              "inmsom_bt_cnn_set_batch")

    def forward(self) -> np.ndarray:
        out = np.zeros((self.batch, self.ny, self.nx, NOUT), dtype=np.float32)
        check(lib().inmsom_bt_cnn_forward(self._h, _f(out)), "inmsom_bt_cnn_forward")
        return out

    def activation(self, layer: int) -> np.ndarray:
        out = np.zeros((self.batch, self.ny, self.nx, CH), dtype=np.float32)
        check(lib().inmsom_bt_cnn_get_activation(self._h, layer, _f(out)), "inmsom_bt_cnn_get_activation")
        return out

    def keep_backprop(self, on: bool = True):
        """Keep dL/dz of every layer of the next backward passes."""
        check(lib().inmsom_bt_cnn_keep_backprop(self._h, int(on)), "inmsom_bt_cnn_keep_backprop")

    def backprop(self, layer: int) -> np.ndarray:
        """dL/dz of layer 0..3 of the last backward pass, [B, H, W, 32]."""
        out = np.empty((self.batch, self.ny, self.nx, 32), dtype=np.float32)
        check(lib().inmsom_bt_cnn_get_backprop(self._h, int(layer), _f(out)), "inmsom_bt_cnn_get_backprop")
        return out

    def loss_grad(self) -> float:
        loss = C.c_double()
        check(lib().inmsom_bt_cnn_loss_grad(self._h, C.byref(loss)), "inmsom_bt_cnn_loss_grad")
        return loss.value

    # ---------------------------------------------------------------- training set
    def dataset_alloc(self, nsamples: int):
        check(lib().inmsom_bt_cnn_dataset_alloc(self._h, nsamples), "inmsom_bt_cnn_dataset_alloc")
        self.nsamples = nsamples

    def dataset_set(self, i: int, x: np.ndarray, target: np.ndarray, mask: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if x.shape != (self.ny, self.nx, CIN0) or t.shape != (self.ny, self.nx, NOUT) or \
                m.shape != (self.ny, self.nx):
            raise ValueError("sample shapes")
        check(lib().inmsom_bt_cnn_dataset_set(self._h, i, _f(x), _f(t), m.ctypes.data_as(_capi._u8p)),
              "inmsom_bt_cnn_dataset_set")

    def train_step(self, idx: Sequence[int], lr: float = 1e-3) -> float:
        a = np.ascontiguousarray(idx, dtype=np.int32)
        if a.size != self.batch:
            raise ValueError(f"need {self.batch} sample indices")
        loss = C.c_double()
        check(lib().inmsom_bt_cnn_train_step(self._h, a.ctypes.data_as(_capi._ip), float(lr), C.byref(loss)),
              "inmsom_bt_cnn_train_step")
        return loss.value

    def eval(self, idx: Sequence[int]) -> float:
        """Loss of device samples ``idx`` (no gradient, no update)."""
        a = np.ascontiguousarray(idx, dtype=np.int32)
        if a.size != self.batch:
            raise ValueError(f"need {self.batch} sample indices")
        loss = C.c_double()
        check(lib().inmsom_bt_cnn_eval(self._h, a.ctypes.data_as(_capi._ip), C.byref(loss)), "inmsom_bt_cnn_eval")
        return loss.value

    def synchronize(self):
        check(lib().inmsom_bt_cnn_synchronize(self._h), "inmsom_bt_cnn_synchronize")
