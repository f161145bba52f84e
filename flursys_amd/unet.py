"""Learned warm start: the U-Net of the learned FCG over the C ABI.

The reference has no learned component; this is the host-side handle of the
gfx950 U-Net (BF16 MFMA convolutions, native training) that the 3D cavity
solver uses with ``preconditioner=PRECOND_LEARNED``: the network predicts
the pressure correction from the initial residual restricted to its grid,
and multigrid FCG continues from that guess (a prediction that does not
lower the residual is undone, and the plain FP64 PCG fallback still holds).

Arrays are float32 [batch][n^3] (x fastest).  ``diag`` is the Laplacian's
diagonal 2 (1/hx^2 + 1/hy^2 + 1/hz^2) at the network's resolution; the
solver's ``coarse_pair`` returns it with each training pair.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import UNetConfig, check, load, ptr

# (c1, c2, cout, level, relu): c1 input channels at the layer's level, c2
# channels of the half-resolution input upsampled into the concatenation.
LAYERS = [
    (16, 0, 16, 0, 1), (16, 0, 16, 0, 1), (16, 0, 32, 1, 1), (32, 0, 32, 1, 1),
    (32, 0, 48, 2, 1), (48, 0, 48, 2, 1), (48, 0, 64, 3, 1), (64, 0, 64, 3, 1),
    (48, 64, 48, 2, 1), (48, 0, 48, 2, 1), (32, 48, 32, 1, 1), (32, 0, 32, 1, 1),
    (16, 32, 16, 0, 1), (16, 0, 16, 0, 1), (16, 0, 16, 0, 0),
]


def layer_table():
    """[(c1, c2, cin, cout, kpad, level, relu, offset)] in parameter order."""
    out, off = [], 0
    for c1, c2, cout, level, relu in LAYERS:
        cin = c1 + c2
        kpad = (27 * cin + 31) // 32 * 32
        out.append((c1, c2, cin, cout, kpad, level, relu, off))
        off += cout * kpad
    return out


def params_to_layers(params: np.ndarray):
    """Split the flat parameter vector into per-layer [cout][cin][3][3][3]
    arrays (kernel axes z, y, x)."""
    layers = []
    for c1, c2, cin, cout, kpad, level, relu, off in layer_table():
        w = params[off:off + cout * kpad].reshape(cout, kpad)[:, :27 * cin]
        layers.append(w.reshape(cout, 27, cin).transpose(0, 2, 1).reshape(cout, cin, 3, 3, 3))
    return layers


class UNet:
    """fl_unet: weights, Adam state and activations live on the device."""

    def __init__(self, n: int = 64, max_batch: int = 8, seed: int = 1,
                 ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.default_context()
        self.n = n
        self.max_batch = max_batch
        lib = load()
        h = C.c_void_p()
        cfg = UNetConfig(n, max_batch, seed)
        check(lib.fl_unet_create(self.ctx.handle, C.byref(cfg), C.byref(h)), self.ctx.handle)
        self.handle = h
        self.num_params = int(lib.fl_unet_num_params(h))

    def close(self):
        if getattr(self, "handle", None):
            load().fl_unet_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _batch(self, a) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.ndim == 3:
            a = a[None]
        a = a.reshape(a.shape[0], -1)
        if a.shape[1] != self.n ** 3:
            raise ValueError(f"expected samples of {self.n}^3 cells, got {a.shape[1]}")
        return np.ascontiguousarray(a)

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        check(load().fl_unet_get_params(self.handle, ptr(out)), self.ctx.handle)
        return out

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.size}")
        check(load().fl_unet_set_params(self.handle, ptr(p)), self.ctx.handle)

    @property
    def gain(self) -> float:
        """Output gain G (the last parameter): predictions are
        (o - mean o) rms(b) G / diag."""
        return float(self.params()[-1])

    def set_gain(self, gain: float) -> None:
        """Sets G (keeps the weights; resets the Adam state)."""
        p = self.params()
        p[-1] = gain
        self.set_params(p)

    def save(self, path: str) -> None:
        np.save(path, self.params(), allow_pickle=False)

    def load(self, path: str) -> None:
        self.set_params(np.load(path, allow_pickle=False))

    def train_step(self, rhs, p, diag: float, lr: float) -> float:
        x, t = self._batch(rhs), self._batch(p)
        if x.shape != t.shape:
            raise ValueError("rhs and p batches differ in shape")
        loss = C.c_float()
        check(load().fl_unet_train_step(self.handle, x.shape[0], ptr(x), ptr(t), float(diag),
                                        float(lr), C.byref(loss)), self.ctx.handle)
        return float(loss.value)

    def infer(self, rhs, diag: float) -> np.ndarray:
        x = self._batch(rhs)
        out = np.empty_like(x)
        check(load().fl_unet_infer(self.handle, x.shape[0], ptr(x), float(diag), ptr(out)),
              self.ctx.handle)
        return out

    def forward_raw(self, rhs) -> np.ndarray:
        """Raw 16-channel output [batch][n^3][16] of the normalised input."""
        x = self._batch(rhs)
        out = np.empty((x.shape[0], self.n ** 3, 16), np.float32)
        check(load().fl_unet_forward_raw(self.handle, x.shape[0], ptr(x), ptr(out)),
              self.ctx.handle)
        return out

    def grads(self, rhs, p, diag: float):
        """(parameter gradient, loss) of the training loss, no update."""
        x, t = self._batch(rhs), self._batch(p)
        g = np.empty(self.num_params, np.float32)
        loss = C.c_float()
        check(load().fl_unet_grads(self.handle, x.shape[0], ptr(x), ptr(t), float(diag), ptr(g),
                                   C.byref(loss)), self.ctx.handle)
        return g, float(loss.value)


def probe_mfma_bf16(a, b, ctx: _lib.Context | None = None) -> np.ndarray:
    """d = a (16x32) @ b (32x16) through the MFMA operand layout."""
    ctx = ctx or _lib.default_context()
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    assert a.shape == (16, 32) and b.shape == (32, 16)
    d = np.empty((16, 16), np.float32)
    check(load().fl_probe_mfma_bf16(ctx.handle, ptr(a), ptr(b), ptr(d)), ctx.handle)
    return d
