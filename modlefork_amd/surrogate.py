"""Learned surrogate of the simulation over the C-ABI (mgpu_sg_*).

The network maps a window's barrier track (position/direction/occupancy channels) and two
scalars (LEF density, processivity) to the window's log1p contact band; forward, backward and
AdamW run as the library's HIP kernels (BF16 on MFMA, FP32 accumulation). It has no counterpart
in the reference. Arrays are fp32 numpy: x [n][in_len][in_ch], y [n][ndiag][out_len].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

FP = C.POINTER(C.c_float)


def _f(a):
    return a.ctypes.data_as(FP)


def default_config(**over) -> L.MgpuSgConfig:
    c = L.MgpuSgConfig()
    L.check(L.lib().mgpu_sg_config_default(C.byref(c)))
    for k, v in over.items():
        setattr(c, k, v)
    return c


class Surrogate:
    def __init__(self, ctx, cfg: L.MgpuSgConfig | None = None):
        self.ctx = ctx
        self.cfg = cfg or default_config()
        self.ptr = C.c_void_p()
        rc = L.lib().mgpu_sg_create(ctx.ptr, C.byref(self.cfg), C.byref(self.ptr))
        if rc != 0:
            raise RuntimeError(f"mgpu_sg_create failed ({rc}): "
                               f"{L.lib().mgpu_sg_last_error(None).decode()}")
        self.nparams = int(L.lib().mgpu_sg_num_params(self.ptr))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"surrogate call failed ({rc}): "
                               f"{L.lib().mgpu_sg_last_error(self.ptr).decode()}")

    @property
    def out_shape(self):
        return (int(self.cfg.ndiag), int(self.cfg.out_len))

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        self._check(L.lib().mgpu_sg_get_params(self.ptr, _f(out)))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        self._check(L.lib().mgpu_sg_set_params(self.ptr, _f(p)))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        n = x.shape[0]
        y = np.zeros((n,) + self.out_shape, np.float32)
        self._check(L.lib().mgpu_sg_forward(self.ptr, _f(x), n, _f(y)))
        return y

    def gradients(self, x: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float()
        self._check(L.lib().mgpu_sg_gradients(self.ptr, _f(x), _f(y), x.shape[0], C.byref(loss),
                                              _f(g)))
        return float(loss.value), g

    def adamw_step(self, g: np.ndarray, lr: float, weight_decay: float):
        g = np.ascontiguousarray(g, np.float32)
        self._check(L.lib().mgpu_sg_adamw_step(self.ptr, _f(g), lr, weight_decay))

    def set_data(self, x: np.ndarray, y: np.ndarray):
        self._x = np.ascontiguousarray(x, np.float32)
        self._y = np.ascontiguousarray(y, np.float32)
        self._check(L.lib().mgpu_sg_set_data(self.ptr, _f(self._x), _f(self._y), x.shape[0]))
        self._x = self._y = None

    def train_step(self, idx: np.ndarray, lr: float, weight_decay: float = 0.0):
        idx = np.ascontiguousarray(idx, np.uint32)
        loss, ms = C.c_float(), C.c_float()
        self._check(L.lib().mgpu_sg_train_step(self.ptr, idx.ctypes.data_as(L.U32P), len(idx), lr,
                                               weight_decay, C.byref(loss), C.byref(ms)))
        return float(loss.value), float(ms.value)

    def predict_timed(self, n: int, repeats: int) -> float:
        ms = C.c_float()
        self._check(L.lib().mgpu_sg_predict_timed(self.ptr, n, repeats, C.byref(ms)))
        return float(ms.value)

    def close(self):
        if self.ptr:
            L.lib().mgpu_sg_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mfma_probe(ctx, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """One 32 x 32 x 16 bf16 MFMA through the kernels' lane layout; a, b are bf16 bit patterns."""
    a = np.ascontiguousarray(a, np.uint16)
    b = np.ascontiguousarray(b, np.uint16)
    d = np.zeros((32, 32), np.float32)
    rc = L.lib().mgpu_debug_mfma(ctx.ptr, a.ctypes.data_as(C.POINTER(C.c_uint16)),
                                 b.ctypes.data_as(C.POINTER(C.c_uint16)), _f(d))
    L.check(rc, ctx.ptr)
    return d


def to_bf16_bits(x: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(x, np.float32).view(np.uint32)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32)


def bf16_round(x: np.ndarray) -> np.ndarray:
    return from_bf16_bits(to_bf16_bits(x))
