"""Learned propagator (mavi_mlp_* and mavi_learned_* in include/mavi_hip.h).

No counterpart in the reference: an addition beside the plain FP64 path,
which stays its fallback and the judge of its error.

  MLP           the network (BF16 weights, FP32 accumulation on the matrix
                cores): inference, loss + gradients, a device-resident data
                set and Adam training
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import _DP, _I64P, _P, check, load

N_IN, N_HID, N_OUT, N_PARAMS = 160, 128, 2, 53890

_U16P = C.POINTER(C.c_uint16)
_FP = C.POINTER(C.c_float)

SIGNATURES = {
    "mavi_mlp_create": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "mavi_mlp_destroy": (C.c_int, [_P]),
    "mavi_mlp_num_params": (C.c_int, []),
    "mavi_mlp_get_params": (C.c_int, [_P, _FP]),
    "mavi_mlp_set_params": (C.c_int, [_P, _FP]),
    "mavi_mlp_forward": (C.c_int, [_P, C.c_int64, _U16P, _FP, C.c_int32]),
    "mavi_mlp_loss_grad": (C.c_int, [_P, C.c_int64, _U16P, _FP, _DP, _FP]),
    "mavi_mlp_dataset_clear": (C.c_int, [_P]),
    "mavi_mlp_dataset_append": (C.c_int, [_P, C.c_int64, _U16P, _FP]),
    "mavi_mlp_dataset_size": (C.c_int, [_P, _I64P]),
    "mavi_mlp_dataset_get": (C.c_int, [_P, C.c_int64, C.c_int64, _U16P, _FP]),
    "mavi_mlp_train": (C.c_int, [_P, C.c_int64, C.c_int32, C.c_float, C.c_uint64, _FP]),
    "mavi_mlp_gemm_check": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _U16P, _U16P,
                                      _FP]),
    "mavi_mlp_time_forward": (C.c_int, [_P, C.c_int64, C.c_int32, _DP]),
}

_bound = False


def _lib_bound():
    global _bound
    lib = load()
    if not _bound:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return lib


def _u16(a):
    a = np.ascontiguousarray(a, dtype=np.uint16)
    return a, a.ctypes.data_as(_U16P)


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_FP)


def gemm_check(ctx: _lib.Context, a_bits, b_bits, m, n, k, ta, tb):
    """Device BF16 GEMM (test hook of the matrix-core tiles)."""
    a, pa = _u16(a_bits)
    b, pb = _u16(b_bits)
    c = np.zeros((m, n), dtype=np.float32)
    check(_lib_bound().mavi_mlp_gemm_check(ctx.handle, m, n, k, int(ta), int(tb), pa, pb, c.ctypes.data_as(_FP)),
          ctx.handle)
    return c


class MLP:
    """Owner of one mavi_mlp handle."""

    def __init__(self, ctx: _lib.Context | None = None, seed: int = 0, device: int = 0):
        lib = _lib_bound()
        self.ctx = ctx or _lib.Context.for_device(device)
        h = _P()
        check(lib.mavi_mlp_create(self.ctx.handle, C.c_uint64(seed), C.byref(h)), self.ctx.handle)
        self.handle = h

    def _call(self, name, *args):
        check(getattr(_lib_bound(), name)(self.handle, *args), self.ctx.handle)

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(N_PARAMS, dtype=np.float32)
        self._call("mavi_mlp_get_params", p.ctypes.data_as(_FP))
        return p

    @params.setter
    def params(self, p):
        p, pp = _f32(p)
        assert p.size == N_PARAMS
        self._call("mavi_mlp_set_params", pp)

    def forward(self, x_bits, fused: bool = True) -> np.ndarray:
        x, px = _u16(x_bits)
        m = x.size // N_IN
        y = np.zeros((m, N_OUT), dtype=np.float32)
        self._call("mavi_mlp_forward", C.c_int64(m), px, y.ctypes.data_as(_FP), 1 if fused else 0)
        return y

    def loss_grad(self, x_bits, t):
        x, px = _u16(x_bits)
        t, pt = _f32(t)
        m = x.size // N_IN
        loss = C.c_double()
        g = np.zeros(N_PARAMS, dtype=np.float32)
        self._call("mavi_mlp_loss_grad", C.c_int64(m), px, pt, C.byref(loss), g.ctypes.data_as(_FP))
        return loss.value, g

    def dataset_clear(self):
        self._call("mavi_mlp_dataset_clear")

    def dataset_append(self, x_bits, t):
        x, px = _u16(x_bits)
        t, pt = _f32(t)
        self._call("mavi_mlp_dataset_append", C.c_int64(x.size // N_IN), px, pt)

    def dataset_size(self) -> int:
        n = C.c_int64()
        check(_lib_bound().mavi_mlp_dataset_size(self.handle, C.byref(n)), self.ctx.handle)
        return n.value

    def dataset_get(self, first: int, n: int):
        x = np.zeros((n, N_IN), dtype=np.uint16)
        t = np.zeros((n, N_OUT), dtype=np.float32)
        self._call("mavi_mlp_dataset_get", C.c_int64(first), C.c_int64(n), x.ctypes.data_as(_U16P),
                   t.ctypes.data_as(_FP))
        return x, t

    def train(self, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(int(steps), 1), dtype=np.float32)
        self._call("mavi_mlp_train", C.c_int64(int(steps)), C.c_int32(int(batch)), C.c_float(lr),
                   C.c_uint64(seed), losses.ctypes.data_as(_FP))
        return losses[:steps]

    def time_forward(self, m: int, reps: int = 20) -> float:
        us = C.c_double()
        check(_lib_bound().mavi_mlp_time_forward(self.handle, C.c_int64(m), C.c_int32(reps), C.byref(us)),
              self.ctx.handle)
        return us.value

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib_bound().mavi_mlp_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
