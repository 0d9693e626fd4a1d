"""Learned propagator (mavi_mlp_* and mavi_learned_* in include/mavi_hip.h).

No counterpart in the reference: an addition beside the plain FP64 path,
which stays its fallback and the yardstick of its error.

  MLP           the network (BF16 weights, FP32 accumulation on the matrix
                cores): inference, loss + gradients, a device-resident data
                set and Adam training
  Learned       the gated learned step on a rings system, feature export and
                training-data collection from plain trajectories
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import _DP, _I64P, _P, check, load

N_IN, N_HID, N_OUT, N_PARAMS = 160, 128, 2, 53890
NOALLOC_ROWS = 8192  # MAVI_MLP_NOALLOC_ROWS

_U16P = C.POINTER(C.c_uint16)
_FP = C.POINTER(C.c_float)

SIGNATURES = {
    "mavi_mlp_create": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "mavi_mlp_destroy": (C.c_int, [_P]),
    "mavi_mlp_num_params": (C.c_int, []),
    "mavi_mlp_get_params": (C.c_int, [_P, _FP]),
    "mavi_mlp_set_params": (C.c_int, [_P, _FP]),
    "mavi_mlp_forward": (C.c_int, [_P, C.c_int64, _U16P, _FP, C.c_int32]),
    "mavi_mlp_forward_noalloc": (C.c_int, [_P, C.c_int64, _U16P, _FP]),
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

    def forward_noalloc(self, x_bits) -> np.ndarray:
        """forward() through mavi_mlp_forward_noalloc: the fused kernel with
        the library's fixed staging array (at most NOALLOC_ROWS rows)."""
        x, px = _u16(x_bits)
        m = x.size // N_IN
        y = np.zeros((m, N_OUT), dtype=np.float32)
        self._call("mavi_mlp_forward_noalloc", C.c_int64(m), px, y.ctypes.data_as(_FP))
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


class LearnedDesc(C.Structure):
    _fields_ = [
        ("k", C.c_int32),
        ("reserved", C.c_int32),
        ("r_feat", C.c_double),
        ("target_scale", C.c_double),
        ("d_min", C.c_double),
        ("max_disp", C.c_double),
        ("e_tol", C.c_double),
        ("seed", C.c_uint64),
    ]


SIGNATURES.update({
    "mavi_learned_create": (C.c_int, [_P, _P, C.POINTER(LearnedDesc), C.POINTER(_P)]),
    "mavi_learned_destroy": (C.c_int, [_P]),
    "mavi_learned_features": (C.c_int, [_P, _U16P, _DP]),
    "mavi_learned_collect": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_uint64]),
    "mavi_learned_run": (C.c_int, [_P, C.c_int64, C.c_int32, _I64P]),
    "mavi_learned_stats": (C.c_int, [_P, _I64P]),
})


class Learned:
    """Learned propagator bound to a rings system (a RingsSystem's device
    system) and optionally a network.  ``run(n)`` advances n learned steps,
    each standing for k Rings step! calls; rejected steps are replaced by k
    plain FP64 steps."""

    def __init__(self, dev: _lib.DeviceSystem, mlp: MLP | None = None, k: int = 10, r_feat: float = 1.25,
                 target_scale: float = 1e-3, d_min: float = 0.0, max_disp: float = 0.0, e_tol: float = 0.0,
                 seed: int = 0):
        lib = _lib_bound()
        self.dev, self.mlp, self.k = dev, mlp, int(k)
        self.desc = LearnedDesc(k=int(k), reserved=0, r_feat=r_feat, target_scale=target_scale, d_min=d_min,
                                max_disp=max_disp, e_tol=e_tol, seed=int(seed))
        h = _P()
        check(lib.mavi_learned_create(dev.handle, mlp.handle if mlp else None, C.byref(self.desc), C.byref(h)),
              dev.ctx.handle)
        self.handle = h
        self.accepted = 0
        self.rejected = 0

    def _call(self, name, *args):
        check(getattr(_lib_bound(), name)(self.handle, *args), self.dev.ctx.handle)

    def features(self):
        """(x bits [n][160], explicit k dt displacement [n][2]) at the current state."""
        n = self.dev.n
        x = np.zeros((n, N_IN), dtype=np.uint16)
        e = np.zeros((n, 2))
        self._call("mavi_learned_features", x.ctypes.data_as(_U16P), e.ctypes.data_as(_DP))
        return x, e

    def collect(self, nsnap: int, stride: int = 0, rows_per_snap: int = 1 << 30, seed: int = 0):
        self._call("mavi_learned_collect", int(nsnap), int(stride), int(min(rows_per_snap, 2**31 - 1)),
                   C.c_uint64(seed))

    def run(self, nsteps: int, use_net: bool = True) -> int:
        acc = C.c_int64()
        self._call("mavi_learned_run", C.c_int64(int(nsteps)), 1 if use_net else 0, C.byref(acc))
        self.accepted += acc.value
        self.rejected += int(nsteps) - acc.value
        return acc.value

    def stats(self) -> dict:
        a = np.zeros(7, dtype=np.int64)
        self._call("mavi_learned_stats", a.ctypes.data_as(_I64P))
        keys = ("accepted", "rejected", "displacement", "overlap", "ring_energy", "neighbours", "nonfinite")
        return {k: int(v) for k, v in zip(keys, a)}

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib_bound().mavi_learned_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
