"""Replica batches on the GPU (mavi_batch_* in include/mavi_hip.h).

The device side of run_experiment_batch (src/experiments.jl:246-300): there
every replica is its own System stepped by its own step function, one Julia
thread per replica.  Here the replicas share one set of device arrays and
every phase of a step is one launch over all replicas of a kind; multi-step
runs replay a captured graph of ``graph_steps`` steps.  Each replica's
trajectory is bit-identical to the same system stepped alone.

    b = Batch()
    i = b.add(state, space_cfg, dynamic_cfg, int_cfg)          # plain or rings
    b.finalize(); b.upload(i, state.pos, state.vel); b.run(100)
    pos, vel, pol = b.download(i)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import _DP, _I64P, _P, SysDesc, check, dptr, load

SIGNATURES = {
    "mavi_batch_create": (C.c_int, [_P, C.POINTER(_P)]),
    "mavi_batch_destroy": (C.c_int, [_P]),
    "mavi_batch_add_replica": (C.c_int, [_P, C.POINTER(SysDesc), C.POINTER(C.c_int32)]),
    "mavi_batch_finalize": (C.c_int, [_P]),
    "mavi_batch_upload": (C.c_int, [_P, C.c_int32, _DP, _DP, _DP]),
    "mavi_batch_download": (C.c_int, [_P, C.c_int32, _DP, _DP, _DP]),
    "mavi_batch_get_forces": (C.c_int, [_P, C.c_int32, _DP]),
    "mavi_batch_run": (C.c_int, [_P, C.c_int64]),
    "mavi_batch_set_option": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "mavi_batch_info": (C.c_int, [_P, _I64P]),
    "mavi_batch_prof_enable": (C.c_int, [_P, C.c_int]),
    "mavi_batch_prof_get": (C.c_int, [_P, C.c_int, _DP, _I64P]),
    "mavi_batch_autotune": (C.c_int, [_P, C.c_int32, _DP, C.POINTER(C.c_int32)]),
    "mavi_batch_noise_get_counter": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "mavi_batch_noise_set_counter": (C.c_int, [_P, C.c_uint64]),
}

MAVI_BOPT_WG_PLAIN = 0
MAVI_BOPT_WG_RINGS = 1
MAVI_BOPT_LIST_PLAIN = 2
MAVI_BOPT_LIST_RINGS = 3
MAVI_BOPT_GRAPH_STEPS = 4
MAVI_BOPT_GRAPHS = 5

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


class Batch:
    """Owner of one mavi_batch handle."""

    def __init__(self, ctx: _lib.Context | None = None, device: int = 0):
        lib = _lib_bound()
        self.ctx = ctx or _lib.Context.for_device(device)
        h = _P()
        check(lib.mavi_batch_create(self.ctx.handle, C.byref(h)), self.ctx.handle)
        self.handle = h
        self.replicas: list[dict] = []
        self.finalized = False
        if not _lib.USE_GRAPHS:
            self.set_option(MAVI_BOPT_GRAPHS, 0)

    def _call(self, name, *args):
        check(getattr(_lib_bound(), name)(self.handle, *args), self.ctx.handle)

    def add(self, state, space_cfg, dynamic_cfg, int_cfg, seed: int = 0) -> int:
        """Append a replica; ``state`` is a SecondLawState (LJ / HarmTrunc) or a
        RingsState (rings, device noise stream ``seed``)."""
        from . import rings as _rings
        from . import systems as _systems

        if isinstance(state, _rings.RingsState):
            d, _, tables = _rings.describe(state, space_cfg, dynamic_cfg, int_cfg, seed)
            rings = True
        else:
            d, _ = _systems.describe(state, space_cfg, dynamic_cfg, int_cfg)
            tables = ()
            rings = False
        i = self.add_desc(d)
        del tables  # read during the call only
        return i

    def add_desc(self, d: SysDesc) -> int:
        """Append a replica from a raw descriptor (host tables read during the call)."""
        idx = C.c_int32()
        self._call("mavi_batch_add_replica", C.byref(d), C.byref(idx))
        self.replicas.append({"rings": int(d.potential) == _lib.MAVI_POT_RINGS, "n": int(d.n),
                              "nrings": int(d.nrings)})
        return idx.value

    def finalize(self):
        self._call("mavi_batch_finalize")
        self.finalized = True

    def upload(self, i: int, pos, vel=None, pol=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        pol = None if pol is None else np.ascontiguousarray(pol, dtype=np.float64)
        self._call("mavi_batch_upload", C.c_int32(i), dptr(pos), dptr(vel), dptr(pol))

    def download(self, i: int):
        """(pos, vel, pol) of replica i; vel is None for rings, pol None for plain."""
        r = self.replicas[i]
        pos = np.zeros((r["n"], 2))
        vel = None if r["rings"] else np.zeros((r["n"], 2))
        pol = np.zeros(r["nrings"]) if r["rings"] else None
        self._call("mavi_batch_download", C.c_int32(i), dptr(pos), dptr(vel), dptr(pol))
        return pos, vel, pol

    def get_forces(self, i: int):
        f = np.zeros((self.replicas[i]["n"], 2))
        self._call("mavi_batch_get_forces", C.c_int32(i), dptr(f))
        return f

    def run(self, nsteps: int):
        self._call("mavi_batch_run", C.c_int64(int(nsteps)))

    def noise_counter(self) -> int:
        c = C.c_uint64()
        self._call("mavi_batch_noise_get_counter", C.byref(c))
        return int(c.value)

    def set_noise_counter(self, counter: int):
        self._call("mavi_batch_noise_set_counter", C.c_uint64(int(counter)))

    def set_option(self, option: int, value: int):
        self._call("mavi_batch_set_option", C.c_int32(int(option)), C.c_int32(int(value)))

    def info(self) -> dict:
        a = np.zeros(6, dtype=np.int64)
        self._call("mavi_batch_info", a.ctypes.data_as(_I64P))
        keys = ("replicas", "slots", "particles", "padded_slots", "chunks", "rings")
        return {k: int(v) for k, v in zip(keys, a)}

    def prof_enable(self, on: bool = True):
        self._call("mavi_batch_prof_enable", 1 if on else 0)

    def prof_get(self, which: int):
        ms = C.c_double()
        n = C.c_int64()
        check(_lib_bound().mavi_batch_prof_get(self.handle, which, C.byref(ms), C.byref(n)), self.ctx.handle)
        return ms.value, n.value

    def autotune(self, trial_steps: int = 10):
        best = C.c_double()
        shape = (C.c_int32 * 4)()
        self._call("mavi_batch_autotune", C.c_int32(int(trial_steps)), C.byref(best), shape)
        return best.value, {"wg_plain": shape[0], "list_plain": shape[1], "wg_rings": shape[2],
                            "list_rings": shape[3]}

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib_bound().mavi_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
