"""ctypes binding of libmavi_hip.so (the C ABI declared in include/mavi_hip.h).

The product path has no CPU fallback: if the library is missing or no GPU is
visible, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .build import library_path

MAVI_OK = 0
MAVI_E_CELL_OVERFLOW = 1
MAVI_E_OUT_OF_BOX = 2
MAVI_E_NONFINITE = 3
MAVI_E_HIP = 4
MAVI_E_BADARG = 5
MAVI_E_UNSUPPORTED = 6

MAVI_WALL_RIGID = 0
MAVI_WALL_PERIODIC = 1
MAVI_POT_LJ = 0
MAVI_POT_HARMTRUNC = 1
MAVI_POT_RINGS = 2
MAVI_POT_SZABO = 3
MAVI_POT_RTP = 4
(MAVI_WK_RIGID_RECT, MAVI_WK_PERIODIC_RECT, MAVI_WK_RIGID_CIRCLE, MAVI_WK_SLIPPERY_LINES, MAVI_WK_SLIPPERY_CIRCLE,
 MAVI_WK_POTENTIAL_CIRCLE, MAVI_WK_POTENTIAL_LINES) = range(7)
MAVI_WMODE = {"outside": 0, "inside": 1, "repulsion": 2}

MAVI_PROF_BIN = 0
MAVI_PROF_FORCE = 1
MAVI_PROF_RINGS = 2
MAVI_PROF_FORCE_REUSE = 3

_STATUS_NAMES = {
    MAVI_E_CELL_OVERFLOW: "E_CELL_OVERFLOW",
    MAVI_E_OUT_OF_BOX: "E_OUT_OF_BOX",
    MAVI_E_NONFINITE: "E_NONFINITE",
    MAVI_E_HIP: "E_HIP",
    MAVI_E_BADARG: "E_BADARG",
    MAVI_E_UNSUPPORTED: "E_UNSUPPORTED",
}


class MaviError(RuntimeError):
    """A non-zero status from libmavi_hip (the Julia glue raises `error(...)`)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class CellOverflowError(MaviError):
    pass


class OutOfBoxError(MaviError):
    pass


class UnsupportedOnGPU(MaviError):
    """Configuration outside the GPU path; the Julia integration keeps Mavi's
    own CPU methods for it (custom DynamicCfg, Szabo/RTP, other geometries)."""


class MaviWall(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("mode", C.c_int32),
        ("pot_kind", C.c_int32),
        ("nlines", C.c_int32),
        ("geom", C.c_double * 4),
        ("pot", C.c_double * 4),
        ("lines", C.POINTER(C.c_double)),
    ]


MAVI_SRC_SOURCE, MAVI_SRC_SINK = 0, 1
MAVI_GEOM_RECT, MAVI_GEOM_CIRCLE = 0, 1


class MaviRingSource(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("geom", C.c_int32),
        ("g", C.c_double * 4),
        ("nbox", C.c_int32),
        ("random_pol", C.c_int32),
        ("pad", C.c_double),
        ("pol", C.c_double),
        ("bbox", C.POINTER(C.c_double)),
        ("spawn", C.POINTER(C.c_double)),
        ("chunk_off", C.POINTER(C.c_int32)),
        ("chunk_ids", C.POINTER(C.c_int32)),
    ]


class SysDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("bottom_left", C.c_double * 2),
        ("size", C.c_double * 2),
        ("wall", C.c_int32),
        ("potential", C.c_int32),
        ("pot", C.c_double * 4),
        ("dt", C.c_double),
        ("cols", C.c_int32),
        ("rows", C.c_int32),
        ("nc", C.c_int32),
        ("reserved0", C.c_int32),
        ("radius", C.c_double),
        ("nmax", C.c_int32),
        ("nrings", C.c_int32),
        ("ntypes", C.c_int32),
        ("reserved1", C.c_int32),
        ("ring_types", C.POINTER(C.c_int32)),
        ("type_np", C.POINTER(C.c_int32)),
        ("type_par", C.POINTER(C.c_double)),
        ("interaction", C.POINTER(C.c_double)),
        ("seed", C.c_uint64),
        ("active", C.c_double * 4),
        ("nwalls", C.c_int32),
        ("reserved2", C.c_int32),
        ("walls", C.POINTER(MaviWall)),
    ]


_P = C.c_void_p
_DP = C.POINTER(C.c_double)
_I64P = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
SIGNATURES = {
    "mavi_abi_version": (C.c_int, []),
    "mavi_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "mavi_ctx_destroy": (C.c_int, [_P]),
    "mavi_last_error": (C.c_char_p, [_P]),
    "mavi_sys_create": (C.c_int, [_P, C.POINTER(SysDesc), C.POINTER(_P)]),
    "mavi_sys_destroy": (C.c_int, [_P]),
    "mavi_sys_upload": (C.c_int, [_P, _DP, _DP, _DP]),
    "mavi_sys_download": (C.c_int, [_P, _DP, _DP, _DP]),
    "mavi_bin": (C.c_int, [_P, _I64P, _I64P]),
    "mavi_forces": (C.c_int, [_P, _DP]),
    "mavi_get_forces": (C.c_int, [_P, _DP]),
    "mavi_step": (C.c_int, [_P, C.c_int64]),
    "mavi_rings_step": (C.c_int, [_P, C.c_int64, _DP]),
    "mavi_sp_step": (C.c_int, [_P, C.c_int64, _DP]),
    "mavi_rings_info": (C.c_int, [_P, _DP, _DP, _DP]),
    "mavi_rings_neighbors_cfg": (C.c_int, [_P, C.c_int, C.c_double, C.c_int]),
    "mavi_rings_neighbors_get": (C.c_int, [_P, _I64P, _I64P]),
    "mavi_sys_noise_get": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "mavi_sys_noise_set": (C.c_int, [_P, C.c_uint64, C.c_uint64]),
    "mavi_rings_var_cfg": (C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32, C.POINTER(MaviRingSource)]),
    "mavi_rings_pol_draws": (C.c_int, [_P, C.POINTER(C.c_double), C.c_int64]),
    "mavi_rings_var_get": (C.c_int, [_P, C.POINTER(C.c_int32), _I64P, _I64P, _I64P, _I64P, _I64P]),
    "mavi_rings_pol_stream_set": (C.c_int, [_P, C.c_int64]),
    "mavi_rings_invasions_cfg": (C.c_int, [_P, C.c_int, C.c_int64, C.c_int32, C.c_int32, C.c_int32]),
    "mavi_rings_invasions_check": (C.c_int, [_P]),
    "mavi_rings_invasions_get": (C.c_int, [_P, _I64P, _I64P, C.c_int64]),
    "mavi_observables": (C.c_int, [_P, _DP, _DP]),
    "mavi_run_record": (C.c_int, [_P, C.c_int64, _DP, _DP]),
    "mavi_prof_enable": (C.c_int, [_P, C.c_int]),
    "mavi_prof_get": (C.c_int, [_P, C.c_int, _DP, _I64P]),
    "mavi_nlist_info": (C.c_int, [_P, _I64P]),
    "mavi_sys_set_tile": (C.c_int, [_P, C.c_int32]),
    "mavi_sys_set_option": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "mavi_autotune": (C.c_int, [_P, C.c_int32, _DP, C.POINTER(C.c_int32)]),
}

MAVI_OPT_TWO_PHASE = 0
MAVI_OPT_GRAPHS = 1
MAVI_OPT_LDS_TILES = 2
MAVI_OPT_REUSE_FORCES = 3
MAVI_OPT_NLIST = 4
MAVI_OPT_NLIST_SKIN = 5
MAVI_OPT_NLIST_BLOCK = 6

# Multi-step calls replay captured hipGraphs unless this is set to False
# before systems are created (kernel traces are then plain launches).
USE_GRAPHS = True

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libmavi_hip.so (built in-tree by build.build_library)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else library_path()
    if not p.exists():
        raise RuntimeError(
            f"libmavi_hip.so not found at {p}: build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def i64ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_I64P)


def check(status: int, ctx) -> None:
    if status == MAVI_OK:
        return
    lib = load()
    msg = lib.mavi_last_error(ctx).decode() if ctx else ""
    cls = {
        MAVI_E_CELL_OVERFLOW: CellOverflowError,
        MAVI_E_OUT_OF_BOX: OutOfBoxError,
        MAVI_E_UNSUPPORTED: UnsupportedOnGPU,
    }.get(status, MaviError)
    raise cls(status, msg)


class Context:
    """One HIP stream on one device (mavi_ctx)."""

    _default: "Context | None" = None

    def __init__(self, device: int = 0):
        lib = load()
        h = _P()
        st = lib.mavi_ctx_create(device, C.byref(h))
        if st != MAVI_OK:
            raise MaviError(st, f"mavi_ctx_create(device={device}) failed (is a GPU visible?)")
        self.handle = h
        self.device = device

    _per_device: dict = {}

    @classmethod
    def default(cls) -> "Context":
        return cls.for_device(int(os.environ.get("MAVI_HIP_DEVICE", "0")))

    @classmethod
    def for_device(cls, device: int) -> "Context":
        """One shared context (and stream) per device for this process."""
        if device not in cls._per_device:
            cls._per_device[device] = Context(device)
        return cls._per_device[device]

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().mavi_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class DeviceSystem:
    """Owner of one mavi_sys handle."""

    def __init__(self, ctx: Context, desc: SysDesc, keep=()):
        lib = load()
        self.ctx = ctx
        self._keep = keep  # arrays the descriptor points into, kept alive during create
        h = _P()
        check(lib.mavi_sys_create(ctx.handle, C.byref(desc), C.byref(h)), ctx.handle)
        self.handle = h
        self.n = int(desc.n)
        self.nrings = int(desc.nrings)
        self.cols = int(desc.cols)
        self.rows = int(desc.rows)
        self.nc = int(desc.nc)
        if not USE_GRAPHS:
            self.set_option(MAVI_OPT_GRAPHS, 0)

    def _call(self, name, *args):
        check(getattr(load(), name)(self.handle, *args), self.ctx.handle)

    def upload(self, pos, vel=None, pol=None):
        self._call("mavi_sys_upload", dptr(pos), dptr(vel), dptr(pol))

    def download(self, pos=None, vel=None, pol=None):
        self._call("mavi_sys_download", dptr(pos), dptr(vel), dptr(pol))

    def bin(self, export: bool = True):
        if not export:
            self._call("mavi_bin", None, None)
            return None, None
        cp = np.zeros((self.cols, self.rows, self.nc), dtype=np.int64)  # (nc,rows,cols) column-major
        counts = np.zeros((self.cols, self.rows), dtype=np.int64)
        self._call("mavi_bin", i64ptr(cp), i64ptr(counts))
        # present them with Julia's index order: chunk_particles[p, row, col]
        return cp.transpose(2, 1, 0), counts.transpose(1, 0)

    def forces(self):
        f = np.zeros((self.n, 2))
        self._call("mavi_forces", dptr(f))
        return f

    def get_forces(self):
        f = np.zeros((self.n, 2))
        self._call("mavi_get_forces", dptr(f))
        return f

    def step(self, nsteps: int):
        self._call("mavi_step", C.c_int64(int(nsteps)))

    def sp_step(self, nsteps: int, noise=None):
        """Self-propelled steps; noise: (nsteps, n) normals (Szabo) or
        (nsteps, n, 2) uniforms (RTP), or None for the device stream."""
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.float64)
        self._call("mavi_sp_step", C.c_int64(int(nsteps)), dptr(noise))

    def rings_step(self, nsteps: int, noise=None):
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.float64)
            assert noise.size >= nsteps * self.nrings
        self._call("mavi_rings_step", C.c_int64(int(nsteps)), dptr(noise))

    def rings_neighbors_cfg(self, enable: bool, tol: float = 1.1, only_rings: bool = False):
        self._call("mavi_rings_neighbors_cfg", 1 if enable else 0, C.c_double(tol), 1 if only_rings else 0)

    def rings_neighbors(self):
        """(counts (n,), lists (n, 15)): 1-based ids ascending, -1 padded."""
        counts = np.zeros(self.n, dtype=np.int64)
        lists = np.zeros((self.n, 15), dtype=np.int64)
        self._call("mavi_rings_neighbors_get", i64ptr(counts), i64ptr(lists))
        return counts, lists

    def noise_state(self):
        """(seed, step counter) of the device noise stream."""
        seed, counter = C.c_uint64(), C.c_uint64()
        self._call("mavi_sys_noise_get", C.byref(seed), C.byref(counter))
        return int(seed.value), int(counter.value)

    def set_noise_state(self, seed: int, counter: int):
        self._call("mavi_sys_noise_set", C.c_uint64(int(seed)), C.c_uint64(int(counter)))

    def rings_var_cfg(self, mask, records):
        """records: list of MaviRingSource (their arrays kept alive by the caller)."""
        m = np.ascontiguousarray(mask, dtype=np.int32)
        arr = (MaviRingSource * max(len(records), 1))(*records)
        self._call("mavi_rings_var_cfg", m.ctypes.data_as(C.POINTER(C.c_int32)), len(records),
                   C.cast(arr, C.POINTER(MaviRingSource)))

    def rings_pol_draws(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._call("mavi_rings_pol_draws", dptr(u), C.c_int64(len(u)))

    def rings_var_state(self):
        """(mask, ids, uids, num_active, pool draws used, built-in stream
        position) of a variable population."""
        nr = self.nrings
        mask = np.zeros(nr, dtype=np.int32)
        ids = np.zeros(nr, dtype=np.int64)
        uids = np.zeros(nr, dtype=np.int64)
        na = np.zeros(1, dtype=np.int64)
        used = np.zeros(1, dtype=np.int64)
        stream = np.zeros(1, dtype=np.int64)
        self._call("mavi_rings_var_get", mask.ctypes.data_as(C.POINTER(C.c_int32)), i64ptr(ids), i64ptr(uids),
                   i64ptr(na), i64ptr(used), i64ptr(stream))
        return mask, ids, uids, int(na[0]), int(used[0]), int(stream[0])

    def rings_pol_stream_set(self, position: int):
        self._call("mavi_rings_pol_stream_set", C.c_int64(int(position)))

    def rings_invasions_cfg(self, scheduled: bool, steps_to_update: int = 0, r_cols: int = 0, r_rows: int = 0,
                            r_nc: int = 0):
        self._call("mavi_rings_invasions_cfg", 1 if scheduled else 0, C.c_int64(int(steps_to_update)),
                   int(r_cols), int(r_rows), int(r_nc))

    def rings_invasions_check(self):
        self._call("mavi_rings_invasions_check")

    def rings_invasions(self) -> np.ndarray:
        """(k, 3) int64 records (invasor, invaded, p_id) of the last check,
        1-based, sorted."""
        count = np.zeros(1, dtype=np.int64)
        self._call("mavi_rings_invasions_get", i64ptr(count), None, C.c_int64(0))
        rec = np.zeros((int(count[0]), 3), dtype=np.int64)
        if len(rec):
            self._call("mavi_rings_invasions_get", i64ptr(count), i64ptr(rec), C.c_int64(len(rec)))
        return rec

    def rings_info(self):
        areas = np.zeros(self.nrings)
        cms = np.zeros((self.nrings, 2))
        cpos = np.zeros((self.n, 2))
        self._call("mavi_rings_info", dptr(areas), dptr(cms), dptr(cpos))
        return areas, cms, cpos

    def observables(self):
        ke = C.c_double()
        pe = C.c_double()
        check(load().mavi_observables(self.handle, C.byref(ke), C.byref(pe)), self.ctx.handle)
        return ke.value, pe.value

    def run_record(self, nsteps: int, with_pe: bool = True):
        ke = np.zeros(nsteps)
        pe = np.zeros(nsteps) if with_pe else None
        self._call("mavi_run_record", C.c_int64(int(nsteps)), dptr(ke), dptr(pe))
        return ke, pe

    def prof_enable(self, on: bool = True):
        self._call("mavi_prof_enable", 1 if on else 0)

    def nlist_info(self) -> dict:
        """Neighbour lists: eligible, running, rebuilds, rolled-back blocks."""
        a = np.zeros(4, dtype=np.int64)
        self._call("mavi_nlist_info", a.ctypes.data_as(_I64P))
        return {"eligible": bool(a[0]), "active": bool(a[1]), "rebuilds": int(a[2]), "rollbacks": int(a[3])}

    def prof_get(self, which: int):
        ms = C.c_double()
        n = C.c_int64()
        check(load().mavi_prof_get(self.handle, which, C.byref(ms), C.byref(n)), self.ctx.handle)
        return ms.value, n.value

    def set_tile(self, wg: int):
        self._call("mavi_sys_set_tile", C.c_int32(int(wg)))

    def autotune(self, trial_steps: int = 20):
        """Pick the fastest pair-kernel shape for this system (state restored)."""
        best = C.c_double()
        shape = (C.c_int32 * 3)()
        self._call("mavi_autotune", C.c_int32(int(trial_steps)), C.byref(best), shape)
        return best.value, {"workgroup": shape[0], "lds_tiles": shape[1], "two_phase": shape[2]}

    def set_option(self, option: int, value: int):
        self._call("mavi_sys_set_option", C.c_int32(int(option)), C.c_int32(int(value)))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                load().mavi_sys_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
