"""ctypes front end of the CPU oracle (liboracle.so, built from mavi_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product (mavi_hip /
libmavi_hip.so) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

KIND_LJ, KIND_HT, KIND_RINGS, KIND_SZABO, KIND_RTP = 0, 1, 2, 3, 4
(WK_RIGID_RECT, WK_PERIODIC_RECT, WK_RIGID_CIRCLE, WK_SLIPPERY_LINES, WK_SLIPPERY_CIRCLE, WK_POTENTIAL_CIRCLE,
 WK_POTENTIAL_LINES) = range(7)
WMODE_OUTSIDE, WMODE_INSIDE, WMODE_REPULSION = 0, 1, 2


def build() -> Path:
    r = subprocess.run(["make", "-B", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB


class OrWall(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("mode", C.c_int32),
        ("pot_kind", C.c_int32),
        ("nlines", C.c_int32),
        ("geom", C.c_double * 4),
        ("pot", C.c_double * 4),
        ("lines", C.POINTER(C.c_double)),
    ]


class OrVarRings(C.Structure):
    _fields_ = [
        ("mask", C.POINTER(C.c_int32)),
        ("ids", C.POINTER(C.c_int64)),
        ("uids", C.POINTER(C.c_int64)),
        ("num_active", C.c_int64),
        ("cms", C.POINTER(C.c_double)),
        ("cpos", C.POINTER(C.c_double)),
        ("cp", C.POINTER(C.c_int64)),
        ("counts", C.POINTER(C.c_int64)),
        ("pol_draws", C.POINTER(C.c_double)),
        ("npol_draws", C.c_int64),
        ("pol_used", C.c_int64),
    ]


class OrSource(C.Structure):
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


SRC_SOURCE, SRC_SINK = 0, 1
GEOM_RECT, GEOM_CIRCLE = 0, 1


class OrSys(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("bl", C.c_double * 2),
        ("size", C.c_double * 2),
        ("periodic", C.c_int32),
        ("kind", C.c_int32),
        ("p", C.c_double * 4),
        ("dt", C.c_double),
        ("cols", C.c_int32),
        ("rows", C.c_int32),
        ("nc", C.c_int32),
        ("nthreads", C.c_int32),
        ("cl", C.c_double),
        ("ch", C.c_double),
        ("radius", C.c_double),
        ("nmax", C.c_int32),
        ("nrings", C.c_int32),
        ("ntypes", C.c_int32),
        ("pad1", C.c_int32),
        ("types", C.POINTER(C.c_int32)),
        ("np_type", C.POINTER(C.c_int32)),
        ("ring_par", C.POINTER(C.c_double)),
        ("inter", C.POINTER(C.c_double)),
        ("ap", C.c_double * 4),
        ("nwalls", C.c_int32),
        ("pad2", C.c_int32),
        ("walls", C.POINTER(OrWall)),
        ("var", C.c_void_p),
        ("nsrc", C.c_int32),
        ("pad3", C.c_int32),
        ("src", C.c_void_p),
    ]


_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_S = C.POINTER(OrSys)

_SIG = {
    "or_jl_div": (C.c_double, [C.c_double, C.c_double]),
    "or_chunk_capacity": (C.c_int64, [C.c_double, C.c_double, C.c_double]),
    "or_cell": (C.c_int, [_S, C.c_double, C.c_double, _I64, _I64]),
    "or_particle_ids": (C.c_int64, [_S, _I64]),
    "or_bin": (C.c_int, [_S, _D, _I64, _I64]),
    "or_neighbors": (None, [C.c_int32, C.c_int32, C.c_int32, _I64, _I32]),
    "or_calc_diff": (None, [_S, _D, _D, _D]),
    "or_pair_force": (None, [_S, _D, C.c_int64, C.c_int64, _D]),
    "or_calc_forces": (C.c_int, [_S, _D, _I64, _I64, _D, _D]),
    "or_calc_forces_threaded": (C.c_int, [_S, _D, _I64, _I64, _D, C.c_int]),
    "or_newton_steps": (C.c_int, [_S, _D, _D, _D, C.c_int64, _D, _D, C.c_int]),
    "or_kinetic_energy": (C.c_double, [C.c_int64, _D]),
    "or_potential_energy_lj": (C.c_double, [_S, _D]),
    "or_rings_continuous_pos": (None, [_S, _D, _D]),
    "or_rings_area": (C.c_double, [_S, _D, C.c_int32]),
    "or_rings_forces": (C.c_int, [_S, _D, _D, _D, _D]),
    "or_rings_steps": (C.c_int, [_S, _D, _D, C.c_int64, _D, _D, _D, _D]),
    "or_rings_var_init": (C.c_int, [_S, _D]),
    "or_sp_steps": (C.c_int, [_S, _D, _D, C.c_int64, _D, _D, C.c_int]),
    "or_wall_forces": (None, [_S, _D, C.c_int64, _D]),
    "or_rings_neighbors": (C.c_int, [_S, _D, C.c_double, C.c_int, _I64, _I64]),
    "or_point_line_intersect": (C.c_int, [_D, _D, _D]),
    "or_rings_invasions": (C.c_int, [_S, _D, _D, C.c_int32, C.c_int32, C.c_int64, _I64, C.c_int64, _I64]),
    "or_walls_apply": (None, [_S, _D, _D, C.c_int64, C.c_double]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        for k, (res, args) in _SIG.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_D)


def _i64(a):
    return None if a is None else a.ctypes.data_as(_I64)


def jl_div(a: float, b: float) -> float:
    return lib().or_jl_div(a, b)


def point_line_intersect(p, a, b) -> bool:
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (p, a, b)]
    return bool(lib().or_point_line_intersect(*[_d(v) for v in arrs]))


def chunk_capacity(cl, ch, r) -> int:
    return int(lib().or_chunk_capacity(cl, ch, r))


class Sys:
    """One oracle system.  Positions are (N, 2) float64 arrays; ids 1-based."""

    def __init__(self, *, n, bl, size, periodic, kind, p, dt=0.0, cols=0, rows=0, nc=0, radius=0.0,
                 nmax=0, nrings=0, types=None, np_type=None, ring_par=None, inter=None, ap=None, walls=None):
        """walls: list of dicts {kind, mode=0, pot_kind=0, geom=(4), pot=(4), lines=(k,4)} (WK_*)."""
        s = OrSys()
        s.n = n
        s.bl[0], s.bl[1] = bl
        s.size[0], s.size[1] = size
        s.periodic = 1 if periodic else 0
        s.kind = kind
        for i, v in enumerate(p):
            s.p[i] = v
        s.dt = dt
        s.cols, s.rows = cols, rows
        if cols > 0:
            s.cl = size[0] / cols
            s.ch = size[1] / rows
        s.nc = nc
        s.radius = radius
        for i, v in enumerate(ap or ()):
            s.ap[i] = v
        self._keep = []
        if walls:
            arr = (OrWall * len(walls))()
            for k, w in enumerate(walls):
                arr[k].kind = w["kind"]
                arr[k].mode = w.get("mode", 0)
                arr[k].pot_kind = w.get("pot_kind", 0)
                for j, v in enumerate(w.get("geom", ())):
                    arr[k].geom[j] = v
                for j, v in enumerate(w.get("pot", ())):
                    arr[k].pot[j] = v
                if w.get("lines") is not None:
                    ln = np.ascontiguousarray(w["lines"], dtype=np.float64).reshape(-1, 4)
                    arr[k].nlines = len(ln)
                    arr[k].lines = ln.ctypes.data_as(_D)
                    self._keep.append(ln)
            s.nwalls = len(walls)
            s.walls = C.cast(arr, C.POINTER(OrWall))
            self._keep.append(arr)
        if kind == KIND_RINGS:
            s.nmax, s.nrings = nmax, nrings
            np_type = np.ascontiguousarray(np.atleast_1d(np_type), dtype=np.int32)
            ring_par = np.ascontiguousarray(ring_par, dtype=np.float64).reshape(-1, 8)
            inter = np.ascontiguousarray(inter, dtype=np.float64)
            s.ntypes = len(np_type)
            s.np_type = np_type.ctypes.data_as(_I32)
            s.ring_par = ring_par.ctypes.data_as(_D)
            s.inter = inter.ctypes.data_as(_D)
            self._keep += [np_type, ring_par, inter]
            if types is not None:
                types = np.ascontiguousarray(types, dtype=np.int32)
                s.types = types.ctypes.data_as(_I32)
                self._keep.append(types)
        self.s = s
        self.n = n
        self.cols, self.rows, self.nc = cols, rows, nc
        self.nrings, self.nmax = nrings, nmax

    @property
    def ptr(self):
        return C.byref(self.s)

    def cell(self, x, y):
        r, c = C.c_int64(), C.c_int64()
        err = lib().or_cell(self.ptr, x, y, C.byref(r), C.byref(c))
        return err, r.value, c.value

    def particle_ids(self):
        k = lib().or_particle_ids(self.ptr, None)
        ids = np.zeros(k, dtype=np.int64)
        lib().or_particle_ids(self.ptr, _i64(ids))
        return ids

    def bin(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        cp = np.zeros((self.cols, self.rows, self.nc), dtype=np.int64)
        counts = np.zeros((self.cols, self.rows), dtype=np.int64)
        err = lib().or_bin(self.ptr, _d(pos), _i64(cp), _i64(counts))
        return err, cp.transpose(2, 1, 0), counts.transpose(1, 0)

    def _bin_raw(self, pos):
        cp = np.zeros((self.cols, self.rows, self.nc), dtype=np.int64)
        counts = np.zeros((self.cols, self.rows), dtype=np.int64)
        err = lib().or_bin(self.ptr, _d(pos), _i64(cp), _i64(counts))
        if err:
            raise RuntimeError(f"oracle binning error {err}")
        return cp, counts

    def forces(self, pos, with_absum=False):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        cp = counts = None
        if self.cols > 0:
            cp, counts = self._bin_raw(pos)
        f = np.zeros((self.n, 2))
        ab = np.zeros(self.n) if with_absum else None
        lib().or_calc_forces(self.ptr, _d(pos), _i64(cp), _i64(counts), _d(f), _d(ab))
        return (f, ab) if with_absum else f

    def forces_threaded(self, pos, nthreads):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        cp, counts = self._bin_raw(pos)
        f = np.zeros((self.n, 2))
        lib().or_calc_forces_threaded(self.ptr, _d(pos), _i64(cp), _i64(counts), _d(f), nthreads)
        return f

    def pair_force(self, pos, i, j):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        f = np.zeros(2)
        lib().or_pair_force(self.ptr, _d(pos), i, j, _d(f))
        return f

    def calc_diff(self, r1, r2):
        a = np.ascontiguousarray(r1, dtype=np.float64)
        b = np.ascontiguousarray(r2, dtype=np.float64)
        out = np.zeros(2)
        lib().or_calc_diff(self.ptr, _d(a), _d(b), _d(out))
        return out

    def newton_steps(self, pos, vel, nsteps, record=False, nthreads=1):
        pos = np.array(pos, dtype=np.float64, order="C")
        vel = np.array(vel, dtype=np.float64, order="C")
        f = np.zeros((self.n, 2))
        ke = np.zeros(nsteps) if record else None
        pe = np.zeros(nsteps) if record else None
        err = lib().or_newton_steps(self.ptr, _d(pos), _d(vel), _d(f), nsteps, _d(ke), _d(pe), nthreads)
        if err:
            raise RuntimeError(f"oracle newton_steps error {err}")
        return pos, vel, f, ke, pe

    def kinetic_energy(self, vel):
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        return lib().or_kinetic_energy(len(vel), _d(vel))

    def potential_energy_lj(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        return lib().or_potential_energy_lj(self.ptr, _d(pos))

    # rings
    def rings_continuous_pos(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        out = np.zeros_like(pos)
        lib().or_rings_continuous_pos(self.ptr, _d(pos), _d(out))
        return out

    def rings_area(self, cpos, ring):
        cpos = np.ascontiguousarray(cpos, dtype=np.float64)
        return lib().or_rings_area(self.ptr, _d(cpos), ring)

    def rings_forces(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        cpos = self.rings_continuous_pos(pos)
        f = np.zeros((self.n, 2))
        areas = np.zeros(self.nrings)
        err = lib().or_rings_forces(self.ptr, _d(pos), _d(cpos), _d(f), _d(areas))
        if err:
            raise RuntimeError(f"oracle rings_forces error {err}")
        return f, areas

    # variable ring population (VarRingsIds) with sources and sinks
    def set_var_rings(self, mask, sources=(), pol_draws=None):
        """Give the rings a variable population: mask (nrings) of active
        rings and a list of source/sink dicts
        {kind: SRC_SOURCE, bbox (nbox,4), spawn (nbox,nmax,2), pad, pol |
         random_pol, chunk_lists: list of 0-based chunk ids per box or None}
        {kind: SRC_SINK, geom: GEOM_RECT/GEOM_CIRCLE, g: (4,)}.
        Call rings_var_init(pos) next (the constructor's sequence)."""
        nr, nm = self.nrings, self.nmax
        ncell = max(self.rows * self.cols, 1)
        self.var = dict(mask=np.ascontiguousarray(mask, dtype=np.int32).copy(),
                        ids=np.zeros(nr, dtype=np.int64), uids=np.zeros(nr, dtype=np.int64),
                        cms=np.zeros((nr, 2)), cpos=np.zeros((nr * nm, 2)),
                        cp=np.zeros(max(self.nc, 1) * ncell, dtype=np.int64), counts=np.zeros(ncell, dtype=np.int64),
                        pol_draws=np.ascontiguousarray(pol_draws if pol_draws is not None else [0.0],
                                                       dtype=np.float64))
        v = OrVarRings()
        d = self.var
        v.mask = d["mask"].ctypes.data_as(_I32)
        v.ids = d["ids"].ctypes.data_as(_I64)
        v.uids = d["uids"].ctypes.data_as(_I64)
        v.cms = _d(d["cms"])
        v.cpos = _d(d["cpos"])
        v.cp = _i64(d["cp"])
        v.counts = _i64(d["counts"])
        v.pol_draws = _d(d["pol_draws"])
        v.npol_draws = 0 if pol_draws is None else len(d["pol_draws"])
        self._var_struct = v
        srcs = (OrSource * max(len(sources), 1))()
        keep = []
        for k, sc in enumerate(sources):
            o = srcs[k]
            o.kind = sc["kind"]
            if sc["kind"] == SRC_SINK:
                o.geom = sc["geom"]
                for j, val in enumerate(sc["g"]):
                    o.g[j] = val
                continue
            bbox = np.ascontiguousarray(sc["bbox"], dtype=np.float64).reshape(-1, 4)
            spawn = np.ascontiguousarray(sc["spawn"], dtype=np.float64).reshape(len(bbox), nm, 2)
            o.nbox = len(bbox)
            o.pad = sc.get("pad", 0.0)
            o.random_pol = 1 if sc.get("random_pol") else 0
            o.pol = sc.get("pol", 0.0)
            o.bbox, o.spawn = _d(bbox), _d(spawn)
            keep += [bbox, spawn]
            if sc.get("chunk_lists") is not None:
                off = np.zeros(len(bbox) + 1, dtype=np.int32)
                off[1:] = np.cumsum([len(c) for c in sc["chunk_lists"]])
                ids = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.int32) for c in sc["chunk_lists"]]
                                                          + [np.zeros(0, np.int32)]), dtype=np.int32)
                o.chunk_off, o.chunk_ids = off.ctypes.data_as(_I32), ids.ctypes.data_as(_I32)
                keep += [off, ids]
        self._src_keep = (srcs, keep)
        self.s.var = C.cast(C.pointer(v), C.c_void_p)
        self.s.nsrc = len(sources)
        self.s.src = C.cast(srcs, C.c_void_p)

    def rings_var_init(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        err = lib().or_rings_var_init(self.ptr, _d(pos))
        if err:
            raise RuntimeError(f"oracle rings_var_init error {err}")

    def var_state(self):
        """(mask, ids, uids, num_active, cms, pol draws used) of the variable population."""
        v = self._var_struct
        d = self.var
        return (d["mask"].copy(), d["ids"].copy(), d["uids"].copy(), int(v.num_active), d["cms"].copy(),
                int(v.pol_used))

    def rings_steps(self, pos, pol, nsteps, noise=None, nthreads=1):
        """nthreads > 1: calc_forces! in Threaded mode (per-thread force slices
        summed in thread order, src/integration.jl:160-193) and per-ring
        threads for the springs, area forces and update."""
        self.s.nthreads = int(nthreads)
        pos = np.array(pos, dtype=np.float64, order="C")
        pol = np.array(pol, dtype=np.float64, order="C")
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.float64)
        f = np.zeros((self.n, 2))
        areas = np.zeros(self.nrings)
        cms = np.zeros((self.nrings, 2))
        err = lib().or_rings_steps(self.ptr, _d(pos), _d(pol), nsteps, _d(noise), _d(f), _d(areas), _d(cms))
        if err:
            raise RuntimeError(f"oracle rings_steps error {err}")
        return pos, pol, f, areas, cms


    def rings_neighbors(self, pos, tol=1.1, only_rings=False):
        """(counts (nslots,), lists (nslots, 15)) of ParticleNeighbors: 1-based
        ids in the reference's visiting order, -1 padded."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        ns = self.nmax * self.nrings
        counts = np.zeros(ns, dtype=np.int64)
        lists = np.zeros((ns, 15), dtype=np.int64)
        err = lib().or_rings_neighbors(self.ptr, _d(pos), tol, 1 if only_rings else 0, _i64(counts), _i64(lists))
        if err:
            raise RuntimeError(f"oracle rings_neighbors error {err}")
        return counts, lists

    def rings_invasions(self, rpts, cms, rcols=0, rrows=0, rnc=0):
        """check_invasions!: list of (invasor, invaded, p_id), 1-based, in the
        reference's push order.  rpts = ring_points (continuous positions for
        a periodic main wall), cms = centres of mass (nrings, 2)."""
        rpts = np.ascontiguousarray(rpts, dtype=np.float64)
        cms = np.ascontiguousarray(cms, dtype=np.float64)
        cap = max(16, self.nmax * self.nrings)
        while True:
            rec = np.zeros((cap, 3), dtype=np.int64)
            count = np.zeros(1, dtype=np.int64)
            err = lib().or_rings_invasions(self.ptr, _d(rpts), _d(cms), rcols, rrows, rnc, _i64(rec), cap,
                                           _i64(count))
            if err:
                raise RuntimeError(f"oracle rings_invasions error {err}")
            if count[0] <= cap:
                return [tuple(int(v) for v in r) for r in rec[:count[0]]]
            cap = int(count[0])

    # walls
    def wall_forces(self, pos, i):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        f = np.zeros(2)
        lib().or_wall_forces(self.ptr, _d(pos), i, _d(f))
        return f

    def walls_apply(self, pos, vel, i, radius):
        """walls! on particle i (0-based); returns the new (pos, vel) arrays."""
        pos = np.array(pos, dtype=np.float64, order="C")
        vel = None if vel is None else np.array(vel, dtype=np.float64, order="C")
        lib().or_walls_apply(self.ptr, _d(pos), _d(vel), i, radius)
        return pos, vel

    # self-propelled (Szabo / run-and-tumble)
    def sp_steps(self, pos, ang, nsteps, noise=None, nthreads=1):
        pos = np.array(pos, dtype=np.float64, order="C")
        ang = np.array(ang, dtype=np.float64, order="C")
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.float64)
        f = np.zeros((self.n, 2))
        err = lib().or_sp_steps(self.ptr, _d(pos), _d(ang), nsteps, _d(noise), _d(f), nthreads)
        if err:
            raise RuntimeError(f"oracle sp_steps error {err}")
        return pos, ang, f


def neighbors(rows, cols, periodic):
    nb = np.zeros(rows * cols * 8, dtype=np.int64)
    nnb = np.zeros(rows * cols, dtype=np.int32)
    lib().or_neighbors(rows, cols, 1 if periodic else 0, _i64(nb), nnb.ctypes.data_as(_I32))
    out = {}
    for c in range(1, cols + 1):
        for r in range(1, rows + 1):
            k = (c - 1) * rows + (r - 1)
            out[(r, c)] = [(int(nb[(k * 4 + e) * 2]), int(nb[(k * 4 + e) * 2 + 1])) for e in range(nnb[k])]
    return out


# --- build an oracle system from the mavi_hip host objects -----------------
def from_system(pos_n, space_cfg, dynamic_cfg, int_cfg, chunks=None, radius=None):
    """Oracle Sys mirroring a plain GPU System configuration."""
    from mavi_hip.configs import (HarmTruncCfg, LenJonesCfg, PeriodicWalls, RunTumbleCfg, SzaboCfg,
                                  get_bounding_box, particle_radius)
    from mavi_hip.systems import wall_records

    g = get_bounding_box(space_cfg.geometry_cfg)
    walls = wall_records(space_cfg)  # mavi_wall kinds and modes are the oracle's
    ap = None
    if isinstance(dynamic_cfg, LenJonesCfg):
        kind, p = KIND_LJ, [dynamic_cfg.sigma, dynamic_cfg.epsilon, dynamic_cfg.cutoff or 0.0, 0.0]
    elif isinstance(dynamic_cfg, HarmTruncCfg):
        kind, p = KIND_HT, [dynamic_cfg.k_rep, dynamic_cfg.k_atr, dynamic_cfg.dist_eq, dynamic_cfg.dist_max]
    elif isinstance(dynamic_cfg, SzaboCfg):
        c = dynamic_cfg
        kind, p = KIND_SZABO, [c.k_rep, c.k_adh, c.r_eq, c.r_max]
        ap = [c.vo, c.mobility, c.relax_time, c.rot_diff]
    elif isinstance(dynamic_cfg, RunTumbleCfg):
        c = dynamic_cfg
        kind, p = KIND_RTP, [c.sigma, c.epsilon, c.cutoff(), 0.0]
        ap = [c.vo, c.tumble_rate, 0.0, 0.0]
    else:
        raise TypeError(type(dynamic_cfg))
    cols = rows = nc = 0
    if int_cfg.chunks_cfg is not None:
        cols, rows = int_cfg.chunks_cfg.num_cols, int_cfg.chunks_cfg.num_rows
        r = particle_radius(dynamic_cfg)
        nc = chunk_capacity(g.length / cols, g.height / rows, r)
    return Sys(n=pos_n, bl=g.bottom_left, size=(g.length, g.height),
               periodic=isinstance(space_cfg.main()[0], PeriodicWalls), kind=kind, p=p, dt=int_cfg.dt,
               cols=cols, rows=rows, nc=nc, radius=radius if radius is not None else particle_radius(dynamic_cfg),
               ap=ap, walls=walls)


def from_rings(state, space_cfg, cfg, int_cfg):
    """Oracle Sys mirroring a rings GPU configuration."""
    from mavi_hip.configs import PeriodicWalls, get_bounding_box
    from mavi_hip.systems import wall_records

    g = get_bounding_box(space_cfg.geometry_cfg)
    cols = rows = nc = 0
    r = min(cfg.particle_radii())
    if int_cfg.chunks_cfg is not None:
        cols, rows = int_cfg.chunks_cfg.num_cols, int_cfg.chunks_cfg.num_rows
        nc = chunk_capacity(g.length / cols, g.height / rows, r)
    inter = cfg.interaction_table()
    return Sys(n=len(state.pos), bl=g.bottom_left, size=(g.length, g.height),
               periodic=isinstance(space_cfg.main()[0], PeriodicWalls), kind=KIND_RINGS, p=[0, 0, 0, 0],
               dt=int_cfg.dt, cols=cols, rows=rows, nc=nc, radius=r, nmax=state.nmax, nrings=state.num_rings,
               types=state.types, np_type=np.atleast_1d(state.num_particles), ring_par=cfg.type_table(),
               inter=inter.reshape(-1), walls=wall_records(space_cfg))


def source_dicts(sources, nmax):
    """Oracle records of the mirror's Source / SinkCfg objects (list order kept)."""
    from mavi_hip.configs import CircleCfg

    out = []
    for s in sources:
        if hasattr(s, "geometry_cfg"):  # SinkCfg
            g = s.geometry_cfg
            if isinstance(g, CircleCfg):
                out.append(dict(kind=SRC_SINK, geom=GEOM_CIRCLE, g=(g.center[0], g.center[1], g.radius, 0.0)))
            else:
                out.append(dict(kind=SRC_SINK, geom=GEOM_RECT, g=(g.bottom_left[0], g.bottom_left[1], g.length,
                                                                    g.height)))
            continue
        spawn = np.stack(s.bbox_spawn_pos)
        if spawn.shape[1] != nmax:
            raise ValueError("spawn_pos must hold num_max_particles points (add_ring! writes every slot)")
        rnd = isinstance(s.cfg.spawn_pol, str)
        out.append(dict(kind=SRC_SOURCE, bbox=np.array(s.bbox), spawn=spawn, pad=float(s.cfg.pad), random_pol=rnd,
                        pol=0.0 if rnd else float(s.cfg.spawn_pol), chunk_lists=s.chunk_ids))
    return out
