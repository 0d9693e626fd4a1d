"""Mirror of Mavi.Rings for the GPU path (src/rings/*.jl).

  RingsCfg, HarmTruncCfg, InteractionMatrix   src/rings/configs.jl:26-164
  RingsState                                  src/rings/states.jl:74-124
  RingsSystem                                 src/rings/rings.jl:231-291
  step(system)                                Rings.Integration.step!, src/rings/integration.jl:522-543
  rectangular_grid, create_circle, random_pol src/rings/init_states.jl
  get_ring_radius                             src/rings/utils.jl:5-7
  InvasionsCfg, IntCfgExtra, RingsIntCfg      src/rings/configs.jl:334-350
  Invasion, RingsSystem.invasions             src/rings/rings.jl:86-100
  equilibrium_area, equilibrium_ring_radius   src/rings/configs.jl:221-328
  ActiveState, VarRingsIds                    src/states.jl:14-20, src/rings/states.jl:20-39, 200-223
  SourceCfg, SinkCfg, Source                  src/rings/sources.jl:31-185, 244-251

Host layout: ``rings_pos`` is an (nrings, nmax, 2) array, i.e. Julia's
Matrix{SVector{2,T}}(nmax, nrings) in memory order, and ``pos`` is its
(nrings*nmax, 2) reshape sharing memory (slot = ring*nmax + p).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .configs import (GPU, ChunksCfg, CircleCfg, IntCfg, PeriodicWalls, RectangleCfg, SpaceCfg, chunk_capacity,
                      get_bounding_box)
from .systems import Chunks, TimeInfo, attach_walls, base_desc, check_inside, wall_kind, wall_records


@dataclass
class HarmTruncCfg:  # src/rings/configs.jl:26-57
    k_rep: float
    k_atr: float
    dist_eq: float
    dist_max: float

    def params(self):
        return [float(self.k_rep), float(self.k_atr), float(self.dist_eq), float(self.dist_max)]

    def particle_radius(self):
        return self.dist_eq / 2.0


@dataclass
class InteractionMatrix:  # src/rings/configs.jl:61-63; matrix[t1][t2]
    matrix: list


_FIELDS = ("p0", "relax_time", "vo", "mobility", "rot_diff", "k_area", "k_spring", "l_spring")


class RingsCfg:  # src/rings/configs.jl:95-150
    def __init__(self, *, p0, relax_time, vo, mobility, rot_diff, k_area, k_spring, l_spring,
                 interaction_finder, num_particles=-1):
        args = dict(p0=p0, relax_time=relax_time, vo=vo, mobility=mobility, rot_diff=rot_diff,
                    k_area=k_area, k_spring=k_spring, l_spring=l_spring)
        lens = [len(v) if np.ndim(v) else 1 for v in args.values()]
        self.num_types = max(lens)
        self.typed = self.num_types > 1 or any(np.ndim(v) for v in args.values())
        for k, v in args.items():
            arr = np.atleast_1d(np.asarray(v, dtype=np.float64))
            if arr.size == 1 and self.num_types > 1:
                arr = np.full(self.num_types, float(arr[0]))
            setattr(self, k, arr)
        self.num_particles = num_particles
        self.interaction_finder = interaction_finder

    def type_table(self) -> np.ndarray:
        return np.stack([getattr(self, f) for f in _FIELDS], axis=1).astype(np.float64)  # (ntypes, 8)

    def interaction_table(self) -> np.ndarray:
        f = self.interaction_finder
        nt = self.num_types
        tab = np.zeros((nt, nt, 4))  # [t2, t1, :] so that memory is column-major [t1, t2]
        for t1 in range(nt):
            for t2 in range(nt):
                cfg = f.matrix[t1][t2] if isinstance(f, InteractionMatrix) else f
                tab[t2, t1] = cfg.params()
        return tab

    def particle_radius(self, type_=1):
        f = self.interaction_finder
        cfg = f.matrix[type_ - 1][type_ - 1] if isinstance(f, InteractionMatrix) else f
        return cfg.particle_radius()

    def particle_radii(self):
        return [self.particle_radius(t + 1) for t in range(self.num_types)]


@dataclass
class ActiveState:  # src/states.jl:14-20
    """Initial active mask of a variable ring population (True: all)."""

    mask: object = True


class VarRingsIds:  # src/rings/states.jl:20-39
    """Active rings of a variable population: mask, ids (1-based; the
    entries past num_active keep what an earlier update left), uids,
    p_ids (0-based slots) and the counts, as calc_active_ids! leaves them."""

    def __init__(self, mask, nmax, ring_np):
        n = len(mask)
        self.mask = np.asarray(mask, dtype=bool).copy()
        self.ids = np.arange(1, n + 1, dtype=np.int64)
        self.uids = np.arange(1, n + 1, dtype=np.int64)
        self.num_active = 0
        self._nmax, self._ring_np = nmax, ring_np
        self.update()

    def update(self):
        """calc_active_ids!, src/rings/states.jl:200-223."""
        act = np.nonzero(self.mask)[0]
        self.ids[:len(act)] = act + 1
        self.num_active = len(act)
        self.p_ids = np.concatenate([r * self._nmax + np.arange(self._ring_np(r + 1)) for r in act]
                                    + [np.zeros(0, dtype=np.int64)]).astype(np.int64)
        self.num_p_active = len(self.p_ids)

    def get_rings_ids(self):
        return self.ids[:self.num_active]


class RingsState:  # src/rings/states.jl:74-124
    def __init__(self, *, rings_pos, pol, num_particles=None, types=None, active_state=None):
        rp = np.asarray(rings_pos, dtype=np.float64)
        if rp.ndim == 3 and rp.shape[0] == 2 and rp.shape[2] != 2:
            rp = np.transpose(rp, (2, 1, 0))  # Julia (2, nmax, nrings) -> (nrings, nmax, 2)
        self.rings_pos = np.ascontiguousarray(rp)
        self.num_rings, self.nmax = self.rings_pos.shape[:2]
        self.pos = self.rings_pos.reshape(self.num_rings * self.nmax, 2)
        self.pol = np.ascontiguousarray(np.asarray(pol, dtype=np.float64)).copy()
        if num_particles is None:
            num_particles = self.nmax
        self.num_particles = num_particles
        self.types = None if types is None else np.asarray(types, dtype=np.int32)
        self.rings_ids = None  # FixRingsIds
        if active_state is not None:
            mask = active_state.mask
            if np.ndim(mask) == 0:
                mask = np.full(self.num_rings, bool(mask))
            self.rings_ids = VarRingsIds(mask, self.nmax, self.ring_num_particles)

    def ring_num_particles(self, ring_id):  # 1-based ring id
        if self.types is None:
            return int(self.num_particles)
        return int(np.asarray(self.num_particles)[self.types[ring_id - 1] - 1])

    def ring_num_particles_of_type(self, type_):  # 1-based type
        return int(np.atleast_1d(np.asarray(self.num_particles))[type_ - 1 if self.types is not None else 0])

    def particle_ids(self) -> np.ndarray:
        """get_particles_ids: FixRingsIds.p_ids or VarRingsIds.p_ids (0-based slots)."""
        if self.rings_ids is not None:
            return self.rings_ids.p_ids
        if self.types is None:
            return np.arange(len(self.pos))
        ids = [r * self.nmax + i for r in range(self.num_rings) for i in range(self.ring_num_particles(r + 1))]
        return np.asarray(ids)


def describe(state: RingsState, space_cfg: SpaceCfg, dynamic_cfg: RingsCfg, int_cfg, seed: int = 0):
    """Descriptor of a rings system for the C ABI, with the checks of
    RingsSystem(...) (src/rings/rings.jl:231-291).  Returns (SysDesc, Chunks |
    None, tables): the descriptor points into ``tables``, which must outlive
    the call that reads it."""
    if dynamic_cfg.typed != (state.types is not None):
        raise ValueError("DynamicCfg types and state.types disagree")
    if not isinstance(space_cfg.main()[0], PeriodicWalls) or not isinstance(int_cfg.device, GPU):
        raise _lib.UnsupportedOnGPU(_lib.MAVI_E_UNSUPPORTED,
                                    "rings run on the GPU with a PeriodicWalls rectangle as the main wall")
    ids = state.particle_ids()
    ok, out_ids = check_inside(state.pos, space_cfg.geometry_cfg, ids)
    if not ok:
        raise ValueError(f"Particles with ids={list(out_ids + 1)} outside space.")
    r = min(dynamic_cfg.particle_radii())  # minimum(particle_radius(dynamic_cfg)), src/systems.jl:26
    chunks = None
    if int_cfg.chunks_cfg is not None:
        chunks = Chunks(int_cfg.chunks_cfg.num_cols, int_cfg.chunks_cfg.num_rows,
                        get_bounding_box(space_cfg.geometry_cfg), r)
    d = base_desc(len(state.pos), space_cfg, dynamic_cfg, int_cfg, chunks, r)
    walls_keep = attach_walls(d, wall_records(space_cfg))
    d.potential = _lib.MAVI_POT_RINGS
    d.nmax, d.nrings, d.ntypes = state.nmax, state.num_rings, dynamic_cfg.num_types
    np_types = np.atleast_1d(np.asarray(state.num_particles, dtype=np.int32))
    tables = (
        np.ascontiguousarray(np_types),
        np.ascontiguousarray(dynamic_cfg.type_table()),
        np.ascontiguousarray(dynamic_cfg.interaction_table()),
        None if state.types is None else np.ascontiguousarray(state.types, dtype=np.int32),
        walls_keep,
    )
    import ctypes as C
    d.type_np = tables[0].ctypes.data_as(C.POINTER(C.c_int32))
    d.type_par = tables[1].ctypes.data_as(C.POINTER(C.c_double))
    d.interaction = tables[2].ctypes.data_as(C.POINTER(C.c_double))
    if tables[3] is not None:
        d.ring_types = tables[3].ctypes.data_as(C.POINTER(C.c_int32))
    d.seed = int(seed)
    return d, chunks, tables


# --- integration configs and invasions (src/rings/configs.jl:334-350,
# src/rings/rings.jl:86-100) ---------------------------------------------
@dataclass
class InvasionsCfg:
    """Check for invasions every steps_to_update steps."""

    steps_to_update: int


@dataclass
class IntCfgExtra:
    r_chunks_cfg: ChunksCfg | None = None
    invasions_cfg: InvasionsCfg | None = None


def RingsIntCfg(*, dt, p_chunks_cfg=None, r_chunks_cfg=None, invasions_cfg=None, device=None) -> IntCfg:
    """IntCfg of a rings system: particle chunks, ring chunks (binning the
    rings' centres of mass) and the invasion check schedule."""
    return IntCfg(dt=dt, chunks_cfg=p_chunks_cfg, device=device or GPU(),
                  extra=IntCfgExtra(r_chunks_cfg, invasions_cfg))


@dataclass(frozen=True)
class Invasion:
    """Particle p_id (scalar index, 1-based) of ring `invasor` lies inside
    ring `invaded`."""

    invasor: int
    invaded: int
    p_id: int


def _type_prop(cfg: RingsCfg, name: str, type_: int) -> float:
    return float(getattr(cfg, name)[type_ - 1 if cfg.num_types > 1 else 0])


def equilibrium_area(cfg: RingsCfg, num_particles: int, type_: int = 1) -> float:
    """get_equilibrium_area (src/rings/configs.jl:246-293): area0 below the
    springs-only limit, otherwise the root of the area/spring force balance.
    The reference solves that balance with NLsolve.jl 4.5.1 (nlsolve from
    f = sqrt(0.5)); this takes Newton steps on the same scalar equation from
    the same start, which reaches the same root to the solver's tolerance."""
    n = num_particles
    l_spring, p0 = _type_prop(cfg, "l_spring", type_), _type_prop(cfg, "p0", type_)
    a0 = (n * l_spring / p0) ** 2  # get_area0, :221-226
    theta = 2 * math.pi / n
    p0_lim = 2 * (n * (1 - math.cos(theta)) / math.sin(theta)) ** 0.5  # get_equilibrium_p0, :233-237
    if a0 < (n * l_spring / p0_lim) ** 2:
        return a0
    k_a, k_m = _type_prop(cfg, "k_area", type_), _type_prop(cfg, "k_spring", type_)

    def get_r(f):
        return math.sqrt(f * 2 * a0 / (n * math.sin(theta)))

    def balance(x):
        f = x * x
        r = get_r(f)
        fa = k_a * (a0 - 10 / 2 * r ** 2 * math.sin(theta)) * r * math.sin(theta)
        fm = k_m * (r * math.sqrt(2 * (1 - math.cos(theta))) - l_spring)
        return fa - 2 * fm * math.sin(theta / 2)

    x = math.sqrt(0.5)
    for _ in range(100):
        h = 1e-7 * max(1.0, abs(x))
        d = (balance(x + h) - balance(x - h)) / (2 * h)
        if d == 0.0:
            break
        step = balance(x) / d
        x -= step
        if abs(step) <= 1e-14 * max(1.0, abs(x)):
            break
    a0_sol = x * x * a0
    return a0 if a0_sol < a0 else a0_sol


def equilibrium_ring_radius(cfg: RingsCfg, num_particles: int, type_: int = 1) -> float:
    """get_ring_radius(dynamic_cfg, type), src/rings/configs.jl:320-328: from
    the centre to the outer edge of the particles at the equilibrium area."""
    area_eq = equilibrium_area(cfg, num_particles, type_)
    n = num_particles
    return (2 * area_eq / (n * math.sin(2 * math.pi / n))) ** 0.5 + cfg.particle_radius(type_)


def ring_chunks_grid(state: RingsState, space_cfg: SpaceCfg, cfg: RingsCfg, r_chunks_cfg: ChunksCfg):
    """(cols, rows, nc) of r_chunks, src/rings/rings.jl:188-202: chunks over
    the bounding box, capacity from the smallest equilibrium ring radius."""
    g = get_bounding_box(space_cfg.geometry_cfg)
    cols, rows = r_chunks_cfg.num_cols, r_chunks_cfg.num_rows
    rings_r = min(equilibrium_ring_radius(cfg, state.ring_num_particles_of_type(t + 1), t + 1)
                  for t in range(cfg.num_types))
    return cols, rows, chunk_capacity(g.length / cols, g.height / rows, rings_r)


# --- sources and sinks (src/rings/sources.jl) ------------------------------
@dataclass
class SourceCfg:  # src/rings/sources.jl:31-49
    """Spawn boxes of rings: size[0] x size[1] boxes around the spawn
    template, each refilled with a ring when no particle lies inside it
    (grown by pad).  spawn_pol: a number, or "random" (2 pi u per empty box)."""

    bottom_left: tuple
    spawn_pos: object
    spawn_pol: object = "random"
    pad: float = 0.0
    offset: tuple = (0.0, 0.0)
    size: tuple = (1, 1)


@dataclass
class SinkCfg:  # src/rings/sources.jl:244-251
    """Rings whose centre of mass lies inside geometry_cfg are removed."""

    geometry_cfg: object


def _rect_add(a, b):
    """Base.:+(::RectangleCfg, ::RectangleCfg), src/configs.jl:68-77, on
    (blx, bly, length, height) tuples."""
    max_x = max(a[0] + a[2], b[0] + b[2])
    max_y = max(a[1] + a[3], b[1] + b[3])
    min_x = min(a[0], b[0])
    min_y = min(a[1], b[1])
    return (min_x, min_y, max_x - min_x, max_y - min_y)


def _check_intersection(r1, r2):
    """check_intersection(::RectangleCfg, ::RectangleCfg), src/configs.jl:79-87."""
    ix1 = r2[0] <= r1[0] <= r2[0] + r2[2]
    ix2 = r2[0] <= r1[0] + r1[2] <= r2[0] + r2[2]
    iy1 = r2[1] <= r1[1] <= r2[1] + r2[3]
    iy2 = r2[1] <= r1[1] + r1[3] <= r2[1] + r2[3]
    return (ix1 or ix2) and (iy1 or iy2)


class Source:  # src/rings/sources.jl:127-185
    """Spawn boxes (bbox: (blx, bly, length, height)), their spawn positions
    and, with particle chunks, ChunksChecker's chunk lists (0-based chunk
    indices row0 * num_cols + col0); without chunks the PosChecker (every
    active particle) decides emptiness."""

    def __init__(self, cfg: SourceCfg, chunks: Chunks | None = None):
        self.cfg = cfg
        sp = np.asarray(cfg.spawn_pos, dtype=np.float64).reshape(-1, 2)
        pad = float(cfg.pad)
        min_x, max_x = float(sp[:, 0].min()), float(sp[:, 0].max())
        min_y, max_y = float(sp[:, 1].min()), float(sp[:, 1].max())
        bbox_length = max_x - min_x + 2 * pad
        bbox_height = max_y - min_y + 2 * pad
        blx, bly = float(cfg.bottom_left[0]), float(cfg.bottom_left[1])
        offx, offy = float(cfg.offset[0]), float(cfg.offset[1])
        self.bbox, bbox_pad = [], []
        for i in range(1, cfg.size[0] + 1):
            for j in range(1, cfg.size[1] + 1):
                bx = blx + ((i - 1) * bbox_length + i * offx)
                by = bly + ((j - 1) * bbox_height + j * offy)
                self.bbox.append((bx, by, bbox_length, bbox_height))
                bbox_pad.append((bx - pad, by - pad, bbox_length + 2 * pad, bbox_height + 2 * pad))
        # bbox_spawn_pos: spawn_pos .+ (box bottom_left - template bottom_left + (pad, pad))
        self.bbox_spawn_pos = [sp + np.array([(b[0] - min_x) + pad, (b[1] - min_y) + pad]) for b in self.bbox]
        self.chunk_ids = None if chunks is None else self._chunks_checker(chunks, bbox_pad)

    @staticmethod
    def _chunks_checker(chunks: Chunks, bbox_vec):
        """ChunksChecker(pos, chunks, bbox_vec), src/rings/sources.jl:64-119."""
        g = bbox_vec[0]
        for b in bbox_vec[1:]:  # sum(bbox_vec): a left fold of +
            g = _rect_add(g, b)
        geo = chunks.geometry_cfg
        cl, ch = chunks.chunk_length, chunks.chunk_height
        tlx, tly = geo.bottom_left[0], geo.bottom_left[1] + geo.height
        found, sx, ex = False, 0, 0
        for i in range(1, chunks.num_cols + 1):
            x1 = tlx + (i - 1) * cl
            x2 = x1 + cl
            if not found and x1 <= g[0] <= x2:
                found, sx = True, i
            if found and x1 >= g[0] + g[2]:
                ex = i - 1
                break
        found, sy, ey = False, 0, 0
        for i in range(1, chunks.num_rows + 1):
            y2 = tly - (i - 1) * ch
            y1 = y2 - ch
            if not found and y1 <= g[1] + g[3] <= y2:
                found, sy = True, i
            if found and y2 <= g[1]:
                ey = i - 1
                break
        ids = [[] for _ in bbox_vec]
        for col in range(sx, ex + 1):
            for row in range(sy, ey + 1):
                # get_chunk_rect, src/chunks.jl:50-58
                rect = (geo.bottom_left[0] + (col - 1) * cl, tly - row * ch, cl, ch)
                for b, lst in zip(bbox_vec, ids):
                    if _check_intersection(rect, b):
                        lst.append((row - 1) * chunks.num_cols + (col - 1))
        return ids


def source_records(sources, nmax: int, keep: list):
    """C ABI records (mavi_ring_source) of Source / SinkCfg objects, in list
    order; the arrays they point into are appended to keep."""
    import ctypes as C

    recs = []
    for src in sources:
        r = _lib.MaviRingSource()
        if isinstance(src, SinkCfg):
            g = src.geometry_cfg
            r.kind = _lib.MAVI_SRC_SINK
            if isinstance(g, CircleCfg):
                r.geom, vals = _lib.MAVI_GEOM_CIRCLE, (g.center[0], g.center[1], g.radius, 0.0)
            else:
                r.geom, vals = _lib.MAVI_GEOM_RECT, (g.bottom_left[0], g.bottom_left[1], g.length, g.height)
            for k, v in enumerate(vals):
                r.g[k] = float(v)
            recs.append(r)
            continue
        spawn = np.ascontiguousarray(np.stack(src.bbox_spawn_pos), dtype=np.float64)
        if spawn.shape[1] != nmax:
            raise ValueError(f"spawn_pos holds {spawn.shape[1]} points; add_ring! fills all {nmax} slots of a ring")
        bbox = np.ascontiguousarray(src.bbox, dtype=np.float64)
        r.kind = _lib.MAVI_SRC_SOURCE
        r.nbox = len(bbox)
        rnd = isinstance(src.cfg.spawn_pol, str)
        if rnd and src.cfg.spawn_pol != "random":
            raise ValueError(src.cfg.spawn_pol)
        r.random_pol = 1 if rnd else 0
        r.pol = 0.0 if rnd else float(src.cfg.spawn_pol)
        r.pad = float(src.cfg.pad)
        r.bbox = bbox.ctypes.data_as(C.POINTER(C.c_double))
        r.spawn = spawn.ctypes.data_as(C.POINTER(C.c_double))
        keep += [bbox, spawn]
        if src.chunk_ids is not None:
            off = np.zeros(len(bbox) + 1, dtype=np.int32)
            off[1:] = np.cumsum([len(c) for c in src.chunk_ids])
            ids = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.int32) for c in src.chunk_ids]
                                                      + [np.zeros(1, np.int32)]), dtype=np.int32)
            r.chunk_off = off.ctypes.data_as(C.POINTER(C.c_int32))
            r.chunk_ids = ids.ctypes.data_as(C.POINTER(C.c_int32))
            keep += [off, ids]
        recs.append(r)
    return recs


@dataclass
class NeighborsCfg:  # src/rings/neighbors.jl:11-15
    """Contact neighbours of the particles: type "all" or "rings" (pairs of
    different rings only), recorded below tol * 2 * particle_radius.
    only_count is accepted for the reference's signature; counts and lists
    are both kept (at most 15 per particle)."""

    only_count: bool = False
    type: str = "all"
    tol: float = 1.1


class RingsSystem:
    """GPU-backed RingsSystem (src/rings/rings.jl:231-291)."""

    def __init__(self, *, state: RingsState, space_cfg: SpaceCfg, dynamic_cfg: RingsCfg, int_cfg,
                 time_info: TimeInfo | None = None, seed: int = 0, ctx: _lib.Context | None = None,
                 p_neighbors_cfg: NeighborsCfg | None = None, source_cfg=None, rng=None, pol_draws=None,
                 pol_stream: int = 0):
        """source_cfg: SourceCfg / SinkCfg (or a list of them), which makes the
        population variable like state.rings_ids (ActiveState).  Random spawn
        polarisations: rng (a numpy Generator) feeds the device's pool of
        uniform draws, else the library's built-in stream keyed by seed;
        pol_draws (pool draws handed out but not used yet) and pol_stream
        (the built-in stream's position) resume a saved run."""
        d, self.chunks, self._tables = describe(state, space_cfg, dynamic_cfg, int_cfg, seed)
        self.state, self.space_cfg, self.dynamic_cfg, self.int_cfg = state, space_cfg, dynamic_cfg, int_cfg
        self.time_info = time_info or TimeInfo()
        self.ctx = ctx or _lib.Context.for_device(int_cfg.device.device)
        self.dev = _lib.DeviceSystem(self.ctx, d, keep=self._tables)
        self.p_neighbors_cfg = p_neighbors_cfg
        if p_neighbors_cfg is not None:
            if p_neighbors_cfg.type not in ("all", "rings"):
                raise ValueError(p_neighbors_cfg.type)
            self.dev.rings_neighbors_cfg(True, p_neighbors_cfg.tol, p_neighbors_cfg.type == "rings")
        self.forces = np.zeros_like(state.pos)
        extra = getattr(int_cfg, "extra", None)
        self.r_chunks_grid = None
        self.invasions_cfg = getattr(extra, "invasions_cfg", None)
        if getattr(extra, "r_chunks_cfg", None) is not None:
            self.r_chunks_grid = ring_chunks_grid(state, space_cfg, dynamic_cfg, extra.r_chunks_cfg)
        self.upload()
        self.sources = None
        self._rng = rng
        self._pool = np.zeros(0)  # draws handed to the device and not known to be used
        self._pool_base = 0       # pool index of _pool[0]
        if state.rings_ids is not None or source_cfg is not None:
            self._var_cfg(source_cfg)
            if pol_draws is not None and len(pol_draws):
                self._pool = np.asarray(pol_draws, dtype=np.float64)
                self.dev.rings_pol_draws(self._pool)
            if pol_stream:
                self.dev.rings_pol_stream_set(pol_stream)
        # the constructor's init sequence (src/rings/rings.jl:270-288): the
        # invasion schedule starts at this step count, then forces!
        self._invasions_cfg_call()
        self.forces[:] = self.dev.forces()
        self.areas, self.cms, self.continuos_pos = self.dev.rings_info()

    def _var_cfg(self, source_cfg):
        """RingsInfo's sources (src/rings/rings.jl:154-186) and the variable
        population on the device."""
        cfgs = [] if source_cfg is None else (list(source_cfg) if isinstance(source_cfg, (list, tuple))
                                              else [source_cfg])
        self.sources = []
        for c in cfgs:
            if isinstance(c, SourceCfg):
                self.sources.append(Source(c, self.chunks))
            elif isinstance(c, SinkCfg):
                self.sources.append(c)
            else:
                raise ValueError(f"Unknown source configuration type: {type(c).__name__}")
        st = self.state
        if st.rings_ids is None:
            st.rings_ids = VarRingsIds(np.ones(st.num_rings, bool), st.nmax, st.ring_num_particles)
        self._src_keep = []
        recs = source_records(self.sources, st.nmax, self._src_keep)
        self._random_boxes = sum(len(s.bbox) for s in self.sources
                                 if isinstance(s, Source) and isinstance(s.cfg.spawn_pol, str))
        self.dev.rings_var_cfg(st.rings_ids.mask.astype(np.int32), recs)

    def _pool_used(self):
        used = self.dev.rings_var_state()[4]
        self._pool = self._pool[used - self._pool_base:]
        self._pool_base = used

    def _supply_pol_draws(self, nsteps):
        """Keep the device's pool of uniform draws (from rng) ahead of the
        next nsteps: at most one draw per random spawn box and step."""
        if self._rng is None or not self._random_boxes:
            return
        self._pool_used()
        need = nsteps * self._random_boxes - len(self._pool)
        if need > 0:
            u = self._rng.random(need)
            self.dev.rings_pol_draws(u)
            self._pool = np.concatenate([self._pool, u])

    def pending_pol_draws(self) -> np.ndarray:
        """Pool draws handed to the device and not used yet."""
        if self.sources is None:
            return np.zeros(0)
        self._pool_used()
        return self._pool.copy()

    def _sync_var(self):
        if self.sources is None:
            return
        mask, ids, uids, na, _, _ = self.dev.rings_var_state()
        r = self.state.rings_ids
        r.mask[:] = mask != 0
        r.update()
        r.ids[:] = ids
        r.uids[:] = uids

    def _invasions_cfg_call(self):
        if self.invasions_cfg is None and self.r_chunks_grid is None:
            return
        cols, rows, nc = self.r_chunks_grid or (0, 0, 0)
        sched = self.invasions_cfg is not None
        self.dev.rings_invasions_cfg(sched, self.invasions_cfg.steps_to_update if sched else 0, cols, rows, nc)

    # --- invasions (InvasionsInfo.list, check_invasions!, src/rings/integration.jl:470-509)
    @property
    def invasions(self) -> list[Invasion]:
        """The last check's invasions, ordered by (invasor, invaded, p_id)."""
        if self.invasions_cfg is None and self.r_chunks_grid is None:
            return []
        return [Invasion(*map(int, r)) for r in self.dev.rings_invasions()]

    def check_invasions(self) -> list[Invasion]:
        """check_invasions!(system) at the last step's ring points."""
        if self.invasions_cfg is None and self.r_chunks_grid is None:
            self.r_chunks_grid = (0, 0, 0)
            self._invasions_cfg_call()
        self.dev.rings_invasions_check()
        return self.invasions

    def upload(self):
        self.dev.upload(self.state.pos, None, self.state.pol)

    def download(self):
        self.dev.download(self.state.pos, None, self.state.pol)
        self.forces[:] = self.dev.get_forces()
        self.areas, self.cms, self.continuos_pos = self.dev.rings_info()
        self._sync_var()

    def advance(self, nsteps: int, noise=None):
        if self.sources is not None and self._rng is not None and self._random_boxes:
            # the pool is topped up in segments, so that it stays small
            done = 0
            while done < nsteps:
                seg = min(256, int(nsteps) - done)
                self._supply_pol_draws(seg)
                if noise is None:
                    self.dev.step(seg)
                else:
                    self.dev.rings_step(seg, np.asarray(noise)[done:done + seg])
                done += seg
        elif noise is None:
            self.dev.step(nsteps)
        else:
            self.dev.rings_step(nsteps, noise)
        self.time_info.num_steps += int(nsteps)
        for _ in range(int(nsteps)):
            self.time_info.time += self.int_cfg.dt

    def calc_forces(self):
        self.forces[:] = self.dev.forces()
        return self.forces

    def update_chunks(self):
        cp, counts = self.dev.bin(export=True)
        self.chunks.chunk_particles = cp
        self.chunks.num_particles_in_chunk = counts
        return cp, counts

    # --- contact neighbours (get_neigh_count / get_neigh_list, src/rings/neighbors.jl:67-71)
    def get_neigh_count(self) -> np.ndarray:
        """Neighbour count per particle slot (0-based index) of the last forces!."""
        counts, _ = self.dev.rings_neighbors()
        return counts

    def get_neigh_list(self, pid: int) -> np.ndarray:
        """1-based ids of the neighbours of particle ``pid`` (1-based), ascending."""
        counts, lists = self.dev.rings_neighbors()
        return lists[pid - 1, :counts[pid - 1]]


def step(system: RingsSystem, noise=None):
    """Rings.Integration.step!; ``noise`` optionally supplies the nrings
    standard-normal draws of this step (randn(system.rng) per ring)."""
    system.upload()
    system.advance(1, noise)
    system.download()


# --- init states (src/rings/init_states.jl, src/rings/utils.jl) ------------
def get_ring_radius(p_radius, num_particles):
    return (p_radius * 2) / (2 * (1 - math.cos(2 * math.pi / num_particles))) ** 0.5


def create_circle(center, radius, num_p):
    theta = 2 * math.pi / num_p
    out = np.empty((num_p, 2))
    for i in range(num_p):
        out[i, 0] = radius * math.cos(i * theta) + center[0]
        out[i, 1] = radius * math.sin(i * theta) + center[1]
    return out


def rectangular_grid(*, num_cols, num_rows, num_particles, p_radius, types=None, pad_x=0.0, pad_y=0.0,
                     radius_k=1.0):
    """Ring lattice, src/rings/init_states.jl:27-72; returns (rings_pos, geometry_cfg)."""
    if types is None:
        num_particles = [num_particles]
        p_radius = [p_radius]
        types = [1] * (num_cols * num_rows)
    ring_r = [get_ring_radius(pr, n) for pr, n in zip(p_radius, num_particles)]
    ring_length = [2 * (rr + pr) for rr, pr in zip(ring_r, p_radius)]
    max_ring_r, max_ring_length, max_np = max(ring_r), max(ring_length), max(num_particles)
    pad_x, pad_y = pad_x * max_ring_r, pad_y * max_ring_r
    rings_pos = np.zeros((num_cols * num_rows, max_np, 2))
    idx = 0
    for col_id in range(num_cols):
        for row_id in range(num_rows):
            cx = pad_x / 2 + col_id * (max_ring_length + pad_x) + max_ring_length / 2
            cy = pad_y / 2 + row_id * (max_ring_length + pad_x) + max_ring_length / 2
            t = types[idx]
            n = num_particles[t - 1]
            rings_pos[idx, :n] = create_circle((cx, cy), ring_r[t - 1] * radius_k, n)
            idx += 1
    geometry_cfg = RectangleCfg(length=num_cols * (pad_x + max_ring_length),
                                height=num_rows * (pad_y + max_ring_length))
    return rings_pos, geometry_cfg


def random_pol(num_rings, rng: np.random.Generator):
    return rng.random(num_rings) * 2 * math.pi
