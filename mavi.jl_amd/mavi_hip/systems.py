"""Mirror of Mavi.States / Mavi.Systems / Mavi.ChunksMod for the GPU path.

System(...) (src/systems.jl:73-114) checks that every particle is inside the
geometry, builds the chunk grid (src/systems.jl:14-28, src/chunks.jl:26-40),
creates the device system and uploads the state.  State arrays live on the
host as (N, 2) float64 arrays, the layout of Vector{SVector{2,Float64}}.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .configs import (
    GPU,
    CircleCfg,
    HarmTruncCfg,
    LenJonesCfg,
    LinesCfg,
    ManyWalls,
    PeriodicWalls,
    PotentialWalls,
    RectangleCfg,
    RigidWalls,
    RunTumbleCfg,
    SlipperyWalls,
    SpaceCfg,
    SzaboCfg,
    as_points,
    chunk_capacity,
    get_bounding_box,
    particle_radius,
)


@dataclass
class SecondLawState:  # src/states.jl:233-259
    pos: np.ndarray
    vel: np.ndarray

    def __post_init__(self):
        self.pos = as_points(self.pos).copy()
        self.vel = as_points(self.vel).copy()
        if self.pos.shape != self.vel.shape:
            raise ValueError("pos and vel must have the same number of particles")


@dataclass
class SelfPropelledState:  # src/states.jl:104-125
    pos: np.ndarray
    pol_angle: np.ndarray

    def __post_init__(self):
        self.pos = as_points(self.pos).copy()
        self.pol_angle = np.array(self.pol_angle, dtype=np.float64).reshape(-1).copy()
        if len(self.pol_angle) != len(self.pos):
            raise ValueError("pos and pol_angle must have the same number of particles")


@dataclass
class TimeInfo:  # src/systems.jl:30-33
    num_steps: int = 0
    time: float = 0.0


class Chunks:
    """Chunk grid geometry (src/chunks.jl:10-40); the chunk arrays themselves
    live on the device and are exported on request (update_chunks!)."""

    def __init__(self, num_cols, num_rows, geometry_cfg: RectangleCfg, particle_r: float):
        self.num_cols = int(num_cols)
        self.num_rows = int(num_rows)
        self.chunk_length = geometry_cfg.length / self.num_cols
        self.chunk_height = geometry_cfg.height / self.num_rows
        self.nc = chunk_capacity(self.chunk_length, self.chunk_height, particle_r)
        self.geometry_cfg = geometry_cfg
        self._dev = None
        self.chunk_particles = None
        self.num_particles_in_chunk = None


def check_inside(pos: np.ndarray, geometry_cfg, ids=None):
    """src/space_checks.jl:9-60: rectangles by their box, circles by the
    distance from the origin (as the reference writes it), other geometries
    unchecked."""
    if isinstance(geometry_cfg, CircleCfg):
        p = pos if ids is None else pos[ids]
        out = np.nonzero(np.sum(p * p, axis=1) > geometry_cfg.radius ** 2)[0]
        if ids is not None:
            out = np.asarray(ids)[out]
        return len(out) == 0, out
    if not isinstance(geometry_cfg, RectangleCfg):
        return True, np.zeros(0, dtype=int)
    bl = np.asarray(geometry_cfg.bottom_left)
    tr = bl + np.asarray((geometry_cfg.length, geometry_cfg.height))
    p = pos if ids is None else pos[ids]
    out = np.nonzero(np.any(p < bl, axis=1) | np.any(p > tr, axis=1))[0]
    if ids is not None:
        out = np.asarray(ids)[out]
    return len(out) == 0, out


def gpu_kind(dynamic_cfg):
    """The trait that routes a DynamicCfg to the GPU: built-in potentials only."""
    if isinstance(dynamic_cfg, LenJonesCfg):
        return _lib.MAVI_POT_LJ
    if isinstance(dynamic_cfg, HarmTruncCfg):
        return _lib.MAVI_POT_HARMTRUNC
    if isinstance(dynamic_cfg, SzaboCfg):
        return _lib.MAVI_POT_SZABO
    if isinstance(dynamic_cfg, RunTumbleCfg):
        return _lib.MAVI_POT_RTP
    return None


def wall_kind(space_cfg: SpaceCfg):
    """Main wall as the descriptor's `wall` (stencil and calc_diff)."""
    w, g = space_cfg.main()
    if isinstance(w, PeriodicWalls) and isinstance(g, RectangleCfg):
        return _lib.MAVI_WALL_PERIODIC
    if isinstance(w, (RigidWalls, SlipperyWalls, PotentialWalls)):
        return _lib.MAVI_WALL_RIGID
    return None


def _pot_record(potential):
    if isinstance(potential, LenJonesCfg):
        return _lib.MAVI_POT_LJ, (potential.sigma, potential.epsilon, 0.0, 0.0)
    if isinstance(potential, HarmTruncCfg):
        return _lib.MAVI_POT_HARMTRUNC, (potential.k_rep, potential.k_atr, potential.dist_eq, potential.dist_max)
    raise _lib.UnsupportedOnGPU(_lib.MAVI_E_UNSUPPORTED, "potential walls take a LenJonesCfg or HarmTruncCfg")


def wall_records(space_cfg: SpaceCfg):
    """The walls list of a SpaceCfg as plain records {kind, mode, pot_kind,
    geom, pot, lines} (the mavi_wall fields), or None for a single rigid or
    periodic rectangle (the descriptor's `wall`)."""
    if not isinstance(space_cfg.wall_type, ManyWalls) and isinstance(space_cfg.geometry_cfg, RectangleCfg) and \
            isinstance(space_cfg.wall_type, (RigidWalls, PeriodicWalls)):
        return None
    out = []
    for w, g in space_cfg.pairs():
        rec = dict(mode=0, pot_kind=0, geom=(0.0, 0.0, 0.0, 0.0), pot=(0.0, 0.0, 0.0, 0.0), lines=None)
        if isinstance(g, RectangleCfg) and isinstance(w, (RigidWalls, PeriodicWalls)):
            rec["kind"] = _lib.MAVI_WK_RIGID_RECT if isinstance(w, RigidWalls) else _lib.MAVI_WK_PERIODIC_RECT
            rec["geom"] = (g.bottom_left[0], g.bottom_left[1], g.length, g.height)
        elif isinstance(g, CircleCfg) and isinstance(w, (RigidWalls, SlipperyWalls, PotentialWalls)):
            rec["kind"] = {RigidWalls: _lib.MAVI_WK_RIGID_CIRCLE, SlipperyWalls: _lib.MAVI_WK_SLIPPERY_CIRCLE,
                           PotentialWalls: _lib.MAVI_WK_POTENTIAL_CIRCLE}[type(w)]
            rec["geom"] = (g.center[0], g.center[1], g.radius, 0.0)
        elif isinstance(g, LinesCfg) and isinstance(w, (SlipperyWalls, PotentialWalls)):
            rec["kind"] = _lib.MAVI_WK_SLIPPERY_LINES if isinstance(w, SlipperyWalls) else _lib.MAVI_WK_POTENTIAL_LINES
            rec["lines"] = np.array([[*ln.p1, *ln.p2] for ln in g.lines], dtype=np.float64)
        else:
            raise _lib.UnsupportedOnGPU(_lib.MAVI_E_UNSUPPORTED,
                                        f"{type(w).__name__} with {type(g).__name__} is not on the GPU path")
        if isinstance(w, PotentialWalls):
            rec["mode"] = _lib.MAVI_WMODE[w.mode]
            rec["pot_kind"], rec["pot"] = _pot_record(w.potential)
        out.append(rec)
    return out


def attach_walls(d: _lib.SysDesc, records):
    """Point the descriptor at a ctypes copy of the records; returns what
    must stay alive until mavi_sys_create returns."""
    import ctypes as C

    if not records:
        return ()
    arr = (_lib.MaviWall * len(records))()
    keep = [arr]
    for k, r in enumerate(records):
        arr[k].kind, arr[k].mode, arr[k].pot_kind = r["kind"], r["mode"], r["pot_kind"]
        for j in range(4):
            arr[k].geom[j] = r["geom"][j]
            arr[k].pot[j] = r["pot"][j]
        if r["lines"] is not None:
            ln = np.ascontiguousarray(r["lines"])
            arr[k].nlines = len(ln)
            arr[k].lines = ln.ctypes.data_as(C.POINTER(C.c_double))
            keep.append(ln)
    d.nwalls = len(records)
    d.walls = C.cast(arr, C.POINTER(_lib.MaviWall))
    return tuple(keep)


def base_desc(n, space_cfg, dynamic_cfg, int_cfg, chunks: Chunks | None, radius) -> _lib.SysDesc:
    g = get_bounding_box(space_cfg.geometry_cfg)  # the chunk space (get_chunks, src/systems.jl:19)
    d = _lib.SysDesc()
    d.n = int(n)
    d.bottom_left[0], d.bottom_left[1] = g.bottom_left
    d.size[0], d.size[1] = g.length, g.height
    d.wall = wall_kind(space_cfg)
    d.dt = float(int_cfg.dt)
    if chunks is not None:
        d.cols, d.rows, d.nc = chunks.num_cols, chunks.num_rows, chunks.nc
    d.radius = float(radius)
    return d


def describe(state, space_cfg: SpaceCfg, dynamic_cfg, int_cfg):
    """Descriptor of a plain system for the C ABI, with the checks of
    System(...) (src/systems.jl:73-114): particles inside the geometry, a
    GPU-capable potential and walls.  Returns (SysDesc, Chunks | None)."""
    ok, out_ids = check_inside(state.pos, space_cfg.geometry_cfg)
    if not ok:
        raise ValueError(f"Particles with ids={list(out_ids + 1)} outside space.")
    kind = gpu_kind(dynamic_cfg)
    wall = wall_kind(space_cfg)
    active = kind in (_lib.MAVI_POT_SZABO, _lib.MAVI_POT_RTP)
    if active != isinstance(state, SelfPropelledState):
        raise TypeError(f"{type(dynamic_cfg).__name__} runs with a "
                        f"{'SelfPropelledState' if active else 'SecondLawState'}")
    if kind is None or wall is None or not isinstance(int_cfg.device, GPU):
        raise _lib.UnsupportedOnGPU(
            _lib.MAVI_E_UNSUPPORTED,
            f"{type(dynamic_cfg).__name__} / {type(space_cfg.wall_type).__name__} / "
            f"{type(int_cfg.device).__name__} is not on the GPU path; Mavi's own CPU "
            "methods run it (the Julia integration dispatches there)",
        )
    r = particle_radius(dynamic_cfg)
    chunks = None
    if int_cfg.chunks_cfg is not None:
        chunks = Chunks(int_cfg.chunks_cfg.num_cols, int_cfg.chunks_cfg.num_rows,
                        get_bounding_box(space_cfg.geometry_cfg), r)
    d = base_desc(len(state.pos), space_cfg, dynamic_cfg, int_cfg, chunks, r)
    d._walls_keep = attach_walls(d, wall_records(space_cfg))
    d.potential = kind
    c = dynamic_cfg
    if kind == _lib.MAVI_POT_LJ:
        d.pot[0], d.pot[1] = c.sigma, c.epsilon
        d.pot[2] = c.cutoff if c.cutoff else 0.0
    elif kind == _lib.MAVI_POT_HARMTRUNC:
        d.pot[0], d.pot[1] = c.k_rep, c.k_atr
        d.pot[2], d.pot[3] = c.dist_eq, c.dist_max
    elif kind == _lib.MAVI_POT_SZABO:
        d.pot[0], d.pot[1], d.pot[2], d.pot[3] = c.k_rep, c.k_adh, c.r_eq, c.r_max
        d.active[0], d.active[1], d.active[2], d.active[3] = c.vo, c.mobility, c.relax_time, c.rot_diff
    else:
        d.pot[0], d.pot[1] = c.sigma, c.epsilon
        d.active[0], d.active[1] = c.vo, c.tumble_rate
    return d, chunks


class System:
    """GPU-backed System (src/systems.jl:45-114)."""

    def __init__(self, *, state, space_cfg: SpaceCfg, dynamic_cfg, int_cfg,
                 time_info: TimeInfo | None = None, ctx: _lib.Context | None = None, seed: int = 0):
        d, self.chunks = describe(state, space_cfg, dynamic_cfg, int_cfg)
        d.seed = int(seed)  # device noise stream of self-propelled systems
        self.state = state
        self.space_cfg = space_cfg
        self.dynamic_cfg = dynamic_cfg
        self.int_cfg = int_cfg
        self.time_info = time_info or TimeInfo()
        self.ctx = ctx or _lib.Context.for_device(int_cfg.device.device)
        self.dev = _lib.DeviceSystem(self.ctx, d)
        self.forces = np.zeros_like(state.pos)
        self.upload()
        if self.chunks is not None:
            self.chunks._dev = self.dev
            self.update_chunks()

    # --- host <-> device ---------------------------------------------------
    @property
    def self_propelled(self) -> bool:
        return isinstance(self.state, SelfPropelledState)

    def upload(self):
        if self.self_propelled:
            self.dev.upload(self.state.pos, None, self.state.pol_angle)
        else:
            self.dev.upload(self.state.pos, self.state.vel)

    def download(self):
        if self.self_propelled:
            self.dev.download(self.state.pos, None, self.state.pol_angle)
        else:
            self.dev.download(self.state.pos, self.state.vel)
        self.forces[:] = self.dev.get_forces()

    # --- reference operations ----------------------------------------------
    def update_chunks(self):
        """update_chunks!(system.chunks): exports the reference arrays."""
        cp, counts = self.dev.bin(export=True)
        self.chunks.chunk_particles = cp
        self.chunks.num_particles_in_chunk = counts
        return cp, counts

    def calc_forces(self):
        """clean_forces! + update_chunks! + calc_forces!(system)."""
        self.forces[:] = self.dev.forces()
        return self.forces

    def get_forces(self):
        return self.forces

    def get_num_total_particles(self):
        return len(self.state.pos)

    def advance(self, nsteps: int, noise=None):
        """nsteps steps of the system's dynamics (newton_step!, or
        szabo_step! / rtp_step! for a SelfPropelledState) with the state
        resident on the device.  noise (self-propelled only) replaces the
        device stream: (nsteps, n) normals for Szabo, (nsteps, n, 2)
        uniforms for run-and-tumble."""
        if noise is not None:
            if not self.self_propelled:
                raise ValueError("host noise is for self-propelled systems")
            self.dev.sp_step(nsteps, noise)
        else:
            self.dev.step(nsteps)
        self.time_info.num_steps += int(nsteps)
        for _ in range(int(nsteps)):
            self.time_info.time += self.int_cfg.dt
