"""Mirror of Mavi.Configs for the GPU path (src/configs.jl).

Names, fields and argument meanings follow the reference; only the types the
GPU path runs are provided: rectangles, circles, lines and lists of them
(ManyGeometries) with RigidWalls, PeriodicWalls, SlipperyWalls,
PotentialWalls and ManyWalls; LenJonesCfg, HarmTruncCfg, SzaboCfg and
RunTumbleCfg; IntCfg/ChunksCfg and the device modes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


# --- geometry (src/configs.jl:34-51) --------------------------------------
@dataclass
class RectangleCfg:
    length: float
    height: float
    bottom_left: tuple = (0.0, 0.0)

    def __post_init__(self):
        self.length = float(self.length)
        self.height = float(self.height)
        self.bottom_left = (float(self.bottom_left[0]), float(self.bottom_left[1]))

    @property
    def size(self):
        return (self.length, self.height)


@dataclass
class CircleCfg:  # src/configs.jl:145-152
    radius: float
    center: tuple = (0.0, 0.0)

    def __post_init__(self):
        self.radius = float(self.radius)
        self.center = (float(self.center[0]), float(self.center[1]))


@dataclass
class Line2D:  # src/configs.jl:92-117
    p1: tuple
    p2: tuple

    def __post_init__(self):
        self.p1 = (float(self.p1[0]), float(self.p1[1]))
        self.p2 = (float(self.p2[0]), float(self.p2[1]))


@dataclass
class LinesCfg:  # src/configs.jl:137-143
    lines: list
    bbox: RectangleCfg | None = None

    def __post_init__(self):
        self.lines = [ln if isinstance(ln, Line2D) else Line2D(*ln) for ln in self.lines]


@dataclass
class ManyGeometries:  # src/configs.jl:30-32
    list: tuple


def get_bounding_box(g) -> RectangleCfg:
    """src/configs.jl:186-225."""
    if isinstance(g, RectangleCfg):
        return g
    if isinstance(g, CircleCfg):
        r = g.radius
        return RectangleCfg(length=2 * r, height=2 * r, bottom_left=(g.center[0] - r, g.center[1] - r))
    if isinstance(g, LinesCfg):
        if g.bbox is not None:
            return g.bbox
        xs = [v for ln in g.lines for v in (ln.p1[0], ln.p2[0])]
        ys = [v for ln in g.lines for v in (ln.p1[1], ln.p2[1])]
        return RectangleCfg(length=max(xs) - min(xs), height=max(ys) - min(ys), bottom_left=(min(xs), min(ys)))
    if isinstance(g, ManyGeometries):
        box = get_bounding_box(g.list[0])
        for h in g.list[1:]:
            b = get_bounding_box(h)  # RectangleCfg + RectangleCfg, src/configs.jl:70-79
            max_x = max(box.bottom_left[0] + box.length, b.bottom_left[0] + b.length)
            max_y = max(box.bottom_left[1] + box.height, b.bottom_left[1] + b.height)
            min_x = min(box.bottom_left[0], b.bottom_left[0])
            min_y = min(box.bottom_left[1], b.bottom_left[1])
            box = RectangleCfg(length=max_x - min_x, height=max_y - min_y, bottom_left=(min_x, min_y))
        return box
    raise TypeError(type(g))


# --- walls (src/configs.jl:235-283) ---------------------------------------
class WallType:
    pass


class RigidWalls(WallType):
    def __repr__(self):
        return "RigidWalls()"


class PeriodicWalls(WallType):
    def __repr__(self):
        return "PeriodicWalls()"


class SlipperyWalls(WallType):
    def __repr__(self):
        return "SlipperyWalls()"


class PotentialWalls(WallType):
    """A wall pushing with a PotentialCfg (LenJonesCfg or HarmTruncCfg);
    mode "outside", "inside" or "repulsion" (src/configs.jl:252-283)."""

    def __init__(self, potential, mode: str = "repulsion"):
        if mode not in ("outside", "inside", "repulsion"):
            raise ValueError(mode)
        self.potential = potential
        self.mode = mode

    def __repr__(self):
        return f"PotentialWalls({self.potential!r}, {self.mode!r})"


@dataclass
class ManyWalls(WallType):  # src/configs.jl:242-244
    list: tuple


@dataclass
class SpaceCfg:  # src/configs.jl:285-301
    wall_type: WallType
    geometry_cfg: object

    @classmethod
    def from_list(cls, pairs):
        """SpaceCfg(spaces_cfg_list): ManyWalls / ManyGeometries of (wall, geometry) pairs."""
        walls, geoms = zip(*pairs)
        return cls(wall_type=ManyWalls(tuple(walls)), geometry_cfg=ManyGeometries(tuple(geoms)))

    def pairs(self):
        if isinstance(self.wall_type, ManyWalls):
            return list(zip(self.wall_type.list, self.geometry_cfg.list))
        return [(self.wall_type, self.geometry_cfg)]

    def main(self):
        """get_main_wall / get_main_geometry, src/configs.jl:303-311."""
        return self.pairs()[0]


# --- dynamics -------------------------------------------------------------
class DynamicCfg:
    pass


class PotentialCfg(DynamicCfg):
    pass


@dataclass
class HarmTruncCfg(PotentialCfg):  # src/configs.jl:341-368
    k_rep: float
    k_atr: float
    dist_eq: float
    dist_max: float


@dataclass
class LenJonesCfg(PotentialCfg):  # src/configs.jl:380-397
    """Lennard-Jones.  ``cutoff`` is not a reference field: None (default) is
    the reference's LenJonesCfg, whose range is set by the chunk stencil only;
    a number truncates the force at that distance (zero beyond it), the
    truncated-LJ form the reference uses for RunTumbleCfg."""

    sigma: float
    epsilon: float
    cutoff: float | None = None


@dataclass
class SzaboCfg(DynamicCfg):  # src/configs.jl:399-408
    """Self-propelled particles of Szabo et al.: linear repulsion/adhesion
    between r_eq and r_max (calc_interaction, src/integration.jl:68-86) and
    polarisations relaxing towards the velocity (update_szabo!, :433-465)."""

    vo: float
    mobility: float
    relax_time: float
    k_rep: float
    k_adh: float
    r_eq: float
    r_max: float
    rot_diff: float


@dataclass
class RunTumbleCfg(DynamicCfg):  # src/configs.jl:410-415
    """Run-and-tumble particles: LJ repulsion truncated at 2^(1/6) sigma
    (calc_interaction, src/integration.jl:88-109) and tumbles at rate
    tumble_rate (update_rtp!, :467-489)."""

    vo: float
    sigma: float
    epsilon: float
    tumble_rate: float

    def cutoff(self) -> float:
        return 2 ** (1 / 6) * self.sigma  # src/integration.jl:96


def particle_radius(dynamic_cfg) -> float:
    """src/configs.jl:418-421."""
    if isinstance(dynamic_cfg, LenJonesCfg):
        return dynamic_cfg.sigma * 2 ** (1 / 6) / 2
    if isinstance(dynamic_cfg, HarmTruncCfg):
        return dynamic_cfg.dist_eq / 2.0
    if isinstance(dynamic_cfg, SzaboCfg):
        return dynamic_cfg.r_eq / 2
    if isinstance(dynamic_cfg, RunTumbleCfg):
        return dynamic_cfg.sigma * 2 ** (1 / 6) / 2
    if hasattr(dynamic_cfg, "particle_radius"):
        return dynamic_cfg.particle_radius()
    raise TypeError(f"no particle_radius for {type(dynamic_cfg).__name__}")


# --- integration (src/configs.jl:471-487) ---------------------------------
class DeviceMode:
    pass


class Sequencial(DeviceMode):
    pass


class Threaded(DeviceMode):
    pass


class GPU(DeviceMode):
    """The MI355X device mode (libmavi_hip)."""

    def __init__(self, device: int = 0):
        self.device = device


@dataclass
class ChunksCfg:
    num_cols: int
    num_rows: int


@dataclass
class IntCfg:
    dt: float
    chunks_cfg: ChunksCfg | None = None
    device: DeviceMode = field(default_factory=GPU)
    extra: object = None


def chunk_capacity(chunk_length: float, chunk_height: float, particle_r: float) -> int:
    """Chunk capacity nc, src/chunks.jl:32-34."""
    nc = (math.ceil(0.5 * chunk_length / particle_r) + 1) * (math.ceil(0.5 * chunk_height / particle_r) + 1)
    return int(math.trunc(math.ceil(nc * 2)))


def as_points(a) -> np.ndarray:
    """Vector{SVector{2,Float64}} as an (N, 2) float64 array (x0,y0,x1,y1,...)."""
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if arr.ndim == 2 and arr.shape[1] == 2:
        return arr
    if arr.ndim == 2 and arr.shape[0] == 2:  # Julia Matrix 2xN
        return np.ascontiguousarray(arr.T)
    raise ValueError("positions must be (N, 2) or (2, N)")
