"""Mirror of Mavi.Configs for the GPU path (src/configs.jl).

Names, fields and argument meanings follow the reference; only the types the
GPU path runs are provided: rectangles with RigidWalls or PeriodicWalls,
LenJonesCfg, HarmTruncCfg, SzaboCfg and RunTumbleCfg, IntCfg/ChunksCfg and
the device modes.
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


# --- walls (src/configs.jl:238-239) ---------------------------------------
class WallType:
    pass


class RigidWalls(WallType):
    def __repr__(self):
        return "RigidWalls()"


class PeriodicWalls(WallType):
    def __repr__(self):
        return "PeriodicWalls()"


@dataclass
class SpaceCfg:  # src/configs.jl:285-288
    wall_type: WallType
    geometry_cfg: RectangleCfg


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
