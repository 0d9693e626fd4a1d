"""mavi_hip -- MI355X (gfx950) step hot path of Mavi.jl behind Mavi's own API.

The compute lives in libmavi_hip.so (hand-written HIP kernels, C ABI in
include/mavi_hip.h); this package is the host-side mirror of the reference
interface used to drive it.
"""
from . import _lib
from ._lib import (
    CellOverflowError,
    Context,
    MaviError,
    OutOfBoxError,
    UnsupportedOnGPU,
)
from .configs import (
    GPU,
    ChunksCfg,
    CircleCfg,
    Line2D,
    LinesCfg,
    ManyGeometries,
    ManyWalls,
    PotentialWalls,
    SlipperyWalls,
    get_bounding_box,
    HarmTruncCfg,
    IntCfg,
    LenJonesCfg,
    PeriodicWalls,
    RectangleCfg,
    RigidWalls,
    RunTumbleCfg,
    Sequencial,
    SpaceCfg,
    SzaboCfg,
    Threaded,
    chunk_capacity,
    particle_radius,
)
from .integration import (
    calc_forces,
    get_step_function,
    kinetic_energy,
    newton_step,
    potential_energy,
    rtp_step,
    run_system,
    szabo_step,
    update_chunks,
)
from .systems import SecondLawState, SelfPropelledState, System, TimeInfo

__all__ = [
    "GPU", "ChunksCfg", "HarmTruncCfg", "IntCfg", "LenJonesCfg", "PeriodicWalls", "RectangleCfg",
    "RigidWalls", "Sequencial", "SpaceCfg", "Threaded", "chunk_capacity", "particle_radius",
    "calc_forces", "get_step_function", "kinetic_energy", "newton_step", "potential_energy",
    "run_system", "update_chunks", "SecondLawState", "System", "TimeInfo", "Context",
    "SzaboCfg", "RunTumbleCfg", "SelfPropelledState", "szabo_step", "rtp_step",
    "CircleCfg", "Line2D", "LinesCfg", "ManyGeometries", "ManyWalls", "PotentialWalls", "SlipperyWalls",
    "get_bounding_box",
    "MaviError", "CellOverflowError", "OutOfBoxError", "UnsupportedOnGPU",
]
