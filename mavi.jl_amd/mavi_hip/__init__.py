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
    HarmTruncCfg,
    IntCfg,
    LenJonesCfg,
    PeriodicWalls,
    RectangleCfg,
    RigidWalls,
    Sequencial,
    SpaceCfg,
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
    run_system,
    update_chunks,
)
from .systems import SecondLawState, System, TimeInfo

__all__ = [
    "GPU", "ChunksCfg", "HarmTruncCfg", "IntCfg", "LenJonesCfg", "PeriodicWalls", "RectangleCfg",
    "RigidWalls", "Sequencial", "SpaceCfg", "Threaded", "chunk_capacity", "particle_radius",
    "calc_forces", "get_step_function", "kinetic_energy", "newton_step", "potential_energy",
    "run_system", "update_chunks", "SecondLawState", "System", "TimeInfo", "Context",
    "MaviError", "CellOverflowError", "OutOfBoxError", "UnsupportedOnGPU",
]
