"""maglua_amd: MagLua's per-timestep spin-dynamics hot path on MI355X (gfx950).

The compute lives in libmsd.so (hand-written HIP kernels behind the C ABI of
include/msd_abi.h); this package is the host-side mirror of MagLua's Lua
object API over that ABI.
"""
from ._lib import MagLuaError, declared_symbols, lib  # noqa: F401
from .checkpoint import (  # noqa: F401
    Checkpointer, checkpointFromString, checkpointLoad, checkpointSave, checkpointToString,
)
from .maglua import (  # noqa: F401
    BUTCHER, LLG, AppliedField, Anisotropy, Dipole, Dipole2D, Ensemble, Exchange, Heun, LongRange, LongRange2D,
    LongRange3D, Magnetostatics2D, Random, SpinSystem,
    Surrogate, Thermal, make_adapt_step_function, make_pc_step_function, make_rk_step_function, philox_device,
    synchronize,
)

__all__ = [
    "MagLuaError", "SpinSystem", "Exchange", "Anisotropy", "AppliedField", "Thermal", "Dipole", "LongRange",
    "LongRange3D", "LongRange2D", "Dipole2D", "Magnetostatics2D",
    "LLG", "Random", "Heun", "Ensemble", "Surrogate", "make_pc_step_function", "make_rk_step_function",
    "make_adapt_step_function", "BUTCHER",
    "synchronize", "philox_device", "lib", "declared_symbols",
    "Checkpointer", "checkpointSave", "checkpointLoad", "checkpointToString", "checkpointFromString",
]
