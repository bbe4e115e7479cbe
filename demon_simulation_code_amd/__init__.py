"""MI355X-native implementation of DEMON's hot path (Yukawa all-pairs force +
RK4 stepping) behind a C ABI (include/demon_gpu.h).

The compute runs in libdemon_gpu.so (hand-written gfx950 HIP); this package
only binds it.  Importing fails loudly when the library is not built.
"""
from ._lib import (  # noqa: F401
    Batch,
    CONFINEMENT,
    DRAG,
    FP32_PAIR,
    FP64,
    LIB_PATH,
    BATCH_NEWTON3,
    PAIRS_GATHER,
    PAIRS_NEWTON3,
    MAGNETIC,
    SHIELDED_COULOMB,
    THERMAL,
    TIME_VARYING_DRAG,
    GRAVITATIONAL,
    THERMAL_LOCALIZED,
    DRIVING,
    ROTATIONAL,
    TIME_VARYING_THERMAL,
    CONFINEMENT_VOID,
    ELECTRIC,
    RECT_CONFINEMENT,
    VERT_ELECTRIC,
    PLUGIN_PARAMS,
    Context,
    DemonError,
    Params,
    Stats,
    hex_lattice,
    lib,
    trajectory_spectrum,
)
from .physics import ELECTRON_CHARGE, EPSILON0, COULOMB, confine_from_k  # noqa: F401

__all__ = [
    "Batch", "Context", "DemonError", "Params", "Stats", "hex_lattice", "lib", "LIB_PATH", "trajectory_spectrum",
    "CONFINEMENT", "DRAG", "SHIELDED_COULOMB", "THERMAL", "TIME_VARYING_DRAG", "MAGNETIC", "FP64", "FP32_PAIR",
    "BATCH_NEWTON3", "PAIRS_GATHER", "PAIRS_NEWTON3", "GRAVITATIONAL", "THERMAL_LOCALIZED", "DRIVING",
    "ROTATIONAL", "TIME_VARYING_THERMAL", "CONFINEMENT_VOID", "ELECTRIC", "RECT_CONFINEMENT", "VERT_ELECTRIC",
    "PLUGIN_PARAMS",
    "ELECTRON_CHARGE", "EPSILON0", "COULOMB", "confine_from_k",
]
