"""MI355X-native TTM2.1-F / TTM3-F force path and ring-polymer propagation.

The compute lives in libpimd_ttm.so (HIP kernels for gfx950 behind the C ABI
of include/pimd_ttm.h); this package is the Python host mirror of the
reference's force_calc / NormalModes / Langevin interface.
"""
from .params import TTMParams, TTM21F, TTM3F, QSPCFW, SPCF, f32  # noqa: F401
from . import systems  # noqa: F401


def __getattr__(name):
    # The ctypes layer is imported lazily so that systems/params work without
    # the built library (e.g. for generating inputs).
    if name in ("TTMForceField", "RingPolymer", "PIMDIntegrator", "nose_hoover", "load_library",
                "PimdError", "EXPORTS", "LIB_PATH"):
        from . import ttm
        return getattr(ttm, name)
    raise AttributeError(name)
