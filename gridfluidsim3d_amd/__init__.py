"""MI355X-native (gfx950) FLIP velocity transfer and advection for
GridFluidSim3D: deterministic P2G gather, trilinear/tricubic G2P, RK1-4
advection and the fused plain FLIP step, behind the C ABI in include/gfsx.h.

The HIP library (lib/libgfsx.so) is required; there is no CPU fallback.
"""
from .gfsx import (EPILOGUE_FLUIDSIM, EPILOGUE_NONE, Context, FlipState, FlipTape,
                   ORDER_KEEP, ORDER_SORTED, TRICUBIC, TRILINEAR, mac_shapes)
from .pybindings import LibraryLoadError, library_path, load_library

__all__ = ["Context", "FlipState", "FlipTape", "EPILOGUE_FLUIDSIM", "EPILOGUE_NONE", "ORDER_KEEP",
           "ORDER_SORTED", "TRICUBIC",
           "TRILINEAR", "mac_shapes", "LibraryLoadError", "library_path",
           "load_library"]
