"""MI355X-native RKDG solver for the 2-D Euler equations on unstructured triangles.

The numerical work runs in librkdg.so (hand-written gfx950 HIP kernels behind a C ABI,
include/rkdg.h); this package is a thin ctypes host layer over it.
"""
from . import _lib
from ._lib import (BASIS_ELEMENT, BASIS_NATIVE, BC_FARFIELD, BC_OUTFLOW, BC_REFLECTION, FLUX_LF, FLUX_ROE,
                   IC_FREESTREAM, IC_RIEMANN4, IC_VORTEX, RkdgError)
from .solver import Batch, Grid, Solver, freestream_state, make_params, mesh_view, riemann4_params, vortex_params

__all__ = ["Batch", "Grid", "Solver", "RkdgError", "freestream_state", "make_params", "mesh_view", "riemann4_params",
           "vortex_params", "BASIS_ELEMENT", "BASIS_NATIVE", "BC_FARFIELD", "BC_OUTFLOW", "BC_REFLECTION",
           "FLUX_LF", "FLUX_ROE", "IC_FREESTREAM", "IC_VORTEX", "IC_RIEMANN4"]
