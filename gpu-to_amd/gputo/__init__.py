"""gputo -- MI355X-native top3D125 MGCG elasticity solver (host side).

The compute path is libgputo.so (HIP kernels for gfx950 behind a C ABI,
include/gputo.h).  This package is the host-side mirror of the reference's
MATLAB-facing interface: `mgcg1` keeps the MEX argument list and semantics of
Cantilever/TOP3D125MGCGCUDA/mgcg1.cu (and the omp variant's WRK form of
Cantilever/omp/mgcg1.c), `Solver` exposes the handle, and `TopOpt` runs the
top3D125mod.m design loop on the device.
"""
from ._lib import (GputoError, PREC_F64, PREC_MIXED_BF16, PREC_MIXED_F32, PRESET_MATLAB_CFG0, PRESET_MGCG9, PRESET_REF_C,
                   load)
from .mgcg import Solver, mgcg1, mgcg9, release_handles
from .problems import cantilever, fixed_x0_face, ke0_h8, uniform_edge_load, matlab_edge_load
from .topopt import TopOpt
from . import log

__all__ = [
    "GputoError", "PREC_F64", "PREC_MIXED_F32", "PREC_MIXED_BF16", "PRESET_REF_C", "PRESET_MATLAB_CFG0", "PRESET_MGCG9", "load",
    "Solver", "mgcg1", "mgcg9", "release_handles", "cantilever", "fixed_x0_face", "ke0_h8", "uniform_edge_load",
    "matlab_edge_load", "TopOpt", "log",
]
