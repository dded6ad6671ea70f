"""MI355X-native pressure projection for 2D_Engine (PROJECTION_2D + its Poisson solve).

The compute path is the HIP library libp2d.so (C ABI in include/p2d.h); this
package is its Python binding, a host-side mirror of the reference's
PROJECTION_2D interface, and the seeded workload generators.  Import it with
importlib.import_module("2d_engine_amd") (the name starts with a digit).
"""
try:  # torch's HIP runtime, when present, must be the first one loaded (see _native)
    import torch  # noqa: F401
except ImportError:
    pass
from . import workloads  # noqa: F401
from ._native import (FP64, MIXED, PREC_F32, PREC_F64, TIME_REINIT_STEP, VISC_DIRICHLET, VISC_REFERENCE,  # noqa: F401
                      Context, P2DError, ViscositySolver, lib, status_string)
from .projection import PROJECTION_2D  # noqa: F401
