"""MI355X-native 2D TMz FDTD stepper (HIP kernels for gfx950 behind a C ABI).

The stepper itself is native (fdtd2d_amd/lib/libfdtd2d.so); this package is
its ctypes face. See include/fdtd2d.h for the C interface.
"""

from ._lib import (  # noqa: F401
    EZ, HX, HY, JZ, MEDIUM_DENSITY, MEDIUM_TABLE, PSI_EZX, PSI_EZY, PSI_HXY, PSI_HYX,
    FDTD2DError, LIB_PATH, lib,
)
from .problem import Plan, Pool, Problem, Setup, Snapshots  # noqa: F401
