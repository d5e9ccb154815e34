"""MI355X-native time-stepping path of FDMAP (SBP-SAT elastic waves and
rate-and-state rupture on multi-block curvilinear grids).

The compute path is libfdmap_hip.so (C ABI: include/fdmap.h); the
Fortran driver under fortran/ is the drop-in host.  This package holds the
Python binding of that ABI.
"""
from . import capi  # noqa: F401
from .problem import Problem, lsrk_coefficients  # noqa: F401

__all__ = ["capi", "Problem", "lsrk_coefficients"]
