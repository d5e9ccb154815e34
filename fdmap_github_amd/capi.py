"""ctypes mirror of include/fdmap.h and the loader of libfdmap_hip.so.

The structures here are byte-for-byte the C structures of the C ABI; the
Fortran host (fdmap_github_amd/fortran) declares the same layouts with
iso_c_binding.  Loading fails loudly when the HIP library has not been
built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfdmap_hip.so")

FDM_NF = 6
FDM_MAX_STAGES = 8
FDM_MAX_SOURCE_PLACES = 16

FDM_OK = 0
FDM_EINVAL = -1
FDM_EDEVICE = -2
FDM_ENOMEM = -3
FDM_ENEWTON = -4

FDM_FP64 = 0
FDM_FP32_BULK = 1

# boundary conditions (boundaries.f90:245-357)
BC = {
    "none": 0,
    "absorbing": 1,
    "free": 2,
    "rigid": 3,
    "absorbing-velocity": 4,
    "free-velocity": 5,
    "absorbing-0": 6,
    "free-0": 7,
    "rigid-0": 8,
    "mms-sin-w": 9,
    "inplane-fault-mms-w": 9,
}
SIDE_L, SIDE_R, SIDE_B, SIDE_T = 0, 1, 2, 3
COUPLING = {"locked": 0, "friction": 1}
LAW = {
    "frictionless": 0,
    "locked": 1,
    "SW": 2,
    "SL": 3,
    "FL": 4,
    "RSL": 5,
    "RFL": 6,
    "RSF": 7,
    "RAL": 8,
}
STF = {"ramp": 0, "sinusoid": 1, "gaussian": 2, "Ricker": 3}

# interface fields in the reference checkpoint order (friction.f90:528-546)
IFACE_FIELDS = [
    "D", "Psi", "DPsi", "trup", "V", "O", "S", "N", "S0", "N0", "Ds", "Dn",
    "DDs", "DDn", "W", "DW", "Nreg", "DNreg",
]
IF = {name: k for k, name in enumerate(IFACE_FIELDS)}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class fdm_side(C.Structure):
    _fields_ = [
        ("bc", C.c_int32),
        ("n", C.c_int32),
        ("normal", _dp),
        ("M", _dp),
        ("F0", _dp),
        ("x", _dp),
        ("y", _dp),
    ]


class fdm_block(C.Structure):
    _fields_ = [
        ("mx", C.c_int32), ("px", C.c_int32), ("my", C.c_int32), ("py", C.c_int32),
        ("plastic", C.c_int32), ("mms", C.c_int32), ("id", C.c_int32), ("pad_", C.c_int32),
        ("rho", C.c_double), ("G", C.c_double), ("M", C.c_double),
        ("gamma", C.c_double), ("nu", C.c_double),
        ("mu", C.c_double), ("beta", C.c_double), ("b", C.c_double),
        ("h", C.c_double), ("eta", C.c_double), ("K", C.c_double),
        ("s0", C.c_double * 6),
        ("side", fdm_side * 4),
        ("Cdiss", C.c_double),
        ("pmlx", C.c_int32),
        ("pmly", C.c_int32),
        ("pml", C.c_double * 7),
    ]


class fdm_iface(C.Structure):
    _fields_ = [
        ("direction", C.c_int32), ("blockm", C.c_int32), ("blockp", C.c_int32),
        ("m", C.c_int32), ("p", C.c_int32), ("coupling", C.c_int32), ("law", C.c_int32),
        ("reg", C.c_int32), ("rup_field", C.c_int32), ("pad_", C.c_int32),
        ("rup_threshold", C.c_double), ("skempton", C.c_double), ("flock", C.c_double),
        ("ld_T", C.c_double),
        ("nhat", _dp), ("rs", _dp), ("sw", _dp), ("S0base", _dp), ("N0base", _dp),
        ("ldS", _dp), ("ldN", _dp), ("outside", _ip), ("Psi0", _dp),
        ("force", C.c_int32), ("pad2_", C.c_int32), ("kn", C.c_double * 8), ("kn_x", _dp),
    ]


class fdm_source_place(C.Structure):
    _fields_ = [
        ("block", C.c_int32), ("i0", C.c_int32), ("j0", C.c_int32), ("pad_", C.c_int32),
        ("dx", C.c_double * 5), ("dy", C.c_double * 5),
    ]


class fdm_source(C.Structure):
    _fields_ = [
        ("nplace", C.c_int32), ("stf", C.c_int32),
        ("t_center", C.c_double), ("t_duration", C.c_double),
        ("Fx", C.c_double), ("Fy", C.c_double),
        ("Mxx", C.c_double), ("Mxy", C.c_double), ("Myy", C.c_double),
        ("place", fdm_source_place * FDM_MAX_SOURCE_PLACES),
    ]


class fdm_problem(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nblocks", C.c_int32), ("nifaces", C.c_int32),
        ("precision", C.c_int32), ("nstage", C.c_int32),
        ("rkA", C.c_double * FDM_MAX_STAGES), ("rkB", C.c_double * FDM_MAX_STAGES),
        ("rkC", C.c_double * (FDM_MAX_STAGES + 1)),
        ("xq", _dp), ("xr", _dp), ("yq", _dp), ("yr", _dp), ("Ji", _dp),
        ("x", _dp), ("y", _dp),
        ("blocks", C.POINTER(fdm_block)),
        ("ifaces", C.POINTER(fdm_iface)),
        ("source", fdm_source),
        ("mat_rho", _dp), ("mat_cs", _dp), ("mat_cp", _dp),
        ("J", _dp),
        ("mode", C.c_int32), ("sbp", C.c_int32),
    ]


class fdm_profile_t(C.Structure):
    _fields_ = [
        ("bulk_ms", C.c_double), ("bulk_launches", C.c_int64),
        ("edge_ms", C.c_double), ("edge_launches", C.c_int64),
        ("bulk_points", C.c_int64),
        ("band_ms", C.c_double), ("band_launches", C.c_int64),
    ]


# entry points declared in include/fdmap.h: (name, restype, argtypes)
_P = C.c_void_p
ENTRY_POINTS = [
    ("fdm_create", C.c_int, [C.POINTER(fdm_problem), C.POINTER(_P)]),
    ("fdm_destroy", None, [_P]),
    ("fdm_last_error", C.c_char_p, [_P]),
    ("fdm_set_fields", C.c_int, [_P, _dp, _dp]),
    ("fdm_get_fields", C.c_int, [_P, _dp, _dp]),
    ("fdm_init_edges", C.c_int, [_P, C.c_double]),
    ("fdm_step", C.c_int, [_P, C.c_double, C.c_double, C.c_int64, C.c_int32]),
    ("fdm_stage1", C.c_int, [_P, C.c_double, C.c_double, C.c_int64]),
    ("fdm_get_iface", C.c_int, [_P, C.c_int32, C.c_int32, _dp]),
    ("fdm_set_iface", C.c_int, [_P, C.c_int32, C.c_int32, _dp]),
    ("fdm_get_iface_stage1", C.c_int, [_P, C.c_int32, C.c_int32, _dp]),
    ("fdm_get_plastic", C.c_int, [_P, _dp, _dp]),
    ("fdm_set_plastic", C.c_int, [_P, _dp, _dp]),
    ("fdm_get_pml", C.c_int, [_P, _dp, _dp]),
    ("fdm_set_pml", C.c_int, [_P, _dp, _dp]),
    ("fdm_max_slip_velocity", C.c_int, [_P, _dp]),
    ("fdm_sync", C.c_int, [_P]),
    ("fdm_profile_enable", C.c_int, [_P, C.c_int32]),
    ("fdm_profile_read", C.c_int, [_P, C.POINTER(fdm_profile_t)]),
]

# entry points of include/fdmap.h that only the MI355X library implements:
# the allocation-free variant of fdm_max_slip_velocity
NOALLOC_ENTRY_POINTS = [
    ("fdm_max_slip_velocity_noalloc", C.c_int, [_P, _dp]),
]

# output streaming of include/fdmap.h (the MI355X library only)
FDM_SNAPSHOT_SLOTS = 4
STREAM_ENTRY_POINTS = [
    ("fdm_snapshot", C.c_int, [_P, C.c_uint32, C.POINTER(C.c_int32)]),
    ("fdm_snapshot_read", C.c_int, [_P, C.c_int32, _dp]),
]


# entry points declared in include/fdmap_ensemble.h (the MI355X library only)
FDM_TUNE_TILE = 0
FDM_TUNE_SWIZZLE = 1
FDM_TUNE_NT_STORE = 2
FDM_TUNE_METRIC_F32 = 3
TILE_SHAPES = ["64x4", "64x8", "128x4"]
ENSEMBLE_ENTRY_POINTS = [
    ("fdm_tune", C.c_int, [_P, C.c_int32, C.c_int32]),
    ("fdm_ensemble_create", C.c_int, [C.POINTER(_P), C.c_int32, C.POINTER(_P)]),
    ("fdm_ensemble_step", C.c_int, [_P, _dp, _dp, C.c_int64, C.c_int32]),
    ("fdm_ensemble_sync", C.c_int, [_P]),
    ("fdm_ensemble_destroy", None, [_P]),
    ("fdm_ensemble_last_error", C.c_char_p, [_P]),
]

# include/fdmap_adjoint.h
FDM_GRAD_S0, FDM_GRAD_A, FDM_GRAD_B, FDM_GRAD_PSI0, FDM_GRAD_NFIELDS = range(5)
ADJOINT_ENTRY_POINTS = [
    ("fdm_adjoint_create", C.c_int,
     [_P, C.c_int32, _ip, _ip, C.c_int32, C.c_int32, C.POINTER(_P)]),
    ("fdm_adjoint_info", C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("fdm_adjoint_set_param", C.c_int, [_P, C.c_int32, _dp]),
    ("fdm_adjoint_forward", C.c_int, [_P, C.c_double, C.c_double, _dp]),
    ("fdm_adjoint_backward", C.c_int, [_P, _dp, _dp, _dp]),
    ("fdm_adjoint_timing", C.c_int, [_P, _dp, _dp, _dp]),
    ("fdm_adjoint_last_error", C.c_char_p, [_P]),
    ("fdm_adjoint_destroy", None, [_P]),
]


def bind(lib, prefix_from="fdm_", prefix_to="fdm_"):
    """Attach restype/argtypes to every entry point of a library that
    implements the fdmap.h interface (optionally under another prefix), and
    to the ensemble and adjoint entry points where the library has them."""
    for name, res, args in ENTRY_POINTS:
        fn = getattr(lib, prefix_to + name[len(prefix_from):])
        fn.restype = res
        fn.argtypes = args
    for name, res, args in (NOALLOC_ENTRY_POINTS + STREAM_ENTRY_POINTS + ENSEMBLE_ENTRY_POINTS
                            + ADJOINT_ENTRY_POINTS):
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    return lib


_lib = None


def load_library(path: str | None = None):
    """Load libfdmap_hip.so; raise if it is missing (no CPU fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libfdmap_hip.so not found at {p}: build it first "
            "(python -c 'import __graft_entry__; __graft_entry__.build()')"
        )
    lib = bind(C.CDLL(p))
    if path is None:
        _lib = lib
    return lib


def dptr(a):
    """double* of a C-contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return _dp()
    assert a.dtype.name == "float64" and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return _ip()
    assert a.dtype.name == "int32" and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)
