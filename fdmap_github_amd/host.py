"""Binding of the Fortran host (libfdmap_host.so).

The Fortran host parses an FDMAP .in problem file exactly as the reference
driver does and fills the fdm_problem description; Python receives a
pointer to that description and can hand it to the MI355X library (or to
any other implementation of include/fdmap.h).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi

HOST_PATH = os.path.join(capi.HERE, "libfdmap_host.so")
EXE_PATH = os.path.join(capi.HERE, "fdmap")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(HOST_PATH):
        raise RuntimeError(f"libfdmap_host.so not found at {HOST_PATH}: build it first")
    capi.load_library()  # the HIP library, resolved by the host's rpath as well
    lib = C.CDLL(HOST_PATH)
    lib.fdmap_host_setup.argtypes = [C.c_char_p, C.c_int32]
    lib.fdmap_host_setup.restype = C.c_int
    lib.fdmap_host_error.argtypes = []
    lib.fdmap_host_error.restype = C.c_char_p
    lib.fdmap_host_problem.argtypes = []
    lib.fdmap_host_problem.restype = C.POINTER(capi.fdm_problem)
    lib.fdmap_host_fields.argtypes = []
    lib.fdmap_host_fields.restype = C.POINTER(C.c_double)
    lib.fdmap_host_info.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32)]
    lib.fdmap_host_info.restype = None
    lib.fdmap_host_coords.argtypes = [C.c_int32]
    lib.fdmap_host_coords.restype = C.POINTER(C.c_double)
    lib.fdmap_host_run.argtypes = [C.c_char_p, C.c_int32]
    lib.fdmap_host_run.restype = C.c_int
    lib.fdmap_host_ckpt_init.argtypes = [C.c_char_p]
    lib.fdmap_host_ckpt_init.restype = C.c_int
    lib.fdmap_host_ckpt_files.argtypes = [C.c_int32, C.c_int32] + [C.c_void_p] * 7
    lib.fdmap_host_ckpt_files.restype = C.c_int
    _lib = lib
    return lib


class HostProblem:
    """A problem read from an .in file by the Fortran host.  The
    description stays valid until the next setup() call."""

    def __init__(self, path: str, precision: str = "fp64"):
        lib = load()
        self.path = path
        prec = capi.FDM_FP64 if precision == "fp64" else capi.FDM_FP32_BULK
        rc = lib.fdmap_host_setup(path.encode(), prec)
        if rc != 0:
            raise RuntimeError(f"fdmap host: {lib.fdmap_host_error().decode()}")
        t, dt, cfl, hmin = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        nt, nx, ny = C.c_int32(), C.c_int32(), C.c_int32()
        lib.fdmap_host_info(C.byref(t), C.byref(dt), C.byref(nt), C.byref(cfl), C.byref(hmin),
                            C.byref(nx), C.byref(ny))
        self.t, self.dt, self.nt = t.value, dt.value, nt.value
        self.CFL, self.hmin = cfl.value, hmin.value
        self.nx, self.ny = nx.value, ny.value
        self.ptr = lib.fdmap_host_problem()
        self.nifaces = self.ptr.contents.nifaces

    def fields(self):
        """initial fields [nF, ny, nx]: 6 in plane strain, 3 (vz, sxz, syz) in
        antiplane shear"""
        p = load().fdmap_host_fields()
        nF = 3 if self.ptr.contents.mode == 3 else 6
        return np.ctypeslib.as_array(p, shape=(6, self.ny, self.nx))[:nF].copy()

    def coords(self, which):
        names = ["x", "y", "xq", "xr", "yq", "yr", "J", "Ji"]
        k = names.index(which) if isinstance(which, str) else int(which)
        p = load().fdmap_host_coords(k)
        return np.ctypeslib.as_array(p, shape=(self.ny, self.nx)).copy()


    def checkpoint_init(self):
        """Parse checkpoint_list and the number file <name>.ckpt as the
        driver does at start; returns the step a run resumes after (0 for a
        fresh start)."""
        lib = load()
        begin = lib.fdmap_host_ckpt_init(self.path.encode())
        if begin < 0:
            raise RuntimeError(f"fdmap host: {lib.fdmap_host_error().decode()}")
        return begin

    def fric_nodes(self):
        P = self.ptr.contents
        return sum(P.ifaces[k].p - P.ifaces[k].m + 1 for k in range(P.nifaces)
                   if P.ifaces[k].coupling == capi.COUPLING["friction"])

    def checkpoint_files(self, op, number, F=None, DF=None, gammap=None, lam=None, IF=None, Wx=None,
                         Wy=None):
        """op 'write' | 'read' | 'delete' on checkpoint number with host
        arrays F, DF [6, ny, nx], gammap, lam [ny, nx], IF [18, nodes of the
        frictional interfaces], Wx, Wy [6, ny, nx] (the PML fields, used when
        a block sets pmlx / pmly); 'read' returns them."""
        lib = load()
        code = {"write": 0, "read": 1, "delete": 2}[op]
        nn = max(1, self.fric_nodes())
        shp = {"F": (6, self.ny, self.nx), "DF": (6, self.ny, self.nx),
               "gammap": (self.ny, self.nx), "lam": (self.ny, self.nx),
               "IF": (len(capi.IFACE_FIELDS), nn), "Wx": (6, self.ny, self.nx), "Wy": (6, self.ny, self.nx)}
        arrs = {}
        for name, v in (("F", F), ("DF", DF), ("gammap", gammap), ("lam", lam), ("IF", IF), ("Wx", Wx),
                        ("Wy", Wy)):
            if op == "read" or v is None:
                arrs[name] = np.zeros(shp[name])
            else:
                arrs[name] = np.ascontiguousarray(v, dtype=np.float64).reshape(shp[name])
        ptrs = [arrs[k].ctypes.data for k in ("F", "DF", "gammap", "lam", "IF", "Wx", "Wy")]
        if lib.fdmap_host_ckpt_files(code, int(number), *ptrs) != 0:
            raise RuntimeError(f"fdmap host: {lib.fdmap_host_error().decode()}")
        return arrs


def run(path: str, precision: str = "fp64") -> int:
    """Run the whole Fortran driver on an .in file in this process."""
    lib = load()
    return lib.fdmap_host_run(path.encode(), capi.FDM_FP64 if precision == "fp64" else capi.FDM_FP32_BULK)
