"""Python handle over an implementation of the fdmap.h interface.

`Context` drives libfdmap_hip.so (the MI355X library).  The same class
can wrap any library exporting the fdmap.h entry points under another
prefix, which is how the tests drive the CPU oracle.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi


class FdmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Context:
    def __init__(self, problem, lib=None, prefix="fdm_"):
        """problem: a Problem, or a pointer to an fdm_problem (e.g. the one
        built by the Fortran host)."""
        self.lib = lib if lib is not None else capi.load_library()
        self.prefix = prefix
        self._problem = problem
        if hasattr(problem, "cstruct"):
            self._cprob = problem.cstruct()
            pp = C.pointer(self._cprob)
            self.nx, self.ny = problem.nx, problem.ny
            self.nF = 3 if getattr(problem, "mode", 2) == 3 else 6
            self.ifaces_n = [itf["p"] - itf["m"] + 1 for itf in problem.ifaces]
        else:
            pp = problem
            self.nx, self.ny = pp.contents.nx, pp.contents.ny
            self.nF = 3 if pp.contents.mode == 3 else 6
            self.ifaces_n = [pp.contents.ifaces[k].p - pp.contents.ifaces[k].m + 1
                             for k in range(pp.contents.nifaces)]
        h = C.c_void_p()
        rc = self._fn("create")(pp, C.byref(h))
        if rc != 0:
            msg = self._fn("last_error")(None)
            raise FdmError(rc, msg.decode() if msg else "create failed")
        self.h = h

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc):
        if rc != 0:
            msg = self._fn("last_error")(self.h)
            raise FdmError(rc, msg.decode() if msg else "error")

    def close(self):
        if getattr(self, "h", None):
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields ----
    def set_fields(self, F, DF=None):
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(self.nF, self.ny, self.nx)
        DFa = None if DF is None else np.ascontiguousarray(DF, dtype=np.float64).reshape(
            self.nF, self.ny, self.nx)
        self._check(self._fn("set_fields")(self.h, capi.dptr(F), capi.dptr(DFa)))

    def get_fields(self):
        F = np.empty((self.nF, self.ny, self.nx))
        DF = np.empty((self.nF, self.ny, self.nx))
        self._check(self._fn("get_fields")(self.h, capi.dptr(F), capi.dptr(DF)))
        return F, DF

    def snapshot(self, fields=None):
        """queue a copy of the current fields (a list of field indices, all by
        default) to host memory; returns the slot to read"""
        fields = range(self.nF) if fields is None else fields
        mask = 0
        for l in fields:
            mask |= 1 << int(l)
        slot = C.c_int32()
        self._check(self._fn("snapshot")(self.h, mask, C.byref(slot)))
        return slot.value, sorted(set(int(l) for l in fields))

    def snapshot_read(self, snap):
        """the fields of a snapshot, [nselected, ny, nx]"""
        slot, fields = snap
        out = np.empty((len(fields), self.ny, self.nx))
        self._check(self._fn("snapshot_read")(self.h, int(slot), capi.dptr(out)))
        return out

    def init_edges(self, t):
        self._check(self._fn("init_edges")(self.h, float(t)))

    def step(self, t, dt, n_first, nsteps):
        self._check(self._fn("step")(self.h, float(t), float(dt), int(n_first), int(nsteps)))

    def stage1(self, t, dt, n):
        self._check(self._fn("stage1")(self.h, float(t), float(dt), int(n)))

    def get_iface(self, k, field, stage1=False):
        f = capi.IF[field] if isinstance(field, str) else int(field)
        out = np.empty(self.ifaces_n[k])
        fn = self._fn("get_iface_stage1" if stage1 else "get_iface")
        self._check(fn(self.h, k, f, capi.dptr(out)))
        return out

    def set_iface(self, k, field, values):
        f = capi.IF[field] if isinstance(field, str) else int(field)
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self._fn("set_iface")(self.h, k, f, capi.dptr(v)))

    def iface_state(self, k, stage1=False):
        return {name: self.get_iface(k, name, stage1) for name in capi.IFACE_FIELDS}

    def get_plastic(self):
        g = np.empty((self.ny, self.nx))
        lam = np.empty((self.ny, self.nx))
        self._check(self._fn("get_plastic")(self.h, capi.dptr(g), capi.dptr(lam)))
        return g, lam

    def set_plastic(self, gammap, lam):
        g = np.ascontiguousarray(gammap, dtype=np.float64)
        l = np.ascontiguousarray(lam, dtype=np.float64)
        assert g.shape == l.shape == (self.ny, self.nx)
        self._check(self._fn("set_plastic")(self.h, capi.dptr(g), capi.dptr(l)))

    def get_pml(self, x=True, y=True):
        """PML auxiliary fields (Wx, Wy), each [6, ny, nx] or None"""
        Wx = np.empty((6, self.ny, self.nx)) if x else None
        Wy = np.empty((6, self.ny, self.nx)) if y else None
        self._check(self._fn("get_pml")(self.h, capi.dptr(Wx), capi.dptr(Wy)))
        return Wx, Wy

    def set_pml(self, Wx=None, Wy=None):
        a = None if Wx is None else np.ascontiguousarray(Wx, dtype=np.float64).reshape(6, self.ny, self.nx)
        b = None if Wy is None else np.ascontiguousarray(Wy, dtype=np.float64).reshape(6, self.ny, self.nx)
        self._check(self._fn("set_pml")(self.h, capi.dptr(a), capi.dptr(b)))

    # ---- device-only helpers ----
    def sync(self):
        self._check(self._fn("sync")(self.h))

    def max_slip_velocity(self, noalloc=False):
        """max of V over the frictional interfaces; noalloc: through the
        entry point that allocates nothing (MI355X library only)"""
        v = C.c_double()
        name = "max_slip_velocity_noalloc" if noalloc else "max_slip_velocity"
        self._check(self._fn(name)(self.h, C.byref(v)))
        return v.value

    def profile_enable(self, on=True):
        self._check(self._fn("profile_enable")(self.h, 1 if on else 0))

    def profile_read(self):
        p = capi.fdm_profile_t()
        self._check(self._fn("profile_read")(self.h, C.byref(p)))
        return dict(bulk_ms=p.bulk_ms, bulk_launches=p.bulk_launches, edge_ms=p.edge_ms,
                    edge_launches=p.edge_launches, bulk_points=p.bulk_points,
                    band_ms=p.band_ms, band_launches=p.band_launches)
