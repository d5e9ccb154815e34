"""Gradients through the discrete adjoint of the time-stepping path
(include/fdmap_adjoint.h), and a torch.autograd.Function over it.

`Adjoint` wraps the C ABI on a `device.Context`.  `FaultInversion` keeps a
context's initial state, so that every evaluation starts from the same
fields and fault state with new friction parameters, and
`ReceiverTraces.apply(problem, S0, a, b)` returns the receiver velocities
[nsteps, nrec, 2] as a torch tensor whose backward pass is the hand-written
adjoint.  The parameter vectors are concatenated over the interfaces (length
`nodes`)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .device import FdmError

PARAMS = {"S0": capi.FDM_GRAD_S0, "a": capi.FDM_GRAD_A, "b": capi.FDM_GRAD_B}


class Adjoint:
    """interval: checkpoint interval in steps; 0 keeps the whole fault trace
    when it fits the trace budget (else square-root checkpointing), -1
    forces square-root checkpointing."""

    def __init__(self, ctx, rec_i, rec_j, nsteps, interval=0):
        self.lib = capi.load_library()
        self.ctx = ctx
        ri = np.ascontiguousarray(rec_i, dtype=np.int32)
        rj = np.ascontiguousarray(rec_j, dtype=np.int32)
        assert ri.shape == rj.shape
        self.nrec, self.nsteps = len(ri), int(nsteps)
        h = C.c_void_p()
        rc = self.lib.fdm_adjoint_create(ctx.h, self.nrec, capi.iptr(ri), capi.iptr(rj),
                                         self.nsteps, int(interval), C.byref(h))
        if rc != 0:
            msg = self.lib.fdm_adjoint_last_error(None)
            raise FdmError(rc, msg.decode() if msg else "fdm_adjoint_create failed")
        self.h = h
        nodes, itv = C.c_int32(), C.c_int32()
        self._check(self.lib.fdm_adjoint_info(h, C.byref(nodes), C.byref(itv)))
        self.nodes, self.interval = nodes.value, itv.value

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.fdm_adjoint_last_error(self.h)
            raise FdmError(rc, msg.decode() if msg else "adjoint error")

    def set_param(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.shape == (self.nodes,), (v.shape, self.nodes)
        self._check(self.lib.fdm_adjoint_set_param(self.h, PARAMS[name], capi.dptr(v)))

    def forward(self, t, dt):
        rec = np.empty((self.nsteps, self.nrec, 2))
        self._check(self.lib.fdm_adjoint_forward(self.h, float(t), float(dt), capi.dptr(rec)))
        return rec

    def backward(self, rec_bar, fields=False):
        """Returns a dict of gradients (S0, a, b, Psi0 per interface node) and,
        with fields=True, the gradient with respect to the initial fields
        [6, ny, nx]."""
        rb = np.ascontiguousarray(rec_bar, dtype=np.float64)
        assert rb.shape == (self.nsteps, self.nrec, 2)
        g = np.empty((capi.FDM_GRAD_NFIELDS, max(1, self.nodes)))
        gF = np.empty((6, self.ctx.ny, self.ctx.nx)) if fields else None
        self._check(self.lib.fdm_adjoint_backward(self.h, capi.dptr(rb), capi.dptr(g),
                                                  capi.dptr(gF)))
        out = {"S0": g[capi.FDM_GRAD_S0], "a": g[capi.FDM_GRAD_A], "b": g[capi.FDM_GRAD_B],
               "Psi0": g[capi.FDM_GRAD_PSI0]}
        if fields:
            out["F0"] = gF
        return out

    def timing(self):
        v = [C.c_double() for _ in range(3)]
        self._check(self.lib.fdm_adjoint_timing(self.h, *[C.byref(x) for x in v]))
        return {"forward_ms": v[0].value, "recompute_ms": v[1].value, "sweep_ms": v[2].value}

    def close(self):
        if getattr(self, "h", None):
            self.lib.fdm_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FaultInversion:
    """A context with its initial state (fields, rates and every interface
    array, as set before construction) and an `Adjoint` over nsteps steps of
    size dt from time t."""

    def __init__(self, ctx, t, dt, rec_i, rec_j, nsteps, interval=0):
        self.ctx, self.t, self.dt = ctx, float(t), float(dt)
        self.F0, self.DF0 = ctx.get_fields()
        self.iface0 = [ctx.iface_state(k) for k in range(len(ctx.ifaces_n))]
        self.adj = Adjoint(ctx, rec_i, rec_j, nsteps, interval)

    def reset(self):
        self.ctx.set_fields(self.F0, self.DF0)
        for k, st in enumerate(self.iface0):
            for name, v in st.items():
                self.ctx.set_iface(k, name, v)
        self.ctx.init_edges(self.t)

    def run(self, S0=None, a=None, b=None, F0=None):
        """Receiver velocities [nsteps, nrec, 2] from the initial state (or
        the initial fields F0 [6, ny, nx]) with the given parameters."""
        for name, v in (("S0", S0), ("a", a), ("b", b)):
            if v is not None:
                self.adj.set_param(name, v)
        self.reset()
        if F0 is not None:
            self.ctx.set_fields(F0, self.DF0)
            self.ctx.init_edges(self.t)
        return self.adj.forward(self.t, self.dt)


def _autograd():
    import torch

    class ReceiverTraces(torch.autograd.Function):
        """ReceiverTraces.apply(problem, S0, a, b) -> float64 tensor [nsteps,
        nrec, 2]; backward runs fdm_adjoint_backward."""

        @staticmethod
        def forward(ctx, problem, S0, a, b):
            ctx.problem = problem
            rec = problem.run(S0.detach().cpu().double().numpy(), a.detach().cpu().double().numpy(),
                              b.detach().cpu().double().numpy())
            return torch.from_numpy(rec)

        @staticmethod
        def backward(ctx, grad_rec):
            g = ctx.problem.adj.backward(grad_rec.detach().cpu().double().contiguous().numpy())
            return (None, torch.from_numpy(g["S0"].copy()), torch.from_numpy(g["a"].copy()),
                    torch.from_numpy(g["b"].copy()))

    return ReceiverTraces


_fn = None


def receiver_traces(problem, S0, a, b):
    """Differentiable receiver traces of `problem` (a FaultInversion) for
    torch tensors S0, a, b."""
    global _fn
    if _fn is None:
        _fn = _autograd()
    return _fn.apply(problem, S0, a, b)
