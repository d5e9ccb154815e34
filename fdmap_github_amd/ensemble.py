"""Batched rupture ensembles: scenario generation, the ensemble executor
(include/fdmap_ensemble.h) and the tile-shape autotuner.

An ensemble groups independent problems, each in its own context, so that
one hipGraph launch advances every member by one RK step.  The autotuner
times the interior tile shapes on a scratch context per (size class,
precision) and caches the choice in a JSON file.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os

import numpy as np

from . import capi, inputs
from .device import Context, FdmError


# ---------------------------------------------------------------------------
# scenarios
# ---------------------------------------------------------------------------

def correlated_lognormal(n, h, mean, sigma_ln, corr_len, rng):
    """Lognormal field over n points of spacing h with the given mean, log
    standard deviation and Gaussian correlation length: exp of a unit-variance
    Gaussian random field (white noise filtered by exp(-k^2 l^2 / 4) in
    wavenumber, which gives the covariance exp(-r^2 / (2 l^2)))."""
    m = 1 << int(math.ceil(math.log2(2 * n)))
    w = rng.standard_normal(m)
    k = 2 * np.pi * np.fft.rfftfreq(m, d=h)
    g = np.fft.irfft(np.fft.rfft(w) * np.exp(-(k * corr_len) ** 2 / 4.0), m)[:n]
    g = (g - g.mean()) / g.std()
    return mean * np.exp(sigma_ln * g - 0.5 * sigma_ln ** 2)


def c4_scenarios(count=64, seed=0, nx_range=(401, 2001), ny_range=(201, 1001), h=0.025):
    """The ensemble configuration: per-scenario block sizes drawn uniformly
    from nx_range x ny_range, a lognormal shear prestress on the fault (mean
    70 MPa, log standard deviation 0.1, correlation length 1 km) and a
    uniform-random rate-and-state a in [0.008, 0.012]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        a = float(rng.uniform(0.008, 0.012))
        S = correlated_lognormal(nx, h, 70.0, 0.1, 1.0, rng)
        out.append(dict(index=k, nx=nx, ny=ny, a=a, S0=S, h=h))
    return out


def write_scenario(directory, sc, nsteps, sn0=120.0, Vinit=1e-12, b=0.014, V0=1e-6,
                   bcT="absorbing"):
    """Write scenario sc as an FDMAP .in file plus its fault stress and state
    files (friction.f90:422-443, 469-489 formats); returns the .in path.  The
    medium carries the normal prestress; the fault's shear prestress is the
    scenario's field, and Psi0 puts every node in steady sliding at Vinit."""
    os.makedirs(directory, exist_ok=True)
    k, nx, ny, a = sc["index"], sc["nx"], sc["ny"], sc["a"]
    stress = os.path.join(directory, f"s{k}_stress.bin")
    state = os.path.join(directory, f"s{k}_state.bin")
    S = np.asarray(sc["S0"], dtype="<f8")
    with open(stress, "wb") as f:
        S.tofile(f)
        np.zeros(nx, dtype="<f8").tofile(f)
    psi = np.array([inputs.rs_state(a, V0, float(s), sn0, Vinit) for s in S], dtype="<f8")
    with open(state, "wb") as f:
        psi.tofile(f)
    text = inputs.c2_planar_rs(name=os.path.join(directory, f"s{k}"), nx=nx, ny=ny, h=sc["h"],
                               nsteps=nsteps, outputs=False, tau0=70.0, sxy0=0.0, sn0=sn0, a=a,
                               b=b, Vinit=Vinit, bcT=bcT)
    text = text.replace(
        "&friction_list\n",
        f"&friction_list\n stress_file = T, stress_filename = '{stress}',\n"
        f" state_file = T, state_filename = '{state}',\n")
    return inputs.write(os.path.join(directory, f"s{k}.in"), text)


# ---------------------------------------------------------------------------
# executor
# ---------------------------------------------------------------------------

class Ensemble:
    """fdm_ensemble over existing contexts (which stay owned by the caller)."""

    def __init__(self, contexts):
        self.lib = capi.load_library()
        self.members = list(contexts)
        arr = (C.c_void_p * len(self.members))(*[c.h.value for c in self.members])
        h = C.c_void_p()
        rc = self.lib.fdm_ensemble_create(arr, len(self.members), C.byref(h))
        if rc != 0:
            msg = self.lib.fdm_ensemble_last_error(None)
            raise FdmError(rc, msg.decode() if msg else "fdm_ensemble_create failed")
        self.h = h

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.fdm_ensemble_last_error(self.h)
            raise FdmError(rc, msg.decode() if msg else "ensemble error")

    def step(self, t, dt, n_first, nsteps):
        t = np.ascontiguousarray(t, dtype=np.float64)
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        assert len(t) == len(dt) == len(self.members)
        self._check(self.lib.fdm_ensemble_step(self.h, capi.dptr(t), capi.dptr(dt), int(n_first),
                                               int(nsteps)))

    def sync(self):
        self._check(self.lib.fdm_ensemble_sync(self.h))

    def close(self):
        if self.h:
            self.lib.fdm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# autotuner
# ---------------------------------------------------------------------------

def size_class(nx, ny):
    """Tuning class of a problem: block width in steps of 400 points, height
    in steps of 200."""
    return (int(nx) // 400, int(ny) // 200)


def tune(make_context, key, cache, steps=4, warmup=2, t=0.0, dt=1e-3):
    """Pick the fastest interior tile shape for one class.  make_context()
    returns a fresh scratch context ready to step (fields set, edges
    initialised); it is stepped with each shape and destroyed.  cache maps
    key -> {"tile": index, "ms": {...}} and is updated in place."""
    if key in cache:
        return cache[key]["tile"]
    lib = capi.load_library()
    ctx = make_context()
    times = {}
    try:
        for q, name in enumerate(capi.TILE_SHAPES):
            ctx._check(lib.fdm_tune(ctx.h, capi.FDM_TUNE_TILE, q))
            ctx.step(t, dt, 1, warmup)
            ctx.sync()
            ctx.profile_enable(True)
            ctx.step(t, dt, warmup + 1, steps)
            p = ctx.profile_read()
            ctx.profile_enable(False)
            times[name] = p["bulk_ms"] / max(1, p["bulk_launches"])
    finally:
        ctx.close()
    best = min(range(len(capi.TILE_SHAPES)), key=lambda q: times[capi.TILE_SHAPES[q]])
    cache[key] = {"tile": best, "ms": times}
    return best


def load_cache(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_cache(path, cache):
    if path:
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)


def apply_tile(ctx, tile):
    lib = capi.load_library()
    ctx._check(lib.fdm_tune(ctx.h, capi.FDM_TUNE_TILE, int(tile)))
