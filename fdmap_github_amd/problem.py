"""Problem container: numpy arrays -> the fdm_problem C structure.

The production host (the Fortran driver under fdmap_github_amd/fortran)
fills fdm_problem itself; this container lets Python code describe a
problem with numpy arrays and hand it to any implementation of the
fdmap.h interface.  Arrays are kept alive by the container for as long as
the C structure is in use.
"""
from __future__ import annotations

import numpy as np

from . import capi


def lsrk_coefficients(order: int):
    """Low-storage RK tables A, B, C(0..nstage) (time_step.f90:25-66)."""
    if order == 1:
        return [0.0], [1.0], [0.0, 1.0]
    if order == 2:
        return [0.0, -1.0], [1.0, 1.0 / 2.0], [0.0, 1.0, 1.0]
    if order == 3:
        return ([0.0, -5.0 / 9.0, -153.0 / 128.0],
                [1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0],
                [0.0, 1.0 / 3.0, 3.0 / 4.0, 1.0])
    if order == 4:  # Kennedy-Carpenter (5,4), time_step.f90:54-63
        A = [0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
             -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0]
        B = [1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
             1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
             2277821191437.0 / 14882151754819.0]
        Cc = [0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
              2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0, 1.0]
        return A, B, Cc
    raise ValueError(f"invalid Runge-Kutta order {order} (init_RK, time_step.f90:33-35)")


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Problem:
    """One problem in the fdm_problem layout (0-based global indices)."""

    def __init__(self, nx: int, ny: int, rk_order: int = 4, precision: str = "fp64", mode: int = 2,
                 sbp: int = 4):
        self.nx, self.ny = int(nx), int(ny)
        # 2: plane strain (6 fields); 3: antiplane shear (3 fields)
        self.mode = int(mode)
        self.sbp = int(sbp)  # first-derivative operator: 2, 4 or 6
        self.rk_order = rk_order
        self.precision = precision
        shape = (self.ny, self.nx)
        self.xq = np.ones(shape)
        self.xr = np.zeros(shape)
        self.yq = np.zeros(shape)
        self.yr = np.ones(shape)
        self.Ji = np.ones(shape)
        self.x = None
        self.y = None
        self.blocks: list[dict] = []
        self.ifaces: list[dict] = []
        self.source: dict | None = None
        # heterogeneous medium: (rho, cs, cp) arrays [ny, nx], or None
        self.medium: tuple | None = None
        # Jacobian [ny, nx], needed by blocks with artificial dissipation
        self.J = None
        self._keep = []

    # ---- construction helpers ----
    def add_block(self, mx, px, my, py, rho, cs, cp, sides, plastic=None, mms=False, ident=None):
        """sides: 4 dicts (L, R, B, T) with keys bc, normal (2,n), M (5,n),
        F0 (6,n), optional x, y (n)."""
        G = rho * cs ** 2
        M = rho * cp ** 2
        lam = M - 2.0 * G
        blk = dict(mx=mx, px=px, my=my, py=py, rho=rho, G=G, M=M,
                   gamma=1.0 - 2.0 * (cs / cp) ** 2, nu=0.5 * lam / (lam + G),
                   K=lam + 2.0 * G / 3.0, plastic=plastic, mms=mms,
                   id=ident if ident is not None else len(self.blocks) + 1, sides=sides)
        self.blocks.append(blk)
        return blk

    # ---- C structure ----
    def cstruct(self) -> capi.fdm_problem:
        keep = []

        def P(a):
            a = _f64(a)
            if a is not None:
                keep.append(a)
            return capi.dptr(a)

        def IP(a):
            if a is None:
                return capi.iptr(None)
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return capi.iptr(a)

        prob = capi.fdm_problem()
        prob.nx, prob.ny = self.nx, self.ny
        prob.nblocks = len(self.blocks)
        prob.nifaces = len(self.ifaces)
        prob.precision = capi.FDM_FP64 if self.precision == "fp64" else capi.FDM_FP32_BULK
        A, B, Cc = lsrk_coefficients(self.rk_order)
        prob.nstage = len(A)
        for k, v in enumerate(A):
            prob.rkA[k] = v
        for k, v in enumerate(B):
            prob.rkB[k] = v
        for k, v in enumerate(Cc):
            prob.rkC[k] = v
        prob.xq, prob.xr, prob.yq, prob.yr, prob.Ji = (
            P(self.xq), P(self.xr), P(self.yq), P(self.yr), P(self.Ji))
        prob.x = P(self.x) if self.x is not None else capi.dptr(None)
        prob.y = P(self.y) if self.y is not None else capi.dptr(None)
        if self.medium is not None:
            prob.mat_rho, prob.mat_cs, prob.mat_cp = (P(a) for a in self.medium)
        else:
            prob.mat_rho = prob.mat_cs = prob.mat_cp = capi.dptr(None)
        prob.J = P(self.J) if self.J is not None else capi.dptr(None)
        prob.mode = self.mode
        prob.sbp = self.sbp

        barr = (capi.fdm_block * max(1, len(self.blocks)))()
        for b, blk in enumerate(self.blocks):
            cb = barr[b]
            cb.mx, cb.px, cb.my, cb.py = blk["mx"], blk["px"], blk["my"], blk["py"]
            for k in ("rho", "G", "M", "gamma", "nu", "K"):
                setattr(cb, k, float(blk[k]))
            pl = blk.get("plastic")
            cb.plastic = 1 if pl else 0
            cb.mms = 1 if blk.get("mms") else 0
            cb.id = int(blk["id"])
            cb.Cdiss = float(blk.get("Cdiss", 0.0))
            # perfectly matched layer: {"x": bool, "y": bool, "n", "sigma",
            # "kappa", "mx", "px", "my", "py"} (material.f90:152-223)
            pml = blk.get("pml")
            if pml:
                cb.pmlx = 1 if pml.get("x") else 0
                cb.pmly = 1 if pml.get("y") else 0
                big = 1.7976931348623157e308
                for k, (name, dflt) in enumerate((("n", 0.0), ("sigma", 0.0), ("kappa", 0.0), ("mx", -big),
                                                  ("px", big), ("my", -big), ("py", big))):
                    cb.pml[k] = float(pml.get(name, dflt))
            if pl:
                for k in ("mu", "beta", "b", "h", "eta"):
                    setattr(cb, k, float(pl.get(k, 0.0)))
            for k, v in enumerate(pl.get("s0", [0.0] * 6) if pl else blk.get("s0", [0.0] * 6)):
                cb.s0[k] = float(v)
            for s in range(4):
                sd = blk["sides"][s]
                n = (blk["py"] - blk["my"] + 1) if s < 2 else (blk["px"] - blk["mx"] + 1)
                cs_ = cb.side[s]
                cs_.bc = capi.BC[sd["bc"]] if isinstance(sd["bc"], str) else int(sd["bc"])
                cs_.n = n
                cs_.normal = P(sd["normal"])
                cs_.M = P(sd["M"])
                cs_.F0 = P(sd["F0"])
                cs_.x = P(sd.get("x"))
                cs_.y = P(sd.get("y"))
        keep.append(barr)
        prob.blocks = barr

        iarr = (capi.fdm_iface * max(1, len(self.ifaces)))()
        for k, itf in enumerate(self.ifaces):
            ci = iarr[k]
            ci.direction = 0 if itf["direction"] in ("x", 0) else 1
            ci.blockm, ci.blockp = itf["blockm"], itf["blockp"]
            ci.m, ci.p = itf["m"], itf["p"]
            ci.coupling = capi.COUPLING[itf.get("coupling", "friction")]
            ci.law = capi.LAW[itf.get("law", "frictionless")]
            ci.reg = 1 if itf.get("reg") else 0
            ci.rup_field = 1 if itf.get("rup_field", "V") == "D" else 0
            ci.rup_threshold = float(itf.get("rup_threshold", 0.0))
            ci.skempton = float(itf.get("skempton", 0.0))
            ci.flock = float(itf.get("flock", 1e10))
            ci.ld_T = float(itf.get("ld_T", 0.0))
            ci.nhat = P(itf["nhat"])
            ci.rs = P(itf.get("rs"))
            ci.sw = P(itf.get("sw"))
            ci.S0base = P(itf.get("S0base"))
            ci.N0base = P(itf.get("N0base"))
            ci.ldS = P(itf.get("ldS"))
            ci.ldN = P(itf.get("ldN"))
            ci.outside = IP(itf.get("outside"))
            ci.Psi0 = P(itf.get("Psi0"))
            if itf.get("force") is not None:
                ci.force = 1
                for q, v in enumerate(itf["force"]):
                    ci.kn[q] = float(v)
                ci.kn_x = P(itf["kn_x"])
        keep.append(iarr)
        prob.ifaces = iarr

        src = self.source
        if src:
            prob.source.nplace = len(src["places"])
            prob.source.stf = capi.STF[src.get("stf", "gaussian")]
            prob.source.t_center = float(src.get("t_center", 0.0))
            prob.source.t_duration = float(src.get("t_duration", 1.0))
            for k in ("Fx", "Fy", "Mxx", "Mxy", "Myy"):
                setattr(prob.source, k, float(src.get(k, 0.0)))
            for k, pl in enumerate(src["places"]):
                cp_ = prob.source.place[k]
                cp_.block, cp_.i0, cp_.j0 = pl["block"], pl["i0"], pl["j0"]
                for q in range(5):
                    cp_.dx[q] = float(pl["dx"][q])
                    cp_.dy[q] = float(pl["dy"][q])
        self._keep = keep
        return prob
