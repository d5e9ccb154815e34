"""Seeded synthetic problems for the parity tests.

Builds curvilinear multi-block grids, their SBP4 metrics, block-side and
interface descriptors, initial fields and friction parameters directly in
numpy, so that the HIP library and the CPU oracle can be run on identical
inputs of any size.  Metrics follow grid.f90:465-534 (SBP derivative of
the coordinates, J = xq*yr - xr*yq, Ji = 1/J); side coefficients follow
init_edges (domain.f90:1349-1459).
"""
from __future__ import annotations

import numpy as np

from fdmap_github_amd.problem import Problem

H00I = 2.823529411764706
DI = np.array([0.083333333333333, -0.666666666666667, 0.0, 0.666666666666667, -0.083333333333333])
DL = np.array([
    [-1.411764705882353, 1.735294117647059, -0.235294117647059, -0.088235294117647, 0.0, 0.0],
    [-0.5, 0.0, 0.5, 0.0, 0.0, 0.0],
    [0.093023255813953, -0.686046511627907, 0.0, 0.686046511627907, -0.093023255813953, 0.0],
    [0.030612244897959, 0.0, -0.602040816326531, 0.0, 0.653061224489796, -0.081632653061224],
])


def sbp_diff(f, order=4):
    """fd.f90:42-80 along the last axis (both ends inside the range), with
    the operator 'SBP<order>' (the SBP2 and SBP6 tables are the oracle's,
    pinned in test_oracle.py)."""
    if order != 4:
        return _sbp_diff_tables(f, order)
    f = np.asarray(f, dtype=np.float64)
    n = f.shape[-1]
    out = np.empty_like(f)
    for i in range(4):
        s = np.zeros(f.shape[:-1])
        for k in range(6):
            s = s + DL[i, k] * f[..., k]
        out[..., i] = s
    for i in range(4):
        s = np.zeros(f.shape[:-1])
        for k in range(6):
            sidx = 5 - k
            s = s + (-DL[i, sidx]) * f[..., n - 1 - sidx]
        out[..., n - 1 - i] = s
    for i in range(4, n - 4):
        s = np.zeros(f.shape[:-1])
        for k in range(-2, 3):
            s = s + DI[2 + k] * f[..., i + k]
        out[..., i] = s
    return out


def _sbp_diff_tables(f, order):
    from oracle import binding as ob

    DIo, DLo, _, _ = ob.sbp_tables(order)
    nbst, nbnd = DLo.shape
    pI = (len(DIo) - 1) // 2
    f = np.asarray(f, dtype=np.float64)
    n = f.shape[-1]
    out = np.empty_like(f)
    for i in range(nbst):
        s = np.zeros(f.shape[:-1])
        for k in range(nbnd):
            s = s + DLo[i, k] * f[..., k]
        out[..., i] = s
    for i in range(nbst):
        s = np.zeros(f.shape[:-1])
        for k in range(nbnd):
            sidx = nbnd - 1 - k
            s = s + (-DLo[i, sidx]) * f[..., n - 1 - sidx]
        out[..., n - 1 - i] = s
    for i in range(nbst, n - nbst):
        s = np.zeros(f.shape[:-1])
        for k in range(-pI, pI + 1):
            s = s + DIo[pI + k] * f[..., i + k]
        out[..., i] = s
    return out


def h00i(order):
    """SAT penalty weight H00i of 'SBP<order>' (fd_coeff.f90:56-70)"""
    return {2: 2.0, 4: H00I, 6: 3.165067037878233}[order]


def rs_psi(a, V0, S, N, V):
    """State Psi at which the regularised law gives slip velocity V at shear
    stress S (inverse of strength, friction.f90:1066-1070)."""
    return a * np.log(2.0 * V0 / V * np.sinh(S / (a * N)))


def build(nx=48, nys=(24, 24), seed=0, warp=0.02, h=0.025, cs=3.464, cp=6.0, rho=2.67,
          coupling="friction", law="RFL", bcs=None, precision="fp64", background=None,
          perturb=1e-3, source=False, plastic=None, nbx=1, rk_order=4, outside_frac=0.0,
          fault_overstress=25.0, ld_T=0.0, reg=False, rup_field="V", Vinit=1e-12,
          skempton=0.0, mms=False, medium=None, Cdiss=None, mode=2,
          inplane=(-120.0, 0.0, -120.0, -120.0), force=None, pml=None, sbp=4):
    """Blocks stacked in y (len(nys) rows) and nbx columns; interfaces
    between them.  medium: amplitude of a smooth seeded heterogeneous medium
    (rho, cs, cp varying by that relative amount around the given values;
    the edge coefficients then use the medium's values at each node).
    Cdiss: artificial dissipation of every block as the rates use it (a
    number, or one per block).  mode 3: antiplane shear, background
    (vz, sxz, syz); inplane (sxx, sxy, syy, szz) is the in-plane prestress
    that frictional interfaces resolve onto the fault.  force: kinematic
    forcing of every frictional interface, a dict of c0, fd, dfdx, vL, vR, xL,
    xR, tmax (distance along the fault: x for horizontal faults, y for
    vertical ones).  pml: a perfectly matched layer for every block (a dict
    of x, y, n, sigma, kappa, mx, px, my, py), or one dict or None per block.
    sbp: first-derivative operator 'SBP<sbp>' (2, 4, 6) for the metrics, the
    SAT weights and the rates.  Returns (Problem, F0 fields [6,ny,nx])."""
    rng = np.random.default_rng(seed)
    nyb = list(nys)
    ny = sum(nyb)
    nxb = [nx] * nbx
    NX = sum(nxb)
    # block index limits
    xm = np.cumsum([0] + nxb[:-1])
    ym = np.cumsum([0] + nyb[:-1])
    # physical coordinates: block edges at multiples of (n-1)*h, blocks meet
    # at the same physical line (interface points are duplicated)
    X = np.empty((ny, NX))
    Y = np.empty((ny, NX))
    xe = np.cumsum([0.0] + [(n - 1) * h for n in nxb])
    ye = np.cumsum([0.0] + [(n - 1) * h for n in nyb])
    xe -= 0.5 * xe[-1]
    ye -= ye[len(nyb) // 2] if len(nyb) > 1 else 0.5 * ye[-1]
    Lx = xe[-1] - xe[0]
    for bx in range(nbx):
        for by in range(len(nyb)):
            q = np.linspace(xe[bx], xe[bx + 1], nxb[bx])
            r = np.linspace(ye[by], ye[by + 1], nyb[by])
            QQ, RR = np.meshgrid(q, r)
            X[ym[by]:ym[by] + nyb[by], xm[bx]:xm[bx] + nxb[bx]] = QQ
            Y[ym[by]:ym[by] + nyb[by], xm[bx]:xm[bx] + nxb[bx]] = RR
    if warp:
        # smooth warp that is identical on both sides of every interface
        ph = rng.uniform(0, 2 * np.pi, 4)
        X2 = X + warp * Lx * 0.05 * np.sin(2 * np.pi * Y / Lx + ph[0]) * np.cos(np.pi * X / Lx + ph[1])
        Y2 = Y + warp * Lx * 0.05 * np.sin(2 * np.pi * X / Lx + ph[2]) * np.cos(np.pi * Y / Lx + ph[3])
        X, Y = X2, Y2
    prob = Problem(NX, ny, rk_order=rk_order, precision=precision, mode=mode, sbp=sbp)
    xq = np.empty_like(X)
    xr = np.empty_like(X)
    yq = np.empty_like(X)
    yr = np.empty_like(X)
    for bx in range(nbx):
        for by in range(len(nyb)):
            sj = slice(ym[by], ym[by] + nyb[by])
            si = slice(xm[bx], xm[bx] + nxb[bx])
            xq[sj, si] = sbp_diff(X[sj, si], sbp)
            yq[sj, si] = sbp_diff(Y[sj, si], sbp)
            xr[sj, si] = sbp_diff(X[sj, si].T, sbp).T
            yr[sj, si] = sbp_diff(Y[sj, si].T, sbp).T
    J = xq * yr - xr * yq
    prob.xq, prob.xr, prob.yq, prob.yr = xq, xr, yq, yr
    prob.Ji = 1.0 / J
    prob.J = J
    prob.x, prob.y = X, Y

    if medium:
        def smooth(k):
            ph = rng.uniform(0, 2 * np.pi, 2)
            return 1.0 + medium * np.sin((2 + k) * np.pi * X / Lx + ph[0]) * np.cos((1 + k) * np.pi * Y / Lx + ph[1])
        prob.medium = (rho * smooth(0), cs * smooth(1), cp * smooth(2))
    nF = 3 if mode == 3 else 6
    bg = np.zeros(nF) if background is None else np.asarray(background, dtype=np.float64)
    assert bg.shape == (nF,)
    G = rho * cs ** 2
    gamma = 1.0 - 2.0 * (cs / cp) ** 2
    nbl = nbx * len(nyb)
    if bcs is None:
        bcs = ("absorbing", "absorbing", "absorbing", "free")
    for by in range(len(nyb)):
        for bx in range(nbx):
            mx, px = xm[bx], xm[bx] + nxb[bx] - 1
            my, py = ym[by], ym[by] + nyb[by] - 1
            sides = []
            for s in range(4):
                if s == 0:
                    idx = (slice(my, py + 1), mx)
                elif s == 1:
                    idx = (slice(my, py + 1), px)
                elif s == 2:
                    idx = (my, slice(mx, px + 1))
                else:
                    idx = (py, slice(mx, px + 1))
                if s < 2:
                    tx, ty = xr[idx], yr[idx]
                    nrm = np.hypot(tx, ty)
                    n1, n2 = (-ty / nrm, tx / nrm) if s == 0 else (ty / nrm, -tx / nrm)
                else:
                    tx, ty = xq[idx], yq[idx]
                    nrm = np.hypot(tx, ty)
                    n1, n2 = (ty / nrm, -tx / nrm) if s == 2 else (-ty / nrm, tx / nrm)
                n = len(n1)
                Jv = J[idx]
                M = np.empty((5, n))
                if prob.medium is not None:
                    # init_edges with the medium at each node (domain.f90:1379-1456)
                    r_, s_, p_ = (a[idx] for a in prob.medium)
                else:
                    r_, s_, p_ = rho, cs, cp
                M[0] = r_ * s_
                M[1] = r_ * p_
                M[2] = 1.0 - 2.0 * (s_ / p_) ** 2
                M[3] = h00i(sbp) * s_ * nrm / Jv
                M[4] = h00i(sbp) * p_ * nrm / Jv
                F0 = np.tile(bg[:, None], (1, n))
                outer = ((s == 0 and bx == 0) or (s == 1 and bx == nbx - 1) or
                         (s == 2 and by == 0) or (s == 3 and by == len(nyb) - 1))
                bc = bcs[s] if outer else "none"
                sides.append(dict(bc=bc, normal=np.vstack([n1, n2]), M=M, F0=F0,
                                  x=X[idx].copy(), y=Y[idx].copy()))
            blk = prob.add_block(mx, px, my, py, rho, cs, cp, sides, plastic=plastic, mms=mms)
            if mode == 3:
                sxx_, sxy_, syy_, szz_ = inplane
                blk["s0"] = [sxx_, sxy_, 0.0, syy_, 0.0, szz_]
            if Cdiss is not None:
                blk["Cdiss"] = Cdiss[len(prob.blocks) - 1] if np.ndim(Cdiss) else float(Cdiss)
            if pml is not None:
                blk["pml"] = pml[len(prob.blocks) - 1] if isinstance(pml, (list, tuple)) else pml
    # interfaces between vertically adjacent blocks (direction 'y') and
    # horizontally adjacent ones ('x'), domain.f90:1648-1682
    for by in range(len(nyb)):
        for bx in range(nbx - 1):
            bm = bx + by * nbx
            bp = bm + 1
            B = prob.blocks[bm]
            nrm_side = B["sides"][1]["normal"]
            n = B["py"] - B["my"] + 1
            prob.ifaces.append(_iface(rng, "x", bm, bp, B["my"], B["py"], nrm_side, n, X[B["my"]:B["py"] + 1, B["px"]],
                                      coupling, law, bg, fault_overstress, outside_frac, ld_T,
                                      reg, rup_field, Vinit, skempton, mode, inplane))
    for bx in range(nbx):
        for by in range(len(nyb) - 1):
            bm = bx + by * nbx
            bp = bm + nbx
            B = prob.blocks[bm]
            nrm_side = B["sides"][3]["normal"]
            n = B["px"] - B["mx"] + 1
            prob.ifaces.append(_iface(rng, "y", bm, bp, B["mx"], B["px"], nrm_side, n, X[B["py"], B["mx"]:B["px"] + 1],
                                      coupling, law, bg, fault_overstress, outside_frac, ld_T,
                                      reg, rup_field, Vinit, skempton, mode, inplane))
    if force is not None:
        kn = [float(force.get(k, d)) for k, d in
              (("c0", 0.0), ("fd", 0.0), ("dfdx", 1.0), ("vL", 0.0), ("vR", 0.0), ("xL", 0.0),
               ("xR", 0.0), ("tmax", 1e300))]
        for itf in prob.ifaces:
            if itf["coupling"] != "friction":
                continue
            m, p_ = itf["m"], itf["p"]
            if itf["direction"] == "y":
                bm = prob.blocks[itf["blockm"]]
                xs = X[bm["py"], m:p_ + 1]
            else:
                bm = prob.blocks[itf["blockm"]]
                xs = Y[m:p_ + 1, bm["px"]]
            itf["force"] = kn
            itf["kn_x"] = np.ascontiguousarray(xs, dtype=np.float64)
    # initial fields: background + smooth random perturbation
    F = np.empty((nF, ny, NX))
    nvel = 1 if mode == 3 else 2
    for l in range(nF):
        scale = perturb * (1.0 if l < nvel else 10.0)
        k1, k2 = rng.uniform(1, 3, 2)
        ph = rng.uniform(0, 2 * np.pi, 2)
        F[l] = bg[l] + scale * np.sin(k1 * np.pi * X / Lx + ph[0]) * np.cos(k2 * np.pi * Y / Lx + ph[1])
    if mms:
        from oracle import binding as ob  # exact solution at t=0 for MMS runs
        for l in range(6):
            for bi, blk in enumerate(prob.blocks):
                sj = slice(blk["my"], blk["py"] + 1)
                si = slice(blk["mx"], blk["px"] + 1)
                F[l, sj, si] = np.vectorize(lambda x, y: ob.mms_sin(x, y, 0.0, blk["id"], l))(X[sj, si], Y[sj, si])
    if source:
        i0 = NX // 2
        j0 = ny // 3
        a_x, a_y = 0.3, 0.6
        prob.source = dict(stf="gaussian", t_center=0.02, t_duration=0.01, Fx=0.0, Fy=0.0,
                           Mxx=1.0, Mxy=0.5, Myy=1.0,
                           places=[dict(block=0, i0=i0 % nx, j0=j0, dx=_delta(a_x, h), dy=_delta(a_y, h))])
    return prob, F


def _delta(a, h):
    """singular_source weights at i0-2..i0+2 (source.f90:327-356)."""
    beta = (1.0 - 2.0 * a) / 4.0 + 0.25
    gam = (2.0 * a - 1.0) / 16.0
    gap = (2.0 * a + 1.0) / 16.0
    return [(-gam) / h, (beta + 3.0 * gam - gap) / h,
            (1.0 - a - 2.0 * beta - 3.0 * gam + 3.0 * gap) / h,
            (a + beta + gam - 3.0 * gap) / h, gap / h]


def _iface(rng, direction, bm, bp, m, p, nrm, n, xs, coupling, law, bg, overstress, outside_frac,
           ld_T, reg, rup_field, Vinit, skempton, mode=2, inplane=None):
    itf = dict(direction=direction, blockm=bm, blockp=bp, m=int(m), p=int(p), coupling=coupling,
               law=law, nhat=nrm.copy(), reg=reg, rup_field=rup_field, rup_threshold=1e-3,
               skempton=skempton, flock=0.6, ld_T=ld_T)
    if coupling != "friction":
        return itf
    a = 0.010 + 0.002 * rng.uniform(-1, 1, n)
    b = np.full(n, 0.014)
    V0 = np.full(n, 1e-6)
    f0 = np.full(n, 0.6)
    L = np.full(n, 0.4)
    fw = np.full(n, 0.1)
    Vw = np.full(n, 0.1)
    itf["rs"] = np.vstack([a, b, V0, f0, L, fw, Vw])
    itf["sw"] = np.vstack([np.zeros(n), np.full(n, 0.677), np.full(n, 0.525), np.full(n, 0.4)])
    # shear/normal prestress resolved on the fault from the background
    n1, n2 = nrm[0], nrm[1]
    if mode == 3:
        # snz of (sxz, syz); snn of the in-plane prestress (interfaces.f90:859-872)
        sxx, sxy, syy = inplane[0], inplane[1], inplane[2]
        S0 = n1 * bg[1] + n2 * bg[2]
    else:
        sxx, sxy, syy = bg[2], bg[3], bg[4]
        S0 = n1 * n2 * (sxx - syy) + (n2 * n2 - n1 * n1) * sxy
    snn = 2.0 * n1 * n2 * sxy + n1 * n1 * sxx + n2 * n2 * syy
    N0 = -snn
    xmid = 0.5 * (xs[0] + xs[-1])
    R = 0.15 * (xs[-1] - xs[0])
    shape = np.exp(-0.5 * ((xs - xmid) / R) ** 2)
    itf["ldS"] = overstress * shape
    itf["ldN"] = np.zeros(n)
    itf["S0base"] = np.zeros(n)
    itf["N0base"] = np.zeros(n)
    Nn = np.maximum(N0, 1.0)
    if law in ("SL", "FL"):
        itf["Psi0"] = np.full(n, 0.6) if np.isscalar(S0) else np.full(n, 0.6)
        itf["Psi0"] = (np.abs(S0) / Nn) - a * np.log(Vinit / V0)
    else:
        itf["Psi0"] = rs_psi(a, V0, np.abs(S0) + 0.0 * shape, Nn, Vinit)
    if outside_frac > 0:
        out = np.zeros(n, dtype=np.int32)
        k = int(outside_frac * n)
        out[:k] = 1
        out[n - k:] = 1
        itf["outside"] = out
    return itf
