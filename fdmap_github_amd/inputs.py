"""FDMAP .in problem files for the benchmark configurations.

Each function returns the text of an input file in the reference's format
(Fortran namelists, '!---BLOCKn---' / '!---IFACEn---' sections and the
free-text output list), parameterised by size so that the same problem can
be run at full size (benchmark) and reduced size (parity tests).
"""
from __future__ import annotations

import math
import os

import numpy as np

RHO, CS, CP = 2.67, 3.464, 6.0


def _f(v):
    """Fortran double literal."""
    return repr(float(v)).replace("e", "d") if "e" in repr(float(v)) else repr(float(v)) + "d0"


def _fdm(fdmethod):
    """the domain_list entry for FDmethod; None leaves the reference's
    default (SBP6, domain.f90:94)"""
    return "" if fdmethod is None else f" FDmethod = '{fdmethod}',"


def rs_state(a, V0, tau, sn, V):
    """Psi such that the regularised law slips at V under tau, sn
    (inverse of friction.f90:1066-1070)."""
    x = tau / (a * sn)
    # a*log(2*V0/V*sinh(x)) without overflow for large x
    return a * (math.log(2.0 * V0 / V) + x + math.log(0.5 * (1.0 - math.exp(-2.0 * x))))


def c1_pointsource(name="c1", n=401, h=0.025, nsteps=1000, outputs=True, ninfo=100):
    """Single elastic block n x n, free surface on top, absorbing elsewhere,
    isotropic moment-tensor source at the centre with a Gaussian source time
    function of width 0.05 s, LSRK(5,4) at CFL 0.5."""
    L = 0.5 * (n - 1) * h
    out = ""
    if outputs:
        out = f"""B
body
vx vy
0d0 1d10 {max(1, nsteps // 10)}
"""
    return f"""&problem_list
 name = '{name}', nt = {nsteps + 1}, CFL = 0.5d0, RKorder = 4, ninfo = {ninfo} /

&domain_list
 mode = 2, FDmethod = 'SBP4', nblocks = 1, nx = {n}, ny = {n} /

!---BLOCK1---

&grid_list
 nx = {n}, ny = {n},
 LB = {_f(-L)},{_f(-L)}, LT = {_f(-L)},{_f(L)},
 RB = {_f(L)},{_f(-L)}, RT = {_f(L)},{_f(L)} /

&boundaries_list bcL = 'absorbing', bcR = 'absorbing', bcB = 'absorbing', bcT = 'free' /

&material_list rho = {_f(RHO)}, cs = {_f(CS)}, cp = {_f(CP)} /

&source_list
 Mxx = 1d0, Myy = 1d0, x0 = 0d0, y0 = 0d0,
 t_center = 0.2d0, t_duration = 0.05d0, source_time_function = 'gaussian' /

!---begin:output_list---
{out}!---end:output_list---
"""


def c2_planar_rs(name="c2", nx=1601, ny=801, h=0.025, nsteps=4000, outputs=True, ninfo=500,
                 law="RFL", plastic=False, curve=None, tau0=70.0, sn0=120.0, overstress=25.0,
                 R=1.5, Vinit=1e-12, a=0.010, b=0.014, sxy0=None, bcT="absorbing",
                 medium=None, Cdiss=0.0, pml=None, fdmethod="SBP4"):
    """Two blocks nx x ny coupled across a horizontal fault by SAT; slip law
    with strong velocity weakening; Gaussian overstress at the fault centre.
    curve: optional path of a fault-profile curve file (FP64 x, y, n1, n2),
    used for the top edge of block 1 and the bottom edge of block 2.  The
    background shear stress is sxy0 (default tau0); Psi0 gives slip velocity
    Vinit under tau0.  bcT: boundary condition of the top of the upper block
    ('free' for a free surface).  medium: path of a heterogeneous medium file
    (write_medium) that both blocks read.  Cdiss: artificial dissipation of
    both blocks (operator_list, domain.f90:167-180).  pml: a perfectly
    matched layer in both blocks, a dict of pml_list entries (pmlx, pmly, n,
    sigma, kappa, mx, px, my, py; material.f90:152-223).  fdmethod: the
    first-derivative operator (None: the reference's default, SBP6)."""
    Lx = 0.5 * (nx - 1) * h
    Ly = (ny - 1) * h
    psi0 = rs_state(a, 1e-6, tau0, sn0, Vinit)
    c1 = f"bndTfile = '{curve}'," if curve else ""
    c2 = f"bndBfile = '{curve}'," if curve else ""
    mat = f"rho = {_f(RHO)}, cs = {_f(CS)}, cp = {_f(CP)}"
    op = f"&operator_list Cdiss = {_f(Cdiss)} /\n\n" if Cdiss else ""
    if pml:
        ent = ", ".join(f"{k} = {('T' if v else 'F') if isinstance(v, bool) else _f(v)}" for k, v in pml.items())
        op += f"&pml_list {ent} /\n\n"
    if plastic:
        mat = f"response = 'plastic', {mat}, mu = 0.6d0, b = 0d0, beta = 0d0"
    if medium:
        mat = f"{mat}, heterogeneous = T, filename = '{medium}'"
    out = ""
    if outputs:
        st = max(1, nsteps // 100)
        out = f"""I
iface1
D V S N Psi
0d0 1d10 {st}
"""
    return f"""&problem_list
 name = '{name}', nt = {nsteps + 1}, CFL = 0.5d0, RKorder = 4, ninfo = {ninfo} /

&domain_list
 mode = 2,{_fdm(fdmethod)} nblocks_x = 1, nblocks_y = 2, nx = {nx}, ny = {2 * ny} /

!---BLOCK1---

{op}&grid_list
 nx = {nx}, ny = {ny}, {c1}
 LB = {_f(-Lx)},{_f(-Ly)}, LT = {_f(-Lx)},0d0,
 RB = {_f(Lx)},{_f(-Ly)}, RT = {_f(Lx)},0d0 /

&boundaries_list bcL = 'absorbing', bcR = 'absorbing', bcB = 'absorbing', bcT = 'none' /

!---BLOCK2---

{op}&grid_list
 nx = {nx}, ny = {ny}, {c2}
 LB = {_f(-Lx)},0d0, LT = {_f(-Lx)},{_f(Ly)},
 RB = {_f(Lx)},0d0, RT = {_f(Lx)},{_f(Ly)} /

&boundaries_list bcL = 'absorbing', bcR = 'absorbing', bcB = 'none', bcT = '{bcT}' /

!---ALL BLOCKS---

&fields_list sxx0 = {_f(-sn0)}, sxy0 = {_f(tau0 if sxy0 is None else sxy0)}, syy0 = {_f(-sn0)}, szz0 = {_f(-sn0)} /

&material_list {mat} /

!---IFACE1---

&interface_list coupling = 'friction', iblockm = 1, iblockp = 2, direction = 'y' /

&friction_list
 friction_law = '{law}',
 rs%a = {_f(a)}, rs%b = {_f(b)}, rs%V0 = 1d-6, rs%f0 = 0.6d0, rs%L = 0.4d0,
 rs%fw = 0.1d0, rs%Vw = 0.1d0,
 Psi0 = {_f(psi0)},
 ld%shape = 'gaussian', ld%S = {_f(overstress)}, ld%R = {_f(R)}, ld%x0 = 0d0 /

!---begin:output_list---
{out}!---end:output_list---
"""


def antiplane_rs(name="m3", nx=301, ny=301, h=0.1, nsteps=1000, outputs=True, law="RAL",
                 sxx0=-50.0, sxz0=30.0, a=0.010, b=0.016, L=0.05, Psi0=0.6, ld_S=5.0,
                 ld_T=0.0, ld_width=4.0, ninfo=100, medium=None, fdmethod="SBP4"):
    """Antiplane shear (mode 3) rupture after the reference's mode-3 example
    (problems/kali-heterogeneous-input.in): two blocks side by side with a
    vertical rate-and-state fault between them, a free surface on top,
    absorbing elsewhere; the fault normal stress is -sxx0 and the shear
    prestress sxz0, with a Gaussian shear load ld_S at mid-depth.  medium:
    path of a heterogeneous medium file (write_medium) of the whole grid
    that both blocks read."""
    mat = f"rho = {_f(RHO)}, cs = {_f(CS)}, cp = {_f(CP)}"
    if medium:
        mat = f"{mat}, heterogeneous = T, filename = '{medium}'"
    Lx = (nx - 1) * h
    Ly = (ny - 1) * h
    out = ""
    if outputs:
        st = max(1, nsteps // 100)
        out = f"""I
iface1
D V S N Psi
0d0 1d10 {st}
B
body
vz sxz syz
0d0 1d10 {max(1, nsteps // 10)}
"""
    return f"""&problem_list
 name = '{name}', nt = {nsteps + 1}, CFL = 0.5d0, RKorder = 4, ninfo = {ninfo} /

&domain_list
 mode = 3,{_fdm(fdmethod)} nblocks_x = 2, nblocks_y = 1, nx = {2 * nx}, ny = {ny} /

!---BLOCK1---

&grid_list
 nx = {nx}, ny = {ny}, mgx = 1, mgy = 1, iblock_x = 1, iblock_y = 1,
 LT = {_f(-Lx)},0d0, RT = 0d0,0d0,
 LB = {_f(-Lx)},{_f(-Ly)}, RB = 0d0,{_f(-Ly)} /

&boundaries_list bcL = 'absorbing', bcR = 'none', bcB = 'absorbing', bcT = 'free' /

!---BLOCK2---

&grid_list
 nx = {nx}, ny = {ny}, mgx = {nx + 1}, mgy = 1, iblock_x = 2, iblock_y = 1,
 LT = 0d0,0d0, RT = {_f(Lx)},0d0,
 LB = 0d0,{_f(-Ly)}, RB = {_f(Lx)},{_f(-Ly)} /

&boundaries_list bcL = 'none', bcR = 'absorbing', bcB = 'absorbing', bcT = 'free' /

!---ALL BLOCKS---

&material_list {mat} /

&fields_list sxx0 = {_f(sxx0)}, sxz0 = {_f(sxz0)} /

!---IFACE1---

&interface_list coupling = 'friction', iblockm = 1, iblockp = 2, direction = 'x', mg = 1, pg = {ny} /

&friction_list
 angle = 90d0, friction_law = '{law}', Psi0 = {_f(Psi0)},
 rs = {_f(a)},{_f(b)},1d-6,0.6d0,{_f(L)},
 ld = 'gaussian',{_f(ld_S)},0d0,{_f(ld_T)},{_f(ld_width)},{_f(-0.5 * Ly)},0d0 / 

!---begin:output_list---
{out}!---end:output_list---
"""


def kali_het(directory, name="kali_het", n=301, nsteps=2000, outputs=True, ninfo=10,
             fdmethod=None):
    """The reference's mode-3 rate-and-state example with fault properties
    from files (problems/kali-heterogeneous-input.in, files as
    problems/kali_input.m makes them): two 30 km x 30 km blocks side by side
    with a vertical aging-law fault between them, a free surface on top; a,
    b, the state and a fault prestress vary along the fault.  n points per
    block side (301 in the example).  The example leaves FDmethod at the
    reference's default, SBP6 (domain.f90:94), as fdmethod=None does; a
    name ('SBP2', 'SBP4', 'SBP6') sets it.  Writes the three fault files next
    to the .in text it returns."""
    os.makedirs(directory, exist_ok=True)
    y = np.linspace(-30.0, 0.0, n)
    a = 0.01 + 0.005 * np.sin(2 * np.pi * y / 10)
    b = 0.016 + 0.004 * np.exp(-0.5 * ((y + 15) / 3) ** 2)
    rs = [a, b, 1e-6 + 0 * y, 0.6 + 0 * y, 0.05 + 0 * y, 0 * y, 1e10 + 0 * y]
    fric = os.path.join(directory, "kali.friction")
    state = os.path.join(directory, "kali.state")
    pre = os.path.join(directory, "kali.prestress")
    # consecutive FP64 arrays along the fault (friction.f90:373-489)
    np.concatenate(rs).astype("<f8").tofile(fric)
    (0.6 + 0.1 * np.exp(y / 5)).astype("<f8").tofile(state)
    np.concatenate([np.cos(2 * np.pi * y / 10), np.sin(2 * np.pi * y / 4)]).astype("<f8").tofile(pre)
    out = ""
    if outputs:
        out = """I
iface1
V S N Psi D
0d0 1d10 1
"""
    return f"""&problem_list
 name = '{os.path.join(directory, name)}', ninfo = {ninfo},
 nt = {nsteps + 1}, CFL = 0.5d0, refine = 1d0 /

&domain_list
 mode = 3,{_fdm(fdmethod)} nx = {n}, ny = {2 * n},
 nblocks = 2, nblocks_x = 2, nblocks_y = 1, nifaces = 1 /

!---BLOCK1---

&grid_list
 nx = {n}, ny = {n}, mgx = 1, mgy = 1, iblock_x = 1, iblock_y = 1,
 LT = -30d0,  0d0, RT = 0d0,  0d0,
 LB = -30d0,-30d0, RB = 0d0,-30d0 /

&boundaries_list
 bcL = 'absorbing', bcR = 'none', bcB = 'absorbing', bcT = 'free' /

!---BLOCK2---

&grid_list
 nx = {n}, ny = {n}, mgx = {n + 1}, mgy = 1, iblock_x = 2, iblock_y = 1,
 LT = 0d0,  0d0, RT = 30d0,  0d0,
 LB = 0d0,-30d0, RB = 30d0,-30d0 /

&boundaries_list
 bcL = 'none', bcR = 'absorbing', bcB = 'absorbing', bcT = 'free' /

!---ALL BLOCKS---

&material_list
 rho = 2.7d0, cs = 3.4641d0, cp = 6d0 /

&fields_list
 sxx0 = -50d0,
 sxz0 =  30d0 /

!---IFACE1---

&interface_list
 coupling = 'friction', iblockm = 1, iblockp = 2,
 direction = 'x', mg = 1, pg = {n} /

&friction_list
 angle = 90d0, xlockm = -25d0, xlockp = 25d0,
 friction_law = 'RAL',
 friction_file = T, filename = '{fric}',
 stress_file = T, stress_filename = '{pre}',
 state_file = T, state_filename = '{state}',
 Psi0 = 0.6d0,
 rs = 0.01d0,0.016d0,1d-6,0.6d0,0.05d0,
 ld = 'gaussian',5d0,0d0,1d-9,4d0,-15d0,0d0 /

!---begin:output_list---
{out}!---end:output_list---
"""


def gp229_sw3(name="gp229", nx=801, nsteps=1000, CFL=1.0, outputs=True, ninfo=10, fdmethod=None):
    """The reference's antiplane slip-weakening example (gp229/gp229sw3ex1.in):
    two 80 km x 40 km blocks stacked in y with a horizontal fault, prestress
    syy0 = -120, syz0 = 70 MPa, slip-weakening friction with forced rupture
    expansion, and a smooth boxcar lowering the shear stress by 3 MPa for
    x > 10 km.  nx points along the fault (801 in the example), nx/2 + 1
    across each block.  The example leaves FDmethod at the reference's
    default, SBP6 (domain.f90:94), as fdmethod=None does; a name sets it."""
    ny = nx // 2 + 1
    out = ""
    if outputs:
        out = f"""I
iface1
D V S N
0d0 1d10 1
B
body
vz sxz syz
0d0 1d10 10
"""
    return f"""&problem_list
 name = '{name}', ninfo = {ninfo}, nt = {nsteps + 1}, CFL = {_f(CFL)}, refine = 1d0 /

&domain_list
 mode = 3,{_fdm(fdmethod)} nx = {nx}, ny = {2 * ny}, nblocks = 2, nblocks_x = 1, nblocks_y = 2,
 nifaces = 1 /

!---BLOCK1---

&grid_list
 nx = {nx}, ny = {ny}, mgx = 1, mgy = 1, iblock_x = 1, iblock_y = 1,
 LT = -40d0,0d0, RT = 40d0,0d0, LB = -40d0,-40d0, RB = 40d0,-40d0 /

&boundaries_list bcL = 'absorbing', bcR = 'absorbing', bcB = 'absorbing', bcT = 'none' /

!---BLOCK2---

&grid_list
 nx = {nx}, ny = {ny}, mgx = 1, mgy = {ny + 1}, iblock_x = 1, iblock_y = 2,
 LT = -40d0,40d0, RT = 40d0,40d0, LB = -40d0,0d0, RB = 40d0,0d0 /

&boundaries_list bcL = 'absorbing', bcR = 'absorbing', bcB = 'none', bcT = 'absorbing' /

!---ALL BLOCKS---

&fields_list syy0 = -120d0, syz0 = 70d0 /

&material_list rho = 2.67d0, cs = 3.464d0, cp = 6d0 /

&operator_list Cdiss = 0d0 /

!---IFACE1---

&interface_list coupling = 'friction', iblockm = 1, iblockp = 2, direction = 'y', mg = 1, pg = {nx} /

&friction_list
 friction_law = 'SW', sw%fs = 0.677d0, sw%fd = 0.525d0, sw%Dc = 0.4d0,
 force = T, kn%fd = 0.525d0, kn%dfdx = 0.2d0, kn%vL = -1d0, kn%vR = 1d0,
 kn%xL = 0.76d0, kn%xR = -0.76d0, kn%tmax = 4d0,
 ld%shape = 'smooth_boxcar', ld%S = -3d0, ld%x0 = 40d0, ld%R = 30d0 /

!---begin:output_list---
{out}!---end:output_list---
"""


def fractal_profile(nx, h, alpha=1e-2, seed=0, lam_min_h=20):
    """Seeded band-limited self-similar profile y(x) over nx points of
    spacing h: power spectrum ~ k^-3 between wavelengths lam_min_h*h and the
    profile length L, scaled so that rms(y) = alpha*L.  Returns x, y, and the
    downward unit normal (n1, n2) = (y', -1)/sqrt(1+y'^2) of the curve-file
    convention (grid.f90:686-695)."""
    rng = np.random.default_rng(seed)
    L = (nx - 1) * h
    x = np.linspace(-0.5 * L, 0.5 * L, nx)
    kmin = 2 * np.pi / L
    kmax = 2 * np.pi / (lam_min_h * h)
    nk = int(np.floor(kmax / kmin))
    k = kmin * np.arange(1, nk + 1)
    amp = k ** -1.5
    ph = rng.uniform(0, 2 * np.pi, nk)
    y = (amp[None, :] * np.cos(np.outer(x, k) + ph[None, :])).sum(axis=1)
    yp = (-amp[None, :] * k[None, :] * np.sin(np.outer(x, k) + ph[None, :])).sum(axis=1)
    s = alpha * L / np.sqrt(np.mean(y ** 2))
    y *= s
    yp *= s
    nrm = np.sqrt(1.0 + yp ** 2)
    return x, y, yp / nrm, -1.0 / nrm


def write_curve(path, x, y, n1, n2):
    """binary FP64 curve file: x(1:n), y(1:n), n1(1:n), n2(1:n)"""
    with open(path, "wb") as f:
        for a in (x, y, n1, n2):
            np.ascontiguousarray(a, dtype="<f8").tofile(f)


def write_medium(path, rho, cs, cp):
    """A heterogeneous medium file (material.f90:278-321): density, S and P
    wave speed of every grid point, each an FP64 array [ny, nx] of the whole
    grid, x fastest."""
    with open(path, "wb") as f:
        for a in (rho, cs, cp):
            np.ascontiguousarray(a, dtype="<f8").tofile(f)
    return path


def smooth_medium(nx, ny, h, amp=0.1, seed=0, nmodes=12):
    """Seeded smooth medium on the whole grid [ny, nx] of spacing h: rho, cs
    and cp times (1 + amp * g), g a sum of nmodes random plane waves with
    wavelengths between 1/8 and the grid's height, scaled to max |g| = 1,
    plus an independent 1/4-amplitude field for cp so the speed ratio varies
    too.  Returns rho, cs, cp."""
    rng = np.random.default_rng(seed)
    Ly = (ny - 1) * h
    y, x = np.meshgrid(np.arange(ny) * h, np.arange(nx) * h, indexing="ij")

    def field():
        g = np.zeros((ny, nx))
        for _ in range(nmodes):
            lam = Ly * rng.uniform(0.125, 1.0)
            th = rng.uniform(0, 2 * np.pi)
            k = 2 * np.pi / lam
            g += np.cos(k * (np.cos(th) * x + np.sin(th) * y) + rng.uniform(0, 2 * np.pi))
        return g / np.abs(g).max()

    g, g2 = field(), field()
    return (RHO * (1 + 0.5 * amp * g), CS * (1 + amp * g),
            CP * (1 + amp * g + 0.25 * amp * g2))


def write(path, text):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write(text)
    return path
