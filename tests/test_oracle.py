"""Pins of the CPU oracle against mathematics that fixes each part of the
path, independently of the oracle's own code.

The reference holds no fixtures or known-answer tests for this path, and
its Fortran cannot be built here without stand-ins for the MPI module, so
every check below comes from a closed form, an identity or a published
formula, chosen so that a dropped term, a wrong sign or index, or a
transposed table fails at least one of them:

- SBP4 tables (fd_coeff.f90:97-126): the published diagonal-norm operator,
  consistency, polynomial exactness, the SBP identity H D + (H D)^T = B,
  and the 1-D application (fd.f90:42-80) against the dense matrix;
- LSRK coefficients (time_step.f90:25-66): the published Carpenter-Kennedy
  and Williamson rationals, the stability polynomial and the stage times;
- friction (friction.f90:593-1264): residual and closed-form inverse of
  every law, the tensile, locked, frictionless and slip-weakening branches,
  and the state-evolution laws' steady states and slopes;
- plasticity (plastic.f90:62-192): return onto the Drucker-Prager surface;
- the manufactured solution (mms.f90:52-234): its source terms are
  u_t - L(u) of its own fields;
- the whole stage (rates.f90:384-936): stage-1 rates against a dense
  split-form operator built here from the SBP matrices, energy decay under
  absorbing boundaries, and the convergence rate of the MMS problem.
"""
import math
from fractions import Fraction as Fr

import numpy as np
import pytest

import problems
from oracle import binding as ob
from oracle.binding import OracleContext

# ---------------------------------------------------------------------------
# SBP4 tables
# ---------------------------------------------------------------------------

# diagonal-norm fourth-order SBP first derivative (Strand 1994; Mattsson and
# Nordstrom 2004), closure rows 0..3 and norm weights
DL_EXACT = [[Fr(-24, 17), Fr(59, 34), Fr(-4, 17), Fr(-3, 34), 0, 0],
            [Fr(-1, 2), 0, Fr(1, 2), 0, 0, 0],
            [Fr(4, 43), Fr(-59, 86), 0, Fr(59, 86), Fr(-4, 43), 0],
            [Fr(3, 98), 0, Fr(-59, 98), 0, Fr(32, 49), Fr(-4, 49)]]
HL_EXACT = [Fr(17, 48), Fr(59, 48), Fr(43, 48), Fr(49, 48)]
DI_EXACT = [Fr(1, 12), Fr(-2, 3), 0, Fr(2, 3), Fr(-1, 12)]


def dense_D(n):
    """Dense SBP4 derivative matrix (unit spacing) from the exact tables;
    the right closure is the antisymmetric mirror of the left one."""
    D = np.zeros((n, n))
    dl = np.array([[float(v) for v in row] for row in DL_EXACT])
    for i in range(4):
        D[i, :6] = dl[i]
        D[n - 1 - i, n - 6:] = -dl[i][::-1]
    for i in range(4, n - 4):
        D[i, i - 2:i + 3] = [float(v) for v in DI_EXACT]
    return D


def test_sbp4_tables_match_published_operator():
    DI, DL, HL, H00i = ob.sbp4_tables()
    assert np.allclose(DI, [float(v) for v in DI_EXACT], rtol=0, atol=5e-15)
    assert np.allclose(DL, [[float(v) for v in r] for r in DL_EXACT], rtol=0, atol=5e-15)
    assert np.allclose(HL, [float(v) for v in HL_EXACT], rtol=0, atol=5e-15)
    assert abs(H00i - 48.0 / 17.0) < 5e-15


def test_sbp4_consistency_and_exactness():
    DI, DL, HL, _ = ob.sbp4_tables()
    k = np.arange(-2, 3)
    for p in range(5):  # interior: exact through degree 4
        assert abs(np.dot(DI, k.astype(float) ** p) - (1.0 if p == 1 else 0.0)) < 1e-14
    for r in range(4):  # closures: exact through degree 2 at row r
        kk = np.arange(6) - r
        for p in range(3):
            assert abs(np.dot(DL[r], kk.astype(float) ** p) - (1.0 if p == 1 else 0.0)) < 1e-14


def test_sbp_identity_with_tables():
    DI, DL, HL, _ = ob.sbp4_tables()
    n = 17
    D = np.zeros((n, n))
    for i in range(4):
        D[i, :6] = DL[i]
        D[n - 1 - i, n - 6:] = -DL[i][::-1]
    for i in range(4, n - 4):
        D[i, i - 2:i + 3] = DI
    h = np.ones(n)
    h[:4] = HL
    h[n - 4:] = HL[::-1]
    Q = np.diag(h) @ D
    B = np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.abs(Q + Q.T - B).max() < 2e-14


# diagonal norms of the published operators: SBP2 (1/2 at the ends), SBP6
# (Strand 1994; Mattsson and Nordstrom 2004, sixth-order interior with
# third-order boundary closure)
HL_EXACT_OP = {2: [Fr(1, 2)],
               4: HL_EXACT,
               6: [Fr(13649, 43200), Fr(12013, 8640), Fr(2711, 4320), Fr(5359, 4320),
                   Fr(7877, 8640), Fr(43801, 43200)]}
# (interior order, closure order) of each operator
ORDERS = {2: (2, 1), 4: (4, 2), 6: (6, 3)}


@pytest.mark.parametrize("order", [2, 4, 6])
def test_sbp_tables_are_summation_by_parts(order):
    """'SBP2', 'SBP4', 'SBP6' (fd_coeff.f90:79-246): the norm is the
    published one, H00i = 1/H(0,0), the dense operator satisfies
    H D + (H D)^T = diag(-1, 0, ..., 0, 1) (summation by parts), and it
    differentiates polynomials exactly through the interior order inside and
    the closure order in the boundary rows (the tables are 15-digit
    decimals, hence the tolerances)"""
    DIo, DLo, HL, H00i = ob.sbp_tables(order)
    nbst, nbnd = DLo.shape
    np.testing.assert_allclose(HL, [float(v) for v in HL_EXACT_OP[order]], rtol=0, atol=5e-15)
    assert abs(H00i - 1.0 / float(HL_EXACT_OP[order][0])) < 5e-15
    n = 31
    D = dense_D_op(n, order)
    h = norm_weights(n, order)
    Q = np.diag(h) @ D
    B = np.zeros((n, n))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.abs(Q + Q.T - B).max() < 5e-14
    pin, pbnd = ORDERS[order]
    x = np.arange(n, dtype=float) - 7.0
    for p in range(pin + 1):
        d = D @ x ** p
        want = p * x ** (p - 1) if p else np.zeros(n)
        scale = max(1.0, np.abs(x).max() ** max(p - 1, 0))
        inner = slice(nbst, n - nbst)
        assert np.abs(d[inner] - want[inner]).max() < 1e-12 * scale, ("interior", p)
        if p <= pbnd:
            assert np.abs(d - want).max() < 1e-12 * scale, ("closure", p)
    # the closure rows are not exact one degree higher (the order is sharp)
    d = D @ x ** (pbnd + 1)
    assert np.abs(d[:nbst] - (pbnd + 1) * x[:nbst] ** pbnd).max() > 1e-6
    # diff (fd.f90:42-80) applies exactly this matrix
    f = np.random.default_rng(order).standard_normal(n)
    np.testing.assert_allclose(ob.diff(f, order), D @ f, rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [9, 12, 31])
def test_diff_matches_dense_operator(n):
    rng = np.random.default_rng(n)
    f = rng.standard_normal(n)
    assert np.allclose(ob.diff(f), dense_D(n) @ f, rtol=0, atol=1e-13)
    x = np.arange(n, dtype=float)
    assert np.allclose(ob.diff(3.0 * x ** 2 - x + 2.0), 6.0 * x - 1.0, rtol=0, atol=1e-11)


# ---------------------------------------------------------------------------
# LSRK coefficients
# ---------------------------------------------------------------------------

LSRK = {
    # Carpenter and Kennedy (1994), five-stage fourth order, solution 3
    4: ([0, Fr(-567301805773, 1357537059087), Fr(-2404267990393, 2016746695238),
         Fr(-3550918686646, 2091501179385), Fr(-1275806237668, 842570457699)],
        [Fr(1432997174477, 9575080441755), Fr(5161836677717, 13612068292357),
         Fr(1720146321549, 2090206949498), Fr(3134564353537, 4481467310338),
         Fr(2277821191437, 14882151754819)]),
    # Williamson (1980) three-stage third order
    3: ([0, Fr(-5, 9), Fr(-153, 128)], [Fr(1, 3), Fr(15, 16), Fr(8, 15)]),
}


def lsrk_apply(A, B, Cc, f, y0, t0, h):
    y, dy = y0, 0.0
    for s in range(len(A)):
        dy = A[s] * dy + h * f(t0 + Cc[s] * h, y)
        y = y + B[s] * dy
    return y


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_lsrk_stability_polynomial_and_stage_times(order):
    A, B, Cc = ob.rk_coeffs(order)
    if order in LSRK:
        Ae, Be = LSRK[order]
        assert np.allclose(A, [float(v) for v in Ae], rtol=1e-15, atol=0)
        assert np.allclose(B, [float(v) for v in Be], rtol=1e-15, atol=0)
    # stability polynomial R(z) = sum c_k z^k: c_k = 1/k! for k <= order
    # (read off by applying the scheme to y' = z y at several z)
    zs = np.linspace(-0.5, 0.5, len(A) + 1)
    R = [lsrk_apply(A, B, Cc, lambda t, y, z=z: z * y, 1.0, 0.0, 1.0) for z in zs]
    coef = np.polyfit(zs, R, len(A))[::-1]
    for k in range(order + 1):
        assert abs(coef[k] - 1.0 / math.factorial(k)) < 1e-9, (k, coef[k])
    # stage times: y' = t^p integrated exactly for p < order
    for p in range(order):
        y1 = lsrk_apply(A, B, Cc, lambda t, y: t ** p, 0.0, 0.3, 0.7)
        assert abs(y1 - (1.0 ** (p + 1) - 0.3 ** (p + 1)) / (p + 1)) < 1e-13, p
    assert Cc[len(A)] == 1.0


# ---------------------------------------------------------------------------
# friction
# ---------------------------------------------------------------------------

RS = np.array([0.010, 0.014, 1e-6, 0.6, 0.4, 0.1, 0.1])  # a, b, V0, f0, L, fw, Vw
ETA = 0.5 * 2.67 * 3.464  # half the shear impedance


@pytest.mark.parametrize("law", ["RSL", "RFL", "RAL", "RSF"])
@pytest.mark.parametrize("Slock,N,Psi", [(70.0, 120.0, 0.70), (95.0, 120.0, 0.45),
                                          (-40.0, 60.0, 0.6), (72.0, 120.0, 0.8)])
def test_regularized_laws_residual_and_inverse(law, Slock, N, Psi):
    a, V0 = RS[0], RS[2]
    rc, V, S, Nout, O = ob.solve_friction_node(law, RS, Slock, N, 0.0, 0.0, ETA, 0.0, Psi, 1e-9)
    assert rc == 0 and Nout == N and O == 0.0
    assert abs(S + ETA * V - Slock) <= 1e-12 * abs(Slock)
    # the strength a N asinh(V e^(Psi/a) / (2 V0)) equals S
    assert abs(a * N * np.arcsinh(V / (2 * V0) * np.exp(Psi / a)) - S) <= 1e4 * 2.3e-16 * 10
    # closed-form inverse
    Vinv = 2 * V0 * np.exp(-Psi / a) * np.sinh(S / (a * N))
    assert abs(Vinv - V) <= 1e-8 * abs(V) + 1e-300


@pytest.mark.parametrize("law", ["SL", "FL"])
def test_unregularized_laws_residual(law):
    a, V0 = RS[0], RS[2]
    for Slock, N, Psi in [(70.0, 120.0, 0.62), (80.0, 110.0, 0.55)]:
        rc, V, S, _, _ = ob.solve_friction_node(law, RS, Slock, N, 0.0, 0.0, ETA, 0.0, Psi, 1e-9)
        assert rc == 0 and V > 0
        assert abs(S + ETA * V - Slock) <= 1e-12 * Slock
        assert abs(N * (a * np.log(V / V0) + Psi) - S) <= 1e-10


def test_loading_tensile_and_simple_branches():
    # Slock = S0 + phis, Nlock = N0 - (1 - skempton) phip
    rc, V, S, N, _ = ob.solve_friction_node("RSL", RS, 60.0, 100.0, 5.0, 3.0, ETA, 0.0, 0.7,
                                            1e-9, skempton=0.2)
    assert N == 100.0 - 0.8 * 5.0
    assert abs(S + ETA * V - 63.0) < 1e-12 * 63.0
    # tensile: no strength, V = Slock / eta
    rc, V, S, N, _ = ob.solve_friction_node("RFL", RS, 30.0, -5.0, 0.0, 0.0, ETA, 0.0, 0.7, 1e-9)
    assert S == 0.0 and V == 30.0 / ETA and N == -5.0
    rc, V, S, N, _ = ob.solve_friction_node("frictionless", RS, 30.0, 50.0, 0.0, 0.0, ETA, 0.0,
                                            0.7, 1e-9)
    assert S == 0.0 and V == 30.0 / ETA
    rc, V, S, N, _ = ob.solve_friction_node("locked", RS, 30.0, 50.0, 0.0, 0.0, ETA, 0.0, 0.7,
                                            1e-9)
    assert V == 0.0 and S == 30.0


def test_slip_weakening_branches():
    sw = np.array([1.0, 0.677, 0.525, 0.4])  # cohesion, fs, fd, Dc
    N = 120.0
    for D, Slock in [(0.0, 70.0), (0.1, 70.0), (0.1, 90.0), (0.5, 70.0), (0.5, -70.0)]:
        Sv = sw[0] + N * (sw[1] - (sw[1] - sw[2]) * min(D, sw[3]) / sw[3])
        rc, V, S, _, _ = ob.solve_friction_node("SW", RS, Slock, N, 0.0, 0.0, ETA, D, 0.0, 0.0,
                                                sw=sw)
        if abs(Slock) >= Sv:
            assert abs(S - np.sign(Slock) * Sv) < 1e-12 * Sv
            assert abs(V - (Slock - S) / ETA) < 1e-14 * abs(Slock)
        else:
            assert V == 0.0 and S == Slock


def _fss(law, V):
    a, b, V0, f0, L, fw, Vw = RS
    fLV = f0 - (b - a) * np.log(V / V0)
    if law in ("SL", "RSL"):
        return fLV
    return fw + (fLV - fw) / (1.0 + (V / Vw) ** 8) ** 0.125  # flash heating


@pytest.mark.parametrize("law", ["SL", "FL", "RSL", "RFL", "RSF", "RAL"])
@pytest.mark.parametrize("V", [1e-9, 1e-3, 0.5, 3.0])
def test_state_evolution_steady_state(law, V):
    a, b, V0, f0, L, fw, Vw = RS
    if law == "RAL":  # regularised aging law
        Psi_ss = f0 - b * np.log(V / V0)
        slope = -(b * V0 / L) * np.exp((f0 - Psi_ss) / b) / b
    elif law in ("SL", "FL"):  # f = a ln(V/V0) + Psi
        Psi_ss = _fss(law, V) - a * np.log(V / V0)
        slope = -V / L
    else:  # f = a asinh(V e^(Psi/a) / (2 V0))
        Psi_ss = a * np.log(2 * V0 / V * np.sinh(_fss(law, V) / a))
        slope = -V / L
        if law != "RSF":
            X = V / (2 * V0) * np.exp(Psi_ss / a)
            slope *= X / np.sqrt(1 + X * X)  # d f / d Psi
    scale = abs(slope) * max(1.0, abs(Psi_ss))
    assert abs(ob.state_rate(law, RS, V, Psi_ss)) <= 1e-10 * scale
    d = 1e-6
    num = (ob.state_rate(law, RS, V, Psi_ss + d) - ob.state_rate(law, RS, V, Psi_ss - d)) / (2 * d)
    assert abs(num - slope) <= 1e-5 * abs(slope)


# ---------------------------------------------------------------------------
# plasticity
# ---------------------------------------------------------------------------


def _dp_state(s, s0):
    sxx, sxy, syy, szz = s0[0] + s[0], s0[1] + s[1], s0[3] + s[2], s0[5] + s[3]
    sig = (sxx + syy + szz) / 3.0
    J2 = ((sxx - syy) ** 2 + (syy - szz) ** 2 + (szz - sxx) ** 2) / 6.0 + sxy ** 2 + s0[2] ** 2 + s0[4] ** 2
    return sig, np.sqrt(J2), np.array([sxx - sig, sxy, syy - sig, szz - sig])


def test_plastic_return_onto_yield_surface():
    G, K, mu, dt = 32.0, 60.0, 0.6, 1e-3
    s0 = np.zeros(6)
    s = np.array([-60.0, 44.0, -100.0, -80.0])
    sig, tau, dev = _dp_state(s, s0)
    assert tau - (-mu * sig) > 0  # above yield
    s1, Ds1, lam, gp = ob.plastic_point(G, K, mu, 0.0, 0.0, 0.0, 0.0, s0, dt, s, np.zeros(4),
                                        0.0, 0.0)
    sig1, tau1, dev1 = _dp_state(s1, s0)
    assert abs(tau1 - (-mu * sig1)) < 1e-12 * tau
    assert abs(sig1 - sig) < 1e-12 * abs(sig)  # beta = 0: no dilatancy
    assert np.allclose(dev1 / tau1, dev / tau, rtol=1e-12, atol=1e-14)  # radial return
    assert abs(lam - (tau + mu * sig) / (dt * G)) < 1e-10 * lam
    # plastic.f90 increments gammap by dt*lambda twice (lines 129 and 177)
    assert abs(gp - 2.0 * dt * lam) < 1e-15 * gp
    # rates pick up -lambda * G * dev / tau
    assert np.allclose(Ds1, -lam * G * dev / tau, rtol=1e-12, atol=1e-12)
    # below yield: nothing changes
    s_el = np.array([-100.0, 10.0, -100.0, -100.0])
    s2, Ds2, lam2, gp2 = ob.plastic_point(G, K, mu, 0.0, 0.0, 0.0, 0.0, s0, dt, s_el,
                                          np.ones(4), 0.0, 0.0)
    assert lam2 == 0.0 and gp2 == 0.0 and np.array_equal(s2, s_el) and np.array_equal(Ds2, np.ones(4))


# ---------------------------------------------------------------------------
# manufactured solution
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("side", [1, 2])
def test_mms_sources_are_residuals_of_the_fields(side):
    G, cs, cp = 32.0, 3.0, 5.0
    rho = G / cs ** 2
    lam = rho * cp ** 2 - 2 * G
    M = lam + 2 * G
    u = lambda x, y, t, l: ob.mms_sin(x, y, t, side, l)

    def d(fn, i, h=1e-4):  # fourth-order central difference along argument i
        def g(*args):
            a = list(args)
            vals = []
            for k in (-2, -1, 1, 2):
                b = list(a)
                b[i] += k * h
                vals.append(fn(*b))
            return (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
        return g

    rng = np.random.default_rng(side)
    for x, y, t in rng.uniform([-0.5, -0.5, 0.0], [0.5, 0.5, 0.3], (6, 3)):
        D = {(l, i): d(lambda X, Y, T, l=l: u(X, Y, T, l), i)(x, y, t) for l in range(6) for i in range(3)}
        L = [(D[2, 0] + D[3, 1]) / rho, (D[3, 0] + D[4, 1]) / rho,
             M * D[0, 0] + lam * D[1, 1], G * (D[1, 0] + D[0, 1]),
             lam * D[0, 0] + M * D[1, 1], lam * (D[0, 0] + D[1, 1])]
        for l in range(6):
            res = D[l, 2] - L[l]
            src = u(x, y, t, 6 + l)
            scale = max(1.0, abs(D[l, 2]), abs(L[l]))
            assert abs(res - src) < 1e-6 * scale, (l, res, src)


# ---------------------------------------------------------------------------
# whole stage
# ---------------------------------------------------------------------------


def dense_D_op(n, order):
    """dense derivative matrix of 'SBP<order>' from the oracle's tables
    (pinned by test_sbp_tables_are_summation_by_parts)"""
    if order == 4:
        return dense_D(n)
    DIo, DLo, _, _ = ob.sbp_tables(order)
    nbst, nbnd = DLo.shape
    pI = (len(DIo) - 1) // 2
    D = np.zeros((n, n))
    for i in range(nbst):
        D[i, :nbnd] = DLo[i]
        D[n - 1 - i, n - nbnd:] = -DLo[i][::-1]
    for i in range(nbst, n - nbst):
        D[i, i - pI:i + pI + 1] = DIo
    return D


def norm_weights(n, order=4):
    """diagonal of the SBP norm H (unit spacing) of 'SBP<order>'"""
    _, _, HL, _ = ob.sbp_tables(order)
    h = np.ones(n)
    h[:len(HL)] = HL
    h[n - len(HL):] = HL[::-1]
    return h


def split_ops(prob, order=4):
    Dq = dense_D_op(prob.nx, order)
    Dr = dense_D_op(prob.ny, order)
    dq = lambda f: f @ Dq.T
    dr = lambda f: Dr @ f
    sq = lambda a, f: 0.5 * (dq(a * f) + a * dq(f))
    sr = lambda a, f: 0.5 * (dr(a * f) + a * dr(f))
    dx = lambda f: prob.Ji * (sq(prob.yr, f) - sr(prob.yq, f))
    dy = lambda f: prob.Ji * (-sq(prob.xr, f) + sr(prob.xq, f))
    return dx, dy


@pytest.mark.parametrize("seed,warp", [(1, 0.0), (2, 0.08)])
def test_stage1_rates_match_dense_split_operator(seed, warp):
    cs, cp, rho = 3.464, 6.0, 2.67
    prob, F = problems.build(nx=13, nys=(15,), seed=seed, warp=warp, cs=cs, cp=cp, rho=rho,
                             background=[0.1, -0.2, -120.0, 70.0, -110.0, -100.0], perturb=0.5,
                             bcs=("absorbing", "free", "absorbing", "free"))
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    G = rho * cs ** 2
    M = rho * cp ** 2
    lam = M - 2 * G
    gamma = lam / M
    nu = lam / (2 * (lam + G))
    dx, dy = split_ops(prob)
    vx, vy, sxx, sxy, syy = F[:5]
    ref = np.empty_like(DF)
    ref[0] = (dx(sxx) + dy(sxy)) / rho
    ref[1] = (dx(sxy) + dy(syy)) / rho
    ref[2] = M * (dx(vx) + gamma * dy(vy))
    ref[3] = G * (dx(vy) + dy(vx))
    ref[4] = M * (gamma * dx(vx) + dy(vy))
    ref[5] = nu * (ref[2] + ref[4])
    # boundary nodes carry SAT penalties on top; every other point, closure
    # rows and corners included, is the operator alone
    inner = (slice(None), slice(1, -1), slice(1, -1))
    scale = np.abs(ref[inner]).max(axis=(1, 2), keepdims=True)
    err = np.abs(DF[inner] - ref[inner]) / scale
    assert err.max() < 1e-12, err.max()


@pytest.mark.parametrize("order,warp", [(2, 0.08), (6, 0.0), (6, 0.08)])
def test_stage1_rates_other_operators_match_dense_split_operator(order, warp):
    """the same split-form check with 'SBP2' and 'SBP6': interior stencils
    of half-width 1 and 3, closures of 1 x 2 and 6 x 9 points and their
    corners (rates.f90:384-917 with the operator's tables), the metrics
    from the same operator"""
    cs, cp, rho = 3.464, 6.0, 2.67
    n = 21 if order == 6 else 13
    prob, F = problems.build(nx=n, nys=(n + 2,), seed=order, warp=warp, cs=cs, cp=cp, rho=rho,
                             background=[0.1, -0.2, -120.0, 70.0, -110.0, -100.0], perturb=0.5,
                             bcs=("absorbing", "free", "absorbing", "free"), sbp=order)
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    G, M = rho * cs ** 2, rho * cp ** 2
    lam = M - 2 * G
    gamma, nu = lam / M, lam / (2 * (lam + G))
    dx, dy = split_ops(prob, order)
    vx, vy, sxx, sxy, syy = F[:5]
    ref = np.empty_like(DF)
    ref[0] = (dx(sxx) + dy(sxy)) / rho
    ref[1] = (dx(sxy) + dy(syy)) / rho
    ref[2] = M * (dx(vx) + gamma * dy(vy))
    ref[3] = G * (dx(vy) + dy(vx))
    ref[4] = M * (gamma * dx(vx) + dy(vy))
    ref[5] = nu * (ref[2] + ref[4])
    inner = (slice(None), slice(1, -1), slice(1, -1))
    scale = np.abs(ref[inner]).max(axis=(1, 2), keepdims=True)
    err = np.abs(DF[inner] - ref[inner]) / scale
    assert err.max() < 1e-12, err.max()
    # and mode 3 with the same operator
    prob3, F3 = problems.build(nx=n, nys=(n + 2,), seed=order + 1, warp=warp, cs=cs, cp=cp,
                               rho=rho, background=[0.1, 30.0, -5.0], perturb=0.5, mode=3,
                               bcs=("absorbing", "free", "absorbing", "free"), sbp=order)
    o3 = OracleContext(prob3)
    o3.set_fields(F3)
    o3.init_edges(0.0)
    o3.stage1(0.0, 1e-3, 1)
    _, DF3 = o3.get_fields()
    dx3, dy3 = split_ops(prob3, order)
    vz, sxz, syz = F3
    ref3 = np.array([(dx3(sxz) + dy3(syz)) / rho, G * dx3(vz), G * dy3(vz)])
    inner3 = (slice(None), slice(1, -1), slice(1, -1))
    scale3 = np.abs(ref3[inner3]).max(axis=(1, 2), keepdims=True)
    assert (np.abs(DF3[inner3] - ref3[inner3]) / scale3).max() < 1e-12


@pytest.mark.parametrize("seed,warp", [(5, 0.0), (6, 0.08)])
def test_stage1_rates_heterogeneous_match_dense_split_operator(seed, warp):
    """rates_h_*_mode2 (rates_heterogeneous.f90:375-917): the material of
    each point scales the same split-form operator, and gamma multiplies the
    terms of the coupled velocity inside it"""
    cs, cp, rho = 3.464, 6.0, 2.67
    prob, F = problems.build(nx=13, nys=(15,), seed=seed, warp=warp, cs=cs, cp=cp, rho=rho,
                             background=[0.1, -0.2, -120.0, 70.0, -110.0, -100.0], perturb=0.5,
                             bcs=("absorbing", "free", "absorbing", "free"), medium=0.3)
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    r, s_, p_ = prob.medium
    assert np.ptp(r) > 0.1 * rho and np.ptp(s_) > 0.1 * cs
    G = r * s_ ** 2
    M = r * p_ ** 2
    gamma = 1 - 2 * (s_ / p_) ** 2
    nu = 0.5 * (M - 2 * G) / (M - G)
    dx, dy = split_ops(prob)
    vx, vy, sxx, sxy, syy = F[:5]
    ref = np.empty_like(DF)
    ref[0] = (dx(sxx) + dy(sxy)) / r
    ref[1] = (dx(sxy) + dy(syy)) / r
    ref[2] = M * (dx(vx) + gamma * dy(vy))
    ref[3] = G * (dx(vy) + dy(vx))
    ref[4] = M * (gamma * dx(vx) + dy(vy))
    ref[5] = nu * (ref[2] + ref[4])
    inner = (slice(None), slice(1, -1), slice(1, -1))
    scale = np.abs(ref[inner]).max(axis=(1, 2), keepdims=True)
    err = np.abs(DF[inner] - ref[inner]) / scale
    assert err.max() < 1e-12, err.max()


def test_uniform_medium_matches_homogeneous_material():
    """a medium file holding the block's own constants: the heterogeneous
    path differs from the homogeneous one only in rounding order"""
    cs, cp, rho = 3.464, 6.0, 2.67
    kw = dict(nx=40, nys=(20, 24), seed=9, warp=0.03, cs=cs, cp=cp, rho=rho,
              background=[0.0, 0.0, -120.0, 70.0, -120.0, -120.0], perturb=1e-2, law="RFL",
              plastic=dict(mu=0.6, beta=0.0, b=0.0, h=0.0, eta=0.0, s0=[0.0] * 6),
              fault_overstress=15.0)
    out = []
    for uniform in (False, True):
        prob, F = problems.build(**kw)
        if uniform:
            prob.medium = (np.full((prob.ny, prob.nx), rho), np.full((prob.ny, prob.nx), cs),
                           np.full((prob.ny, prob.nx), cp))
        o = OracleContext(prob)
        o.set_fields(F)
        o.init_edges(0.0)
        o.step(0.0, 1e-3, 1, 20)
        out.append((o.get_fields()[0], o.get_iface(0, "V"), o.get_plastic()[0]))
    a, b = out
    assert np.abs(a[0] - b[0]).max() / np.abs(a[0]).max() < 1e-13
    assert np.abs(a[1] - b[1]).max() / np.abs(a[1]).max() < 1e-12
    assert np.abs(a[2] - b[2]).max() <= 1e-12 * max(1.0, np.abs(a[2]).max())


@pytest.mark.parametrize("n", [12, 17, 40])
def test_dissipation_operator_is_negative_semidefinite_in_the_norm(n):
    """AL/AI/AR of fd_coeff.f90:103-124, 240-246: H A is symmetric and
    negative semidefinite, with null space exactly the linear functions
    (Mattsson, Svard and Nordstrom 2004); the interior rows are the fourth
    difference -(1, -4, 6, -4, 1)/12"""
    A = np.array([ob.diss1d(np.eye(n)[k]) for k in range(n)]).T
    _, _, HL, _ = ob.sbp4_tables()
    h = np.ones(n)
    h[:4] = HL
    h[-4:] = HL[::-1]
    HA = h[:, None] * A
    assert np.abs(HA - HA.T).max() < 1e-14
    ev = np.linalg.eigvalsh(0.5 * (HA + HA.T))
    assert ev.max() < 1e-14
    assert np.sum(np.abs(ev) < 1e-12) == 2
    x = np.arange(n, dtype=float)
    assert np.abs(A @ np.ones(n)).max() < 1e-14 and np.abs(A @ x).max() < 1e-12
    for i in range(4, n - 4):
        np.testing.assert_allclose(A[i, i - 2:i + 3], np.array([-1, 4, -6, 4, -1]) / 12.0,
                                   rtol=0, atol=1e-15)


@pytest.mark.parametrize("medium", [None, 0.3])
def test_stage1_dissipation_matches_dense_operator(medium):
    """rates_dissipation_*_mode2 (rates.f90:1397-2021, rates_heterogeneous.f90:
    1376-1997): the rates with Cdiss exceed those without by
    Cdiss Ji (A_q + A_r)(J F) in each block, fields 1..5, with A the
    one-dimensional operator; DF6 follows through szz"""
    kw = dict(nx=26, nys=(17, 19), seed=4, warp=0.06, background=[0.1, -0.2, -120.0, 70.0, -110.0, -100.0],
              perturb=0.5, coupling="locked", medium=medium)
    out = []
    for Cd in (0.0, 2000.0):
        prob, F = problems.build(Cdiss=Cd, **kw)
        o = OracleContext(prob)
        o.set_fields(F)
        o.init_edges(0.0)
        o.stage1(0.0, 1e-3, 1)
        out.append(o.get_fields()[1])
    diff = out[1] - out[0]
    JF = prob.J[None] * F
    ref = np.zeros_like(diff)
    for b in prob.blocks:
        sj, si = slice(b["my"], b["py"] + 1), slice(b["mx"], b["px"] + 1)
        for l in range(5):
            blk = JF[l, sj, si]
            aq = np.array([ob.diss1d(row) for row in blk])
            ar = np.array([ob.diss1d(col) for col in blk.T]).T
            ref[l, sj, si] = 2000.0 * prob.Ji[sj, si] * (aq + ar)
    nu = prob.blocks[0]["nu"]
    if medium is None:
        ref[5] = nu * (ref[2] + ref[4])
    scale = np.abs(out[0]).max(axis=(1, 2), keepdims=True)
    err = np.abs(diff - ref)[:5 if medium else 6] / scale[:5 if medium else 6]
    assert np.all(np.abs(ref[:5]).max(axis=(1, 2)) > 0.05 * scale[:5, 0, 0])
    assert err.max() < 1e-12, err.max()


@pytest.mark.parametrize("medium", [None, 0.3])
def test_stage1_dissipation_mode3_matches_dense_operator(medium):
    """rates_dissipation_*_mode3 (rates.f90:939-1394, rates_heterogeneous.f90:
    920-1373): the antiplane rates with Cdiss exceed those without by
    Cdiss Ji (A_q + A_r)(J F) in each block, for vz, sxz and syz"""
    kw = dict(nx=26, nys=(17, 19), seed=5, warp=0.06, background=[0.1, 30.0, -20.0],
              perturb=0.5, coupling="locked", medium=medium, mode=3,
              inplane=(-120.0, 0.0, -100.0, -110.0))
    out = []
    for Cd in (0.0, 2000.0):
        prob, F = problems.build(Cdiss=Cd, **kw)
        o = OracleContext(prob)
        o.set_fields(F)
        o.init_edges(0.0)
        o.stage1(0.0, 1e-3, 1)
        out.append(o.get_fields()[1])
    diff = out[1] - out[0]
    JF = prob.J[None] * F
    ref = np.zeros_like(diff)
    for b in prob.blocks:
        sj, si = slice(b["my"], b["py"] + 1), slice(b["mx"], b["px"] + 1)
        for l in range(3):
            blk = JF[l, sj, si]
            aq = np.array([ob.diss1d(row) for row in blk])
            ar = np.array([ob.diss1d(col) for col in blk.T]).T
            ref[l, sj, si] = 2000.0 * prob.Ji[sj, si] * (aq + ar)
    scale = np.abs(out[0]).max(axis=(1, 2), keepdims=True)
    err = np.abs(diff - ref) / scale
    assert np.all(np.abs(ref).max(axis=(1, 2)) > 0.05 * scale[:, 0, 0])
    assert err.max() < 1e-12, err.max()


def energy(prob, F, rho, cs, cp, order=4):
    W = np.outer(norm_weights(prob.ny, order), norm_weights(prob.nx, order)) / prob.Ji
    G = rho * cs ** 2
    lam = rho * cp ** 2 - 2 * G
    nu = lam / (2 * (lam + G))
    vx, vy, sxx, sxy, syy = F[:5]
    e = 0.5 * rho * (vx ** 2 + vy ** 2) + ((1 - nu) * (sxx ** 2 + syy ** 2) - 2 * nu * sxx * syy
                                         + 2 * sxy ** 2) / (4 * G)
    return float((W * e).sum())


@pytest.mark.parametrize("bcs,order", [(("absorbing",) * 4, 4),
                                       (("absorbing", "free", "absorbing", "free"), 4),
                                       (("absorbing", "free", "absorbing", "free"), 6),
                                       (("absorbing", "free", "absorbing", "free"), 2)])
def test_energy_decays_with_absorbing_boundaries(bcs, order):
    cs, cp, rho, h = 3.464, 6.0, 2.67, 0.025
    prob, F = problems.build(nx=31, nys=(27,), seed=4, warp=0.06, h=h, cs=cs, cp=cp, rho=rho,
                             background=[0.0] * 6, perturb=1.0, bcs=bcs, sbp=order)
    dt = 0.5 * h / np.hypot(cs, cp)
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    E = [energy(prob, F, rho, cs, cp, order)]
    for n in range(1, 61):
        o.step(0.0, dt, n, 1)
        E.append(energy(prob, o.get_fields()[0], rho, cs, cp, order))
    E = np.array(E)
    assert (np.diff(E) <= 1e-12 * E[0]).all(), np.diff(E).max() / E[0]
    assert E[-1] < 0.95 * E[0]


# ---- antiplane shear (mode 3) ----

@pytest.mark.parametrize("seed,warp", [(11, 0.0), (12, 0.08)])
def test_stage1_rates_mode3_match_dense_split_operator(seed, warp):
    """rates_interior_mode3 / rates_boundary_mode3 (rates.f90:8-381):
    DF1 = (dx sxz + dy syz)/rho, DF2 = G dx vz, DF3 = G dy vz with the
    split-form curvilinear operator, closures and corners included"""
    cs, cp, rho = 3.464, 6.0, 2.67
    prob, F = problems.build(nx=13, nys=(15,), seed=seed, warp=warp, cs=cs, cp=cp, rho=rho,
                             background=[0.1, 30.0, -20.0], perturb=0.5, mode=3,
                             bcs=("absorbing", "free", "rigid", "free"))
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    assert DF.shape == (3, prob.ny, prob.nx)
    G = rho * cs ** 2
    dx, dy = split_ops(prob)
    vz, sxz, syz = F
    ref = np.array([(dx(sxz) + dy(syz)) / rho, G * dx(vz), G * dy(vz)])
    inner = (slice(None), slice(1, -1), slice(1, -1))
    scale = np.abs(ref[inner]).max(axis=(1, 2), keepdims=True)
    err = np.abs(DF[inner] - ref[inner]) / scale
    assert err.max() < 1e-12, err.max()


@pytest.mark.parametrize("seed,warp", [(13, 0.0), (14, 0.08)])
def test_stage1_rates_mode3_heterogeneous_match_dense_split_operator(seed, warp):
    """rates_h_interior_mode3 / rates_h_boundary_mode3 (rates_heterogeneous.
    f90:8-372): the point's rho and G scale the same split-form operator,
    DF1 = (dx sxz + dy syz)/rho, DF2 = G dx vz, DF3 = G dy vz"""
    cs, cp, rho = 3.464, 6.0, 2.67
    prob, F = problems.build(nx=13, nys=(15,), seed=seed, warp=warp, cs=cs, cp=cp, rho=rho,
                             background=[0.1, 30.0, -20.0], perturb=0.5, mode=3,
                             bcs=("absorbing", "free", "rigid", "free"), medium=0.3)
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    r, s_, _ = prob.medium
    assert np.ptp(r) > 0.1 * rho and np.ptp(s_) > 0.1 * cs
    G = r * s_ ** 2
    dx, dy = split_ops(prob)
    vz, sxz, syz = F
    ref = np.array([(dx(sxz) + dy(syz)) / r, G * dx(vz), G * dy(vz)])
    inner = (slice(None), slice(1, -1), slice(1, -1))
    scale = np.abs(ref[inner]).max(axis=(1, 2), keepdims=True)
    err = np.abs(DF[inner] - ref[inner]) / scale
    assert err.max() < 1e-12, err.max()


def test_mode3_uniform_medium_matches_homogeneous_material():
    """antiplane shear with a medium file holding the block's own constants:
    the heterogeneous path differs from the homogeneous one only in
    rounding order, through a rupture"""
    cs, cp, rho = 3.464, 6.0, 2.67
    kw = dict(nx=40, nys=(20, 24), seed=9, warp=0.03, cs=cs, cp=cp, rho=rho,
              background=[0.0, 0.0, 70.0], perturb=1e-2, law="RFL", fault_overstress=25.0,
              mode=3, inplane=(-120.0, 0.0, -100.0, -110.0))
    out = []
    for uniform in (False, True):
        prob, F = problems.build(**kw)
        if uniform:
            prob.medium = (np.full((prob.ny, prob.nx), rho), np.full((prob.ny, prob.nx), cs),
                           np.full((prob.ny, prob.nx), cp))
        o = OracleContext(prob)
        o.set_fields(F)
        o.init_edges(0.0)
        o.step(0.0, 1e-3, 1, 40)
        out.append((o.get_fields()[0], o.get_iface(0, "V")))
    a, b = out
    assert np.abs(a[0] - b[0]).max() / np.abs(a[0]).max() < 1e-13
    assert np.abs(a[1] - b[1]).max() / np.abs(a[1]).max() < 1e-12
    assert not np.array_equal(a[0], b[0])  # the two paths round differently


def energy3(prob, F, rho, cs):
    _, _, HL, _ = ob.sbp4_tables()

    def hw(n):
        h = np.ones(n)
        h[:4] = HL
        h[n - 4:] = HL[::-1]
        return h
    W = np.zeros((prob.ny, prob.nx))
    for b in prob.blocks:
        sj, si = slice(b["my"], b["py"] + 1), slice(b["mx"], b["px"] + 1)
        W[sj, si] += np.outer(hw(b["py"] - b["my"] + 1), hw(b["px"] - b["mx"] + 1)) / prob.Ji[sj, si]
    G = rho * cs ** 2
    vz, sxz, syz = F
    return float((W * (0.5 * rho * vz ** 2 + (sxz ** 2 + syz ** 2) / (2 * G))).sum())


@pytest.mark.parametrize("bcs,coupling", [(("absorbing",) * 4, None),
                                          (("free", "rigid", "free", "rigid"), None),
                                          (("absorbing", "free", "rigid", "free"), "locked"),
                                          (("free", "free", "rigid", "free"), "frictionless")])
def test_mode3_energy_does_not_grow(bcs, coupling):
    """set_bc_mode3 (boundaries.f90:155-213), the mode-3 SAT (time_step.f90:
    829-830) and the mode-3 interfaces (interfaces.f90:676-712, 806-897) are
    energy stable: the discrete energy never grows, and absorbing sides
    drain it"""
    cs, cp, rho, h = 3.464, 6.0, 2.67, 0.025
    nys = (27,) if coupling is None else (15, 17)
    kw = dict(coupling="locked") if coupling == "locked" else dict(coupling="friction", law=coupling)
    if coupling is None:
        kw = {}
    prob, F = problems.build(nx=31, nys=nys, seed=4, warp=0.06, h=h, cs=cs, cp=cp, rho=rho,
                             background=[0.0] * 3, perturb=1.0, bcs=bcs, mode=3,
                             inplane=(-50.0, 0.0, -50.0, -50.0), fault_overstress=0.0, **kw)
    dt = 0.5 * h / np.hypot(cs, cp)
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    E = [energy3(prob, F, rho, cs)]
    for n in range(1, 61):
        o.step(0.0, dt, n, 1)
        E.append(energy3(prob, o.get_fields()[0], rho, cs))
    E = np.array(E)
    assert (np.diff(E) <= 1e-12 * E[0]).all(), np.diff(E).max() / E[0]
    if "absorbing" in bcs:
        assert E[-1] < 0.95 * E[0]


def test_mode3_steady_sliding_under_resolved_prestress():
    """frictional_interface_mode3: S0 is the background (sxz, syz) resolved
    on the fault and N0 the plus block's in-plane prestress (interfaces.f90:
    859-876).  With the state for steady sliding at V = 1e-12 under those,
    the unperturbed fault keeps sliding at that rate"""
    prob, F = problems.build(nx=40, nys=(20, 22), seed=3, warp=0.0, background=[0.0, 0.0, 40.0],
                             perturb=0.0, mode=3, law="RFL", fault_overstress=0.0,
                             inplane=(-90.0, 0.0, -80.0, -85.0), bcs=("free",) * 4)
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.step(0.0, 1e-3, 1, 5)
    V = o.get_iface(0, "V")
    N = o.get_iface(0, "N")
    np.testing.assert_allclose(N, 80.0, rtol=1e-13)
    np.testing.assert_allclose(o.get_iface(0, "S0"), 40.0, rtol=1e-13)
    np.testing.assert_allclose(V, 1e-12, rtol=1e-6)


@pytest.mark.parametrize("mode", [2, 3])
def test_kinematic_forcing_moves_the_rupture_tips(mode):
    """force = T (friction.f90:717-745, 765-767): with a static friction no
    prestress overcomes, only the forced strength c0 + N max(fkm, fkp) lets
    the fault slip; its tips xL + vL t and xR + vR t bound the slipping
    nodes, and inside them the fault slides at the forced strength"""
    v, dt = 2.0, 1e-3
    kn = dict(c0=0.0, fd=0.525, dfdx=20.0, vL=-v, vR=v, xL=0.05, xR=-0.05, tmax=1e9)
    if mode == 2:
        bg, kw = [0.0, 0.0, -120.0, 70.0, -120.0, -120.0], {}
    else:
        bg, kw = [0.0, 0.0, 70.0], dict(inplane=(-120.0, 0.0, -120.0, -120.0))
    prob, F = problems.build(nx=81, nys=(20, 22), seed=5, warp=0.0, background=bg, perturb=0.0,
                             law="SW", fault_overstress=0.0, mode=mode, force=kn, **kw)
    for itf in prob.ifaces:
        itf["sw"][1] = 50.0  # static friction out of reach
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    nsteps = 200
    o.step(0.0, dt, 1, nsteps)
    o.stage1(0.0, dt, nsteps + 1)
    t = nsteps * dt
    x = prob.ifaces[0]["kn_x"]
    V = o.get_iface(0, "V", stage1=True)
    S = o.get_iface(0, "S", stage1=True)
    N = o.get_iface(0, "N", stage1=True)
    tip = -0.05 + v * t
    # beyond the tips the forced strength exceeds the prestress by the
    # margin dfdx (|x| - tip) N: locked, exactly
    far = np.abs(x) > tip + 0.02
    assert np.all(V[far] == 0.0)
    inside = np.abs(x) < tip - 0.01
    assert inside.sum() > 5 and np.all(V[inside] > 0.0)
    np.testing.assert_allclose(S[inside], 0.525 * N[inside], rtol=1e-14)


def mms_error(n, tend=0.04, warp=0.05, order=4):
    G, cs, cp = 32.0, 3.0, 5.0
    rho = G / cs ** 2
    h = 1.0 / (n - 1)
    prob, F = problems.build(nx=n, nys=(n,), seed=8, warp=warp, h=h, cs=cs, cp=cp, rho=rho,
                             mms=True, bcs=("mms-sin-w",) * 4, sbp=order)
    nsteps = int(np.ceil(tend / (0.5 * h / np.hypot(cs, cp))))
    dt = tend / nsteps
    o = OracleContext(prob, threads=4)
    o.set_fields(F)
    o.init_edges(0.0)
    o.step(0.0, dt, 1, nsteps)
    Fo = o.get_fields()[0]
    ex = np.empty_like(Fo)
    for l in range(6):
        ex[l] = np.vectorize(lambda x, y: ob.mms_sin(x, y, tend, 1, l))(prob.x, prob.y)
    _, _, HL, _ = ob.sbp4_tables()
    err = Fo[:5] - ex[:5]
    return np.sqrt((err ** 2).sum(axis=(1, 2)) * h * h).sum()


@pytest.mark.parametrize("order,rate", [(4, 2.7), (6, 3.5), (2, 1.7)])
def test_mms_convergence_rate(order, rate):
    """the manufactured solution converges at the operator's global rate:
    about 3 for SBP4 (second-order closures), 4 for SBP6 (third-order
    closures), 2 for SBP2 (first-order closures)"""
    e = [mms_error(n, order=order) for n in (21, 41, 81)]
    rates = np.log2(np.array(e[:-1]) / np.array(e[1:]))
    print("MMS errors", order, e, "rates", rates)
    assert rates[-1] > rate, rates


def to_nt(F, n1, n2):
    """velocity and in-plane stress in the (tangent, normal) frame, tangent
    t = (n2, -n1): (v.t, v.n, t.s.t, n.s.t, n.s.n)"""
    vx, vy, sxx, sxy, syy = F[:5]
    t1, t2 = n2, -n1
    return (vx * t1 + vy * t2, vx * n1 + vy * n2,
            t1 * t1 * sxx + 2 * t1 * t2 * sxy + t2 * t2 * syy,
            n1 * t1 * sxx + (n1 * t2 + n2 * t1) * sxy + n2 * t2 * syy,
            n1 * n1 * sxx + 2 * n1 * n2 * sxy + n2 * n2 * syy)


@pytest.mark.parametrize("bcs", [("absorbing", "free", "rigid", "absorbing"),
                                 ("free", "rigid", "absorbing", "free")])
def test_boundary_hat_states_satisfy_characteristic_conditions(bcs):
    """At a boundary node the stage-1 rate is the operator plus the penalty
    Ks Fs(Fhat - F) + Kp Fp(Fhat - F); rotated to (t, n) it yields the hat
    differences.  Every hat state keeps the outgoing characteristics
    snt - Zs vt and snn - Zp vn, meets its own condition (absorbing: the
    incoming snt + Zs vt and snn + Zp vn vanish; free: snt = snn = 0; rigid:
    vt = vn = 0), and moves stt by gamma times the change of snn, as a
    P wave does."""
    cs, cp, rho = 3.464, 6.0, 2.67
    prob, F = problems.build(nx=13, nys=(14,), seed=12, warp=0.07, cs=cs, cp=cp, rho=rho,
                             background=[0.0] * 6, perturb=3.0, bcs=bcs)
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    G, M = rho * cs ** 2, rho * cp ** 2
    gamma = (M - 2 * G) / M
    dx, dy = split_ops(prob)
    vx, vy, sxx, sxy, syy = F[:5]
    ref = [(dx(sxx) + dy(sxy)) / rho, (dx(sxy) + dy(syy)) / rho, M * (dx(vx) + gamma * dy(vy)),
           G * (dx(vy) + dy(vx)), M * (gamma * dx(vx) + dy(vy))]
    SAT = np.array([DF[l] - ref[l] for l in range(5)])
    blk = prob.blocks[0]
    ny, nx = prob.ny, prob.nx
    for s, side in enumerate(blk["sides"]):
        idx = [(slice(None), 0), (slice(None), nx - 1), (0, slice(None)), (ny - 1, slice(None))][s]
        n1, n2 = side["normal"]
        Zs, Zp, gam, Ks, Kp = side["M"]
        pen = to_nt(SAT[(slice(None),) + idx], n1, n2)
        vt, vn, stt, snt, snn = to_nt(F[(slice(None),) + idx], n1, n2)
        dvt, dvn, dstt, dsnt, dsnn = pen[0] / Ks, pen[1] / Kp, pen[2] / Kp, pen[3] / Ks, pen[4] / Kp
        k = slice(1, -1)  # corners carry two penalties
        h = dict(vt=(vt + dvt)[k], vn=(vn + dvn)[k], snt=(snt + dsnt)[k], snn=(snn + dsnn)[k])
        sc = np.abs(np.concatenate([snn, snt, Zp[k].max() * vn, Zs[k].max() * vt])).max()
        tol = 1e-9 * sc
        assert np.abs((h["snt"] - Zs[k] * h["vt"]) - (snt - Zs * vt)[k]).max() < tol, (s, "S out")
        assert np.abs((h["snn"] - Zp[k] * h["vn"]) - (snn - Zp * vn)[k]).max() < tol, (s, "P out")
        assert np.abs(dstt - gam * dsnn)[k].max() < tol, (s, "stt")
        bc = bcs[s]
        if bc == "absorbing":
            assert np.abs(h["snt"] + Zs[k] * h["vt"]).max() < tol, (s, bc)
            assert np.abs(h["snn"] + Zp[k] * h["vn"]).max() < tol, (s, bc)
        elif bc == "free":
            assert np.abs(h["snt"]).max() < tol and np.abs(h["snn"]).max() < tol, (s, bc)
        elif bc == "rigid":
            assert np.abs(h["vt"]).max() < tol / Zs.max() and np.abs(h["vn"]).max() < tol / Zp.max()


@pytest.mark.parametrize("bcs", [("absorbing-velocity", "free-velocity", "free-velocity",
                                  "absorbing-velocity"),
                                 ("free-velocity", "absorbing-velocity", "absorbing-velocity",
                                  "free-velocity")])
def test_velocity_boundary_hat_states(bcs):
    """'absorbing-velocity' and 'free-velocity' (boundaries.f90:333-342)
    keep the grid's traction in the hat state and change only the velocity,
    both relative to the side's initial fields F0: absorbing-velocity takes
    the velocity at which the hat's incoming characteristics vanish,
    snt + Zs vt = snn + Zp vn = 0; free-velocity the velocity a free surface
    gives, the one that keeps the outgoing characteristics at zero traction,
    -Zs vt = snt - Zs vt(grid), -Zp vn = snn - Zp vn(grid).  Neither moves
    stt (snn is unchanged).  The initial fields are far from zero so that F0
    matters."""
    cs, cp, rho = 3.464, 6.0, 2.67
    bg = np.array([0.3, -0.2, -120.0, 70.0, -110.0, -100.0])
    prob, F = problems.build(nx=13, nys=(14,), seed=13, warp=0.07, cs=cs, cp=cp, rho=rho,
                             background=list(bg), perturb=3.0, bcs=bcs)
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    G, M = rho * cs ** 2, rho * cp ** 2
    gamma = (M - 2 * G) / M
    dx, dy = split_ops(prob)
    vx, vy, sxx, sxy, syy = F[:5]
    ref = [(dx(sxx) + dy(sxy)) / rho, (dx(sxy) + dy(syy)) / rho, M * (dx(vx) + gamma * dy(vy)),
           G * (dx(vy) + dy(vx)), M * (gamma * dx(vx) + dy(vy))]
    SAT = np.array([DF[l] - ref[l] for l in range(5)])
    FD = F - bg[:, None, None]
    ny, nx = prob.ny, prob.nx
    blk = prob.blocks[0]
    for s, side in enumerate(blk["sides"]):
        idx = [(slice(None), 0), (slice(None), nx - 1), (0, slice(None)), (ny - 1, slice(None))][s]
        n1, n2 = side["normal"]
        Zs, Zp, gam, Ks, Kp = side["M"]
        pen = to_nt(SAT[(slice(None),) + idx], n1, n2)
        vt, vn, stt, snt, snn = to_nt(FD[(slice(None),) + idx], n1, n2)
        dvt, dvn, dstt, dsnt, dsnn = pen[0] / Ks, pen[1] / Kp, pen[2] / Kp, pen[3] / Ks, pen[4] / Kp
        k = slice(1, -1)  # corners carry two penalties
        sc = np.abs(np.concatenate([snn, snt, Zp * vn, Zs * vt])).max()
        tol = 1e-9 * sc
        vtol = tol / Zs.max()
        # the traction and stt are kept
        assert np.abs(dsnt[k]).max() < tol and np.abs(dsnn[k]).max() < tol, (s, "traction")
        assert np.abs(dstt[k]).max() < tol, (s, "stt")
        vt_h, vn_h = (vt + dvt)[k], (vn + dvn)[k]
        if bcs[s] == "absorbing-velocity":
            assert np.abs(snt[k] + Zs[k] * vt_h).max() < tol, (s, "S in")
            assert np.abs(snn[k] + Zp[k] * vn_h).max() < tol, (s, "P in")
        else:
            assert np.abs(-Zs[k] * vt_h - (snt - Zs * vt)[k]).max() < tol, (s, "S out")
            assert np.abs(-Zp[k] * vn_h - (snn - Zp * vn)[k]).max() < tol, (s, "P out")
        # the velocity changes: the conditions are not vacuous here
        assert np.abs(dvt[k]).max() > 1e3 * vtol and np.abs(dvn[k]).max() > 1e3 * vtol


def test_locked_interface_hat_states():
    """Locked interface (interfaces.f90:715-772): the hat states of the two
    sides have equal velocity and traction, each side keeps its outgoing
    characteristics (snt -/+ Zs vt, snn -/+ Zp vn along the minus side's
    normal) and moves stt by gamma times its change of snn."""
    cs, cp, rho = 3.464, 6.0, 2.67
    prob, F = problems.build(nx=13, nys=(12, 13), seed=21, warp=0.06, cs=cs, cp=cp, rho=rho,
                             background=[0.2, -0.1, -120.0, 70.0, -110.0, -100.0], perturb=3.0,
                             coupling="locked", bcs=("absorbing", "free", "absorbing", "free"))
    # a jump across the interface, so that the hat states differ from both sides
    rng = np.random.default_rng(5)
    F[:, 12:] += rng.uniform(0.5, 2.0, (6, 1, 1)) * np.cos(2.0 * prob.x[12:] + rng.uniform(0, 6, (6, 1, 1)))
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    G, M = rho * cs ** 2, rho * cp ** 2
    gamma = (M - 2 * G) / M
    hat = {}
    for b, (s, j) in enumerate([(3, 11), (2, 12)]):
        blk = prob.blocks[b]
        sj = slice(blk["my"], blk["py"] + 1)
        sub = type("P", (), {})()
        sub.nx, sub.ny = prob.nx, blk["py"] - blk["my"] + 1
        for k in ("xq", "xr", "yq", "yr", "Ji"):
            setattr(sub, k, getattr(prob, k)[sj])
        dx, dy = split_ops(sub)
        vx, vy, sxx, sxy, syy = F[:5, sj]
        ref = [(dx(sxx) + dy(sxy)) / rho, (dx(sxy) + dy(syy)) / rho, M * (dx(vx) + gamma * dy(vy)),
               G * (dx(vy) + dy(vx)), M * (gamma * dx(vx) + dy(vy))]
        r = j - blk["my"]
        SAT = np.array([DF[l, j] - ref[l][r] for l in range(5)])
        n1, n2 = prob.blocks[0]["sides"][3]["normal"]  # minus side's outward normal
        Zs, Zp, gam, Ks, Kp = blk["sides"][s]["M"]
        pen = to_nt(SAT, n1, n2)
        base = to_nt(F[:, j], n1, n2)
        d = [pen[0] / Ks, pen[1] / Kp, pen[2] / Kp, pen[3] / Ks, pen[4] / Kp]
        hat[b] = dict(vt=base[0] + d[0], vn=base[1] + d[1], snt=base[3] + d[3],
                      snn=base[4] + d[4], base=base, Zs=Zs, Zp=Zp, dstt=d[2], dsnn=d[4], gam=gam)
    k = slice(1, -1)
    m, p = hat[0], hat[1]
    sc = np.abs(np.concatenate([m["base"][4], m["base"][3], p["base"][4]])).max()
    tol = 1e-9 * sc
    for f in ("snt", "snn"):
        assert np.abs(m[f] - p[f])[k].max() < tol, f
    for f in ("vt", "vn"):
        assert np.abs(m[f] - p[f])[k].max() * m["Zp"].max() < tol, f
    for h, sgn in ((m, -1.0), (p, 1.0)):
        vt, vn, _, snt, snn = h["base"]
        assert np.abs((h["snt"] + sgn * h["Zs"] * h["vt"]) - (snt + sgn * h["Zs"] * vt))[k].max() < tol
        assert np.abs((h["snn"] + sgn * h["Zp"] * h["vn"]) - (snn + sgn * h["Zp"] * vn))[k].max() < tol
        assert np.abs(h["dstt"] - h["gam"] * h["dsnn"])[k].max() < tol


@pytest.mark.parametrize("seed", [33, 34])
def test_frictional_interface_hat_states_and_law(seed):
    """Frictional interface (interfaces.f90:900-1012 with solve_friction):
    with equal initial fields on both sides, the hat states have equal
    traction and normal velocity, a tangential velocity jump equal to the
    slip velocity V, keep each side's outgoing characteristics, and the
    stored S, N are the hat tractions plus the load; (V, S, N, Psi) satisfy
    the regularised slip law with flash heating (RFL)."""
    cs, cp, rho = 3.464, 6.0, 2.67
    prob, F = problems.build(nx=17, nys=(12, 13), seed=seed, warp=0.05, cs=cs, cp=cp, rho=rho,
                             background=[0.0, 0.0, -120.0, 70.0, -120.0, -120.0], perturb=2.0,
                             law="RFL", fault_overstress=25.0,
                             bcs=("absorbing", "absorbing", "absorbing", "free"))
    rng = np.random.default_rng(seed)
    F[:, 12:] += rng.uniform(0.5, 2.0, (6, 1, 1)) * np.cos(2.0 * prob.x[12:] + rng.uniform(0, 6, (6, 1, 1)))
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    st = {f: o.get_iface(0, f) for f in ("V", "S", "N", "Psi", "O")}
    G, M = rho * cs ** 2, rho * cp ** 2
    gamma = (M - 2 * G) / M
    itf = prob.ifaces[0]
    n1, n2 = itf["nhat"]
    hat = {}
    for b, (s, j) in enumerate([(3, 11), (2, 12)]):
        blk = prob.blocks[b]
        sj = slice(blk["my"], blk["py"] + 1)
        sub = type("P", (), {})()
        sub.nx, sub.ny = prob.nx, blk["py"] - blk["my"] + 1
        for k in ("xq", "xr", "yq", "yr", "Ji"):
            setattr(sub, k, getattr(prob, k)[sj])
        dx, dy = split_ops(sub)
        vx, vy, sxx, sxy, syy = F[:5, sj]
        ref = [(dx(sxx) + dy(sxy)) / rho, (dx(sxy) + dy(syy)) / rho, M * (dx(vx) + gamma * dy(vy)),
               G * (dx(vy) + dy(vx)), M * (gamma * dx(vx) + dy(vy))]
        r = j - blk["my"]
        SAT = np.array([DF[l, j] - ref[l][r] for l in range(5)])
        Zs, Zp, gam, Ks, Kp = blk["sides"][s]["M"]
        pen = to_nt(SAT, n1, n2)
        base = to_nt(F[:, j], n1, n2)
        d = [pen[0] / Ks, pen[1] / Kp, pen[2] / Kp, pen[3] / Ks, pen[4] / Kp]
        hat[b] = dict(vt=base[0] + d[0], vn=base[1] + d[1], snt=base[3] + d[3],
                      snn=base[4] + d[4], base=base, Zs=Zs, Zp=Zp, dstt=d[2], dsnn=d[4], gam=gam)
    k = slice(1, -1)
    m, p = hat[0], hat[1]
    tol = 1e-9 * 120.0
    vtol = tol / m["Zs"].max()
    assert np.abs(m["snt"] - p["snt"])[k].max() < tol
    assert np.abs(m["snn"] - p["snn"])[k].max() < tol
    assert np.abs(m["vn"] - p["vn"])[k].max() < vtol
    assert np.abs((p["vt"] - m["vt"]) - st["V"])[k].max() < vtol
    assert np.abs(st["O"]).max() == 0.0
    for h, sgn in ((m, -1.0), (p, 1.0)):
        vt, vn, _, snt, snn = h["base"]
        assert np.abs((h["snt"] + sgn * h["Zs"] * h["vt"]) - (snt + sgn * h["Zs"] * vt))[k].max() < tol
        assert np.abs((h["snn"] + sgn * h["Zp"] * h["vn"]) - (snn + sgn * h["Zp"] * vn))[k].max() < tol
        assert np.abs(h["dstt"] - h["gam"] * h["dsnn"])[k].max() < tol
    # stored fault tractions are the hat tractions plus the load (ramp = 1 at t = 0 with ld_T = 0)
    S_load = itf["S0base"] + itf["ldS"]
    N_load = itf["N0base"] + itf["ldN"]
    assert np.abs(st["S"] - (m["snt"] + S_load))[k].max() < tol
    assert np.abs(st["N"] - (-m["snn"] - N_load))[k].max() < tol
    a, V0 = itf["rs"][0], itf["rs"][2]
    Sf = a * st["N"] * np.arcsinh(st["V"] / (2 * V0) * np.exp(st["Psi"] / a))
    assert np.abs(Sf - st["S"]).max() < 1e4 * 2.3e-16 * 200
    assert (st["V"][k] > 0).any() and (st["N"] > 0).all()


def test_frictional_interface_takes_initial_fields_crosswise():
    """The reference's call at interfaces.f90:795 passes the minus side's
    initial fields F0m where frictional_interface_mode2 (interfaces.f90:900)
    declares its plus side's, F0p, and the plus side's where it declares
    F0m.  With equal prestress on both sides nothing shows.  Here each block
    rests at its own uniform prestress, the minus block's shear stress 10 MPa
    below the plus block's, on a flat fault.  Then, in closed form:

    - the fault is loaded by the minus side's resolved prestress S0 = 60:
      a N asinh(V e^(Psi/a) / 2V0) = 60 - eta V, with eta = Zs/2;
    - each side's hat state is completed with the other side's initial
      fields, so the penalties move the minus side's snt by +10 - eta V and
      vt by (10 - eta V)/Zs, the plus side's snt by -10 - eta V and vt by
      (10 + eta V)/Zs; normal stress and normal velocity stay.

    Reading the arguments in their declared order would load the fault with
    70 (a slip velocity about e^8 times larger) and leave every penalty at
    about eta V."""
    from scipy.optimize import brentq

    cs, cp, rho = 3.464, 6.0, 2.67
    bg_p = np.array([0.0, 0.0, -120.0, 70.0, -120.0, -120.0])
    bg_m = bg_p.copy()
    bg_m[3] = 60.0
    prob, F = problems.build(nx=17, nys=(12, 13), seed=35, warp=0.0, cs=cs, cp=cp, rho=rho,
                             background=list(bg_p), perturb=0.0, law="RFL", fault_overstress=0.0,
                             bcs=("absorbing",) * 4)
    F[:, :12] = bg_m[:, None, None]
    for side in prob.blocks[0]["sides"]:
        side["F0"] = np.tile(bg_m[:, None], (1, side["F0"].shape[1]))
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    _, DF = o.get_fields()
    V, S, S0, N, Psi = (o.get_iface(0, f) for f in ("V", "S", "S0", "N", "Psi"))
    itf = prob.ifaces[0]
    a, V0 = itf["rs"][0], itf["rs"][2]
    Zs = rho * cs
    eta = 0.5 * Zs
    # the load is the minus side's prestress
    assert np.all(S0 == 60.0) and np.all(N == 120.0)
    want = np.array([brentq(lambda v: a[k] * N[k] * np.arcsinh(v / (2 * V0[k]) * np.exp(Psi[k] / a[k]))
                            - (60.0 - eta * v), 0.0, 1.0, xtol=1e-300, rtol=1e-15)
                     for k in range(len(V))])
    np.testing.assert_allclose(V, want, rtol=1e-10)
    np.testing.assert_allclose(S, 60.0 - eta * V, rtol=1e-14)
    other = brentq(lambda v: a[0] * 120.0 * np.arcsinh(v / (2 * V0[0]) * np.exp(Psi[0] / a[0]))
                   - (70.0 - eta * v), 0.0, 1.0, xtol=1e-300, rtol=1e-15)
    assert other > 1e3 * V[0]
    # penalties of the stage-1 rates at the fault rows (uniform fields: the
    # operator adds nothing); flat fault, normal (0, 1): t = x, n = y
    n1, n2 = itf["nhat"]
    assert np.allclose(n1, 0.0) and np.allclose(n2, 1.0)
    k = slice(1, -1)
    for b, j, s, dsnt, dvt in ((0, 11, 3, 10.0 - eta * V, (10.0 - eta * V) / Zs),
                              (1, 12, 2, -10.0 - eta * V, (10.0 + eta * V) / Zs)):
        Zs_, Zp_, gam, Ks, Kp = prob.blocks[b]["sides"][s]["M"]
        pen = to_nt(DF[:5, j], n1, n2)
        d = [pen[0] / Ks, pen[1] / Kp, pen[2] / Kp, pen[3] / Ks, pen[4] / Kp]
        np.testing.assert_allclose(d[3][k], dsnt[k], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(d[0][k], dvt[k], rtol=1e-12, atol=1e-15)
        assert np.abs(d[1][k]).max() < 1e-12 and np.abs(d[4][k]).max() < 1e-12


def _fault_problem(seed, **kw):
    prob, F = problems.build(nx=17, nys=(12, 13), seed=seed, warp=0.05, cs=3.464, cp=6.0, rho=2.67,
                             background=[0.0, 0.0, -120.0, 70.0, -120.0, -120.0], perturb=2.0,
                             law="RFL", bcs=("absorbing", "absorbing", "absorbing", "free"), **kw)
    rng = np.random.default_rng(seed)
    F[:, 12:] += rng.uniform(0.5, 2.0, (6, 1, 1)) * np.cos(2.0 * prob.x[12:] + rng.uniform(0, 6, (6, 1, 1)))
    return prob, F


def test_lock_region_is_coulomb_friction():
    """Outside the lock region (friction.f90:650-672) the fault obeys
    Coulomb friction with coefficient flock: |S| <= flock N, slipping only at
    |S| = flock N, in the direction of S."""
    prob, F = _fault_problem(41, outside_frac=0.3, fault_overstress=60.0)
    out = prob.ifaces[0]["outside"].astype(bool)
    flock = prob.ifaces[0]["flock"]
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    V, S, N = (o.get_iface(0, f)[out] for f in ("V", "S", "N"))
    assert (np.abs(S) <= flock * N * (1 + 1e-14)).all()
    slip = V != 0.0
    assert slip.any() and (~slip).any()
    assert np.allclose(np.abs(S[slip]), flock * N[slip], rtol=1e-14, atol=0)
    assert (np.sign(S[slip]) == np.sign(V[slip])).all()


@pytest.mark.parametrize("T,t", [(0.01, 0.004), (0.05, 0.02)])
def test_load_ramp(T, t):
    """load_stress with ramp time T (friction.f90:1339-1374, 1455-1484): the
    stored shear traction is the hat traction plus S0base + ldS (1 - e^(-t/T))."""
    prob, F = _fault_problem(43, ld_T=T)
    itf = prob.ifaces[0]
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    dt = 1e-3
    n = int(round(t / dt)) + 1
    o.stage1(0.0, dt, n)
    S, N = o.get_iface(0, "S"), o.get_iface(0, "N")
    S0, N0 = o.get_iface(0, "S0"), o.get_iface(0, "N0")
    n1, n2 = itf["nhat"]
    sxx, sxy, syy = -120.0, 70.0, -120.0
    snt = n1 * n2 * (sxx - syy) + (n2 * n2 - n1 * n1) * sxy
    ramp = 1.0 - np.exp(-t / T)
    # S0, N0: the resolved initial traction of the plus side plus the ramped load
    assert np.allclose(S0, snt + itf["S0base"] + itf["ldS"] * ramp, rtol=1e-13, atol=1e-12)
    assert (N > 0).all() and np.isfinite(S).all()


def test_fault_state_rates_over_one_small_step():
    """set_rates_friction (friction.f90:1101-1124) and the fault update: over
    one short step D grows by the time integral of |V|, W by that of S V and
    Psi by that of its state rate (trapezoid rule, second order in dt)."""
    prob, F = _fault_problem(47)
    dt = 1e-6
    rs = prob.ifaces[0]["rs"]
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    st = []
    for n in (1, 2):
        o.stage1(0.0, dt, n)
        s = {f: o.get_iface(0, f).copy() for f in ("V", "S", "Psi", "D", "W")}
        s["rate"] = np.array([ob.state_rate("RFL", rs[:, k], s["V"][k], s["Psi"][k])
                              for k in range(len(s["V"]))])
        st.append(s)
        if n == 1:
            o.step(0.0, dt, 1, 1)
    a, b = st
    k = slice(1, -1)
    trap = lambda f: 0.5 * (f(a) + f(b))[k] * dt
    assert np.allclose((b["D"] - a["D"])[k], trap(lambda s: np.abs(s["V"])), rtol=1e-5, atol=1e-20)
    assert np.allclose((b["W"] - a["W"])[k], trap(lambda s: s["S"] * s["V"]), rtol=1e-5, atol=1e-20)
    # increments below a few ulps of Psi (about 0.7) are lost to rounding
    assert np.allclose((b["Psi"] - a["Psi"])[k], trap(lambda s: s["rate"]), rtol=1e-5, atol=5e-16)


@pytest.mark.parametrize("stf", ["gaussian", "Ricker", "ramp", "sinusoid"])
def test_point_source_moment_and_time_function(stf):
    """The singular source (source.f90:196-254) adds A(t) M to the stress
    rates through a discrete delta whose weights sum to 1/h in x and in y, so
    sum(DF) h^2 = -A(t) M on a zero field, and A(t) has the defining
    properties of each time function."""
    h = 0.025
    prob, F = problems.build(nx=21, nys=(19,), seed=2, warp=0.0, h=h, background=[0.0] * 6,
                             perturb=0.0, source=True, bcs=("absorbing",) * 4)
    tc, td = 0.02, 0.01
    prob.source.update(stf=stf, t_center=tc, t_duration=td, Fx=0.3, Fy=-0.2, Mxx=1.0, Mxy=0.5,
                       Myy=-0.7)
    rho = prob.blocks[0]["rho"]

    def amp(t):
        o = OracleContext(prob)
        o.set_fields(F)
        o.init_edges(0.0)
        o.stage1(t, 1.0, 1)
        DF = o.get_fields()[1]
        mom = DF.sum(axis=(1, 2)) * h * h
        A = -mom[2] / 1.0
        assert np.allclose(mom[:5], [A * 0.3 / rho, -A * 0.2 / rho, -A, -0.5 * A, 0.7 * A],
                           rtol=1e-12, atol=1e-15)
        return A

    if stf == "gaussian":
        assert abs(amp(tc) - 1.0) < 1e-14
        assert abs(amp(tc + td) - np.exp(-0.5)) < 1e-14 and abs(amp(tc - td) - np.exp(-0.5)) < 1e-14
    elif stf == "Ricker":
        assert abs(amp(tc) - 1.0) < 1e-14
        z = td / (np.pi * np.sqrt(2.0))
        assert abs(amp(tc + z)) < 1e-14 and abs(amp(tc - z)) < 1e-14
        assert amp(tc + 2 * z) < 0
    elif stf == "ramp":
        assert amp(tc - 0.1 * td) == 0.0 and amp(tc + 2 * td) == 1.0
        assert abs(amp(tc + 0.25 * td) - 0.25) < 1e-14
    else:
        assert abs(amp(tc)) < 1e-14 and abs(amp(tc + 0.25 * td) - 1.0) < 1e-14
        assert abs(amp(tc + 0.5 * td)) < 1e-14
