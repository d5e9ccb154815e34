"""Perfectly matched layer of the oracle (rates_interior_mode2_pml,
rates_boundary_mode2_pml, rates.f90:2024-2619; pml_type, material.f90:9-15,
152-223), pinned by what the mathematics fixes:

- with no damping a layer block is the plain SBP operator on a Cartesian
  grid (the curvilinear split form reduces to it there);
- the auxiliary fields enter the rates through the elastic operator's own
  x and y coefficient matrices (read off the non-layer operator applied to
  linear fields, not retyped) and decay at sigma + kappa;
- the auxiliary rates are sigma(x) times the field derivatives, which SBP
  differentiates exactly for linear fields, closures included;
- a graded layer in front of a reflecting side absorbs what the side would
  send back.
"""
import numpy as np
import pytest

import problems
from oracle.binding import OracleContext

BIG = 1.7976931348623157e308
RHO, CS, CP = 2.67, 3.464, 6.0


def _cart(pml, background=(0.1, -0.2, -120.0, 70.0, -110.0, -100.0), seed=1, nx=17, nys=(19,), **kw):
    return problems.build(nx=nx, nys=nys, seed=seed, warp=0.0, background=list(background), perturb=0.5,
                          bcs=("absorbing", "free", "absorbing", "free"), pml=pml, cs=CS, cp=CP, rho=RHO,
                          **kw)


def _stage1(prob, F, W=None):
    o = OracleContext(prob)
    o.set_fields(F)
    if W is not None:
        o.set_pml(*W)
    o.init_edges(0.0)
    o.stage1(0.0, 1e-3, 1)
    return o, o.get_fields()[1]


def test_layer_without_damping_is_the_cartesian_operator():
    prob, F = _cart(dict(x=True, y=True, sigma=0.0, kappa=0.0))
    plain, _ = _cart(None)
    o, DF = _stage1(prob, F)
    _, ref = _stage1(plain, F)
    scale = np.abs(ref).max(axis=(1, 2), keepdims=True)
    err = np.abs(DF - ref) / scale
    assert err.max() < 1e-12, err.max()
    DWx, DWy = o.get_pml_rates()
    assert not DWx.any() and not DWy.any()


def _coefficient_matrix(axis):
    """columns: the non-layer operator's rates of the field that is the
    coordinate (x or y) in component k, at an interior point"""
    plain, _ = _cart(None, background=(0.0,) * 6)
    A = np.zeros((5, 6))
    for k in range(6):
        F = np.zeros((6, plain.ny, plain.nx))
        F[k] = plain.x if axis == 0 else plain.y
        _, DF = _stage1(plain, F)
        A[:, k] = DF[:5, plain.ny // 2, plain.nx // 2]
    return A


def test_auxiliary_fields_couple_through_the_elastic_operator_and_decay():
    Ax, Ay = _coefficient_matrix(0), _coefficient_matrix(1)
    # a few entries fixed by the equations of motion and Hooke's law
    G, M = RHO * CS ** 2, RHO * CP ** 2
    np.testing.assert_allclose([Ax[0, 2], Ay[0, 3], Ax[2, 0], Ay[4, 1], Ax[3, 1]],
                               [1 / RHO, 1 / RHO, M, M, G], rtol=1e-12)
    sigma, kappa = 2.5, 0.3
    prob, _ = _cart(dict(x=True, y=True, sigma=sigma, kappa=kappa), background=(0.0,) * 6)
    rng = np.random.default_rng(7)
    Wx = rng.standard_normal((6, prob.ny, prob.nx))
    Wy = rng.standard_normal((6, prob.ny, prob.nx))
    o, DF = _stage1(prob, np.zeros((6, prob.ny, prob.nx)), (Wx, Wy))
    expect = -(np.einsum("lk,kji->lji", Ax, Wx) + np.einsum("lk,kji->lji", Ay, Wy))
    err = np.abs(DF[:5] - expect).max() / np.abs(expect).max()
    assert err < 1e-12, err
    DWx, DWy = o.get_pml_rates()
    np.testing.assert_array_equal(DWx, -(sigma + kappa) * Wx)
    np.testing.assert_array_equal(DWy, -(sigma + kappa) * Wy)


@pytest.mark.parametrize("n", [2.0, 3.0])
def test_auxiliary_rates_are_the_graded_damping_times_the_derivatives(n):
    probe, _ = _cart(None)
    xs, ys = probe.x, probe.y
    px = xs.min() + 0.6 * (xs.max() - xs.min())
    my = ys.min() + 0.5 * (ys.max() - ys.min())
    sigma = 40.0
    layer = dict(x=True, y=True, n=n, sigma=sigma, kappa=0.0, mx=-BIG, px=px, my=my, py=BIG)
    prob, _ = _cart(layer)
    rng = np.random.default_rng(11)
    cx, cy = rng.uniform(-2, 2, 6), rng.uniform(-2, 2, 6)
    F = cx[:, None, None] * prob.x + cy[:, None, None] * prob.y
    o, _ = _stage1(prob, F)
    DWx, DWy = o.get_pml_rates()
    sx = sigma * np.maximum(0.0, prob.x - px) ** n
    sy = sigma * np.maximum(0.0, my - prob.y) ** n
    assert (sx > 0).any() and (sx == 0).any() and (sy > 0).any() and (sy == 0).any()
    ex = cx[:, None, None] * sx
    ey = cy[:, None, None] * sy
    assert np.abs(DWx - ex).max() < 1e-11 * np.abs(ex).max()
    assert np.abs(DWy - ey).max() < 1e-11 * np.abs(ey).max()


def _energy(prob, F, sl):
    G, M = RHO * CS ** 2, RHO * CP ** 2
    vx, vy, sxx, sxy, syy = (f[sl] for f in F[:5])
    lam = M - 2 * G
    # plane-strain strain energy from the in-plane stresses
    exx = ((lam + 2 * G) * sxx - lam * syy) / (4 * G * (lam + G))
    eyy = ((lam + 2 * G) * syy - lam * sxx) / (4 * G * (lam + G))
    return float((0.5 * RHO * (vx ** 2 + vy ** 2) + 0.5 * (sxx * exx + syy * eyy) + sxy ** 2 / (2 * G)).sum())


def _pulse_run(with_layer, far_side):
    """a right-going P pulse in the left of two blocks; the right block is
    a graded layer (or plain), its far side free or absorbing.  Returns
    the left block's energy (relative) once the reflection is back in it."""
    h, nx = 0.025, 41
    L = (nx - 1) * h
    probe, _ = problems.build(nx=nx, nys=(31,), nbx=2, warp=0.0, h=h, coupling="locked")
    x_if = probe.x[0, nx]
    sigma0 = 1.5 * CP / L * np.log(1e3)
    layer = dict(x=True, y=False, n=2.0, sigma=sigma0 / L ** 2, kappa=0.0, mx=-BIG, px=x_if)
    prob, _ = problems.build(nx=nx, nys=(31,), nbx=2, warp=0.0, h=h, coupling="locked", cs=CS, cp=CP,
                             rho=RHO, background=[0.0] * 6, bcs=("absorbing", far_side, "absorbing", "absorbing"),
                             pml=[None, layer] if with_layer else None)
    xc = prob.x[0, nx // 2]
    yc = 0.5 * (prob.y.min() + prob.y.max())
    g = np.exp(-((prob.x - xc) ** 2 + (prob.y - yc) ** 2) / (2 * 0.08 ** 2))
    F = np.zeros((6, prob.ny, prob.nx))
    # right-going P pulse: vx = -sxx/(rho cp)
    F[2] = g
    F[0] = -g / (RHO * CP)
    dt = 0.5 * h / np.hypot(CS, CP)
    o = OracleContext(prob)
    o.set_fields(F)
    o.init_edges(0.0)
    left = (slice(None), slice(0, nx))
    E0 = _energy(prob, F, left)
    # out 0.5 L, across the right block and back: 2.5 L at cp, and a margin
    o.step(0.0, dt, 1, int(round(1.2 * 2.5 * L / CP / dt)))
    Ff = o.get_fields()[0]
    assert np.isfinite(Ff).all()
    return _energy(prob, Ff, left) / E0, o


def test_graded_layer_absorbs_what_a_free_side_reflects():
    """the energy a free far side sends back into the left block, beyond
    what stays there with an absorbing far side, mostly vanishes behind a
    graded layer (the left block keeps a share of the initial stress field
    in every run, which the differences remove)"""
    E_abs, _ = _pulse_run(False, "absorbing")
    E_free, _ = _pulse_run(False, "free")
    E_pml, o = _pulse_run(True, "free")
    returned = E_free - E_abs
    assert returned > 0.02
    assert abs(E_pml - E_abs) < 0.05 * returned, (E_pml - E_abs) / returned
    Wx, _ = o.get_pml(x=True, y=False)
    assert np.abs(Wx[:, :, :41]).max() == 0.0 and np.abs(Wx).max() > 0.0


def test_dissipation_block_ignores_its_layer():
    """set_rates_all (time_step.f90:241-246): a block with dissipation takes
    set_rates_dissipation, whatever its pml_list says"""
    prob, F = _cart(dict(x=True, y=True, sigma=5.0, kappa=1.0), Cdiss=0.3)
    plain, _ = _cart(None, Cdiss=0.3)
    _, DF = _stage1(prob, F)
    _, ref = _stage1(plain, F)
    np.testing.assert_array_equal(DF, ref)


def test_layer_refused_in_antiplane_shear():
    prob, F = problems.build(nx=12, nys=(12,), mode=3, background=[0.0] * 3, pml=dict(x=True))
    with pytest.raises(RuntimeError, match="PML"):
        OracleContext(prob)
