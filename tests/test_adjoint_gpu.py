"""The discrete adjoint (include/fdmap_adjoint.h) against central finite
differences of the forward path itself.

- Locked, curvilinear, four blocks: the receiver traces are linear in the
  initial fields, so the misfit is quadratic and its central difference is
  exact up to rounding; the gradient with respect to the initial fields
  must match it to 1e-9.  This pins the transposed SBP4 operator with its
  closures, the SAT penalties, the free and absorbing boundary hats, the
  locked-interface hats and the RK recurrences.
- Rate-and-state friction (RFL, RSL, RAL): directional derivatives with
  respect to the fault shear prestress, a, b and the initial state, central
  differences with eps = 1e-6, relative error <= 1e-6; the best of the
  steps 1e-7..1e-4 within 2e-7 (the forward path's rounding noise, which
  the unstable rupture amplifies, dominates the difference quotient at
  small steps).
- Checkpoint intervals 1, 7 (a partial last interval), sqrt(N), N and the
  automatic choice give identical gradients; the
  torch.autograd.Function gives the gradient of the direct call.
"""
import numpy as np
import pytest

import problems
from fdmap_github_amd.adjoint import FaultInversion, receiver_traces
from fdmap_github_amd.device import Context, FdmError

pytestmark = pytest.mark.gpu

CS, CP, H = 3.464, 6.0, 0.025
DT = 0.5 * H / np.sqrt(CS ** 2 + CP ** 2)
BG = [0.0, 0.0, -120.0, 70.0, -120.0, -120.0]


def top_receivers(prob, nrec):
    i = np.linspace(3, prob.nx - 4, nrec).round().astype(np.int32)
    return i, np.full(nrec, prob.ny - 1, dtype=np.int32)


def misfit(rec, target):
    return 0.5 * DT * float(np.sum((rec - target) ** 2))


def setup(precision="fp64", nsteps=120, interval=0, nrec=8, **kw):
    prob, F = problems.build(precision=precision, **kw)
    ctx = Context(prob)
    ctx.set_fields(F)
    ctx.init_edges(0.0)
    ri, rj = top_receivers(prob, nrec)
    return prob, F, ctx, FaultInversion(ctx, 0.0, DT, ri, rj, nsteps, interval)


def test_initial_field_gradient_of_linear_problem(gpu):
    prob, F, ctx, inv = setup(nx=40, nys=(24, 28), nbx=2, seed=3, warp=0.05, coupling="locked",
                              background=BG, perturb=1e-2, nsteps=90)
    rng = np.random.default_rng(4)
    target = inv.run() + 1e-3 * rng.standard_normal((90, 8, 2))
    rec = inv.run()
    g = inv.adj.backward(DT * (rec - target), fields=True)["F0"]
    assert np.all(g[5] == 0.0)
    for trial in range(2):
        dF = rng.standard_normal(F.shape) * np.array([1e-2, 1e-2, 1e-1, 1e-1, 1e-1, 1e-1])[:, None, None]
        eps = 1e-2
        fd = (misfit(inv.run(F0=F + eps * dF), target) - misfit(inv.run(F0=F - eps * dF), target)) / (2 * eps)
        ad = float(np.sum(g * dF))
        print("linear trial", trial, "fd", fd, "adjoint", ad, flush=True)
        assert abs(fd - ad) <= 1e-9 * abs(fd)


def friction_case(law, seed):
    prob, F, ctx, inv = setup(nx=64, nys=(30, 34), seed=seed, warp=0.02, background=BG,
                              perturb=1e-3, law=law, nsteps=150)
    itf = prob.ifaces[0]
    p0 = {"S0": itf["S0base"].copy(), "a": itf["rs"][0].copy(), "b": itf["rs"][1].copy()}
    rng = np.random.default_rng(seed + 100)
    n = len(p0["S0"])
    bump = np.convolve(rng.standard_normal(n + 8), np.ones(9) / 9.0, mode="valid")
    target = inv.run(S0=p0["S0"] + 2.0 * bump / np.abs(bump).max(), a=p0["a"], b=p0["b"])
    return prob, ctx, inv, p0, target, rng


def check_direction(label, J, ad, rng_eps=(1e-7, 1e-6, 1e-5, 1e-4)):
    """Central differences of J along one direction: at eps = 1e-6 within
    1e-6 of the adjoint, and at the best step (rounding noise of the forward
    path grows as 1/eps, truncation as eps^2) within 2e-7."""
    rel = {}
    for eps in rng_eps:
        fd = (J(eps) - J(-eps)) / (2 * eps)
        rel[eps] = abs(fd - ad) / abs(ad)
        print(label, "eps %.0e fd % .12e adjoint % .12e rel %.2e" % (eps, fd, ad, rel[eps]), flush=True)
    assert rel[1e-6] <= 1e-6, rel
    assert min(rel.values()) <= 2e-7, rel


@pytest.mark.parametrize("law,seed", [("RFL", 11), ("RSL", 12), ("RAL", 13)])
def test_friction_gradients_match_central_differences(gpu, law, seed):
    """Directions with N(0, 1) entries times a scale per parameter: 10 MPa
    for the shear prestress (70 MPa on the fault), 10 for b, 1 for the
    initial state, a itself for a (the misfit is strongly nonlinear in a:
    an absolute step of 1e-6 is already in the eps^2 truncation regime
    there)."""
    prob, ctx, inv, p0, target, rng = friction_case(law, seed)
    rec = inv.run(**p0)
    vmax = np.abs(ctx.get_iface(0, "V")).max()
    g = inv.adj.backward(DT * (rec - target))
    print(law, "misfit", misfit(rec, target), "max V", vmax, flush=True)
    assert vmax > 1e-3  # the rupture is under way
    n = len(p0["S0"])
    for name in ("S0", "a", "b"):
        d = rng.standard_normal(n) * {"S0": 10.0, "a": p0["a"], "b": 10.0}[name]

        def J(e, name=name, d=d):
            p = dict(p0)
            p[name] = p0[name] + e * d
            return misfit(inv.run(**p), target)

        check_direction(f"{law} {name}", J, float(g[name] @ d))
    # initial state variable
    psi0 = inv.iface0[0]["Psi"].copy()
    d = rng.standard_normal(n)

    def Jpsi(e):
        inv.iface0[0]["Psi"] = psi0 + e * d
        try:
            return misfit(inv.run(**p0), target)
        finally:
            inv.iface0[0]["Psi"] = psi0

    check_direction(f"{law} Psi0", Jpsi, float(g["Psi0"] @ d))


def test_checkpoint_intervals_and_autograd_agree(gpu):
    import torch

    prob, ctx, inv, p0, target, rng = friction_case("RFL", 21)
    grads = []
    ri, rj = top_receivers(prob, 8)
    for interval, expect in ((1, 1), (-1, 13), (150, 150), (0, 150), (7, 7)):
        other = FaultInversion(ctx, 0.0, DT, ri, rj, 150, interval)
        other.F0, other.DF0, other.iface0 = inv.F0, inv.DF0, inv.iface0
        print("interval", interval, "->", other.adj.interval, flush=True)
        assert other.adj.interval == expect
        rec = other.run(**p0)
        grads.append(other.adj.backward(DT * (rec - target)))
        t = other.adj.timing()
        assert t["forward_ms"] > 0 and t["sweep_ms"] > 0
        if other.adj.interval < 150:
            assert t["recompute_ms"] > 0
        other.adj.close()
    for g in grads[1:]:
        for k in ("S0", "a", "b", "Psi0"):
            assert np.array_equal(g[k], grads[0][k]), k
    S0 = torch.tensor(p0["S0"], requires_grad=True)
    a = torch.tensor(p0["a"], requires_grad=True)
    b = torch.tensor(p0["b"], requires_grad=True)
    rec = receiver_traces(inv, S0, a, b)
    J = 0.5 * DT * ((rec - torch.from_numpy(target)) ** 2).sum()
    J.backward()
    for name, x in (("S0", S0), ("a", a), ("b", b)):
        ref = grads[0][name]
        assert np.allclose(x.grad.numpy(), ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_unsupported_problems_are_refused(gpu):
    prob, F = problems.build(nx=40, nys=(20, 20), seed=1, precision="fp32")
    with pytest.raises(FdmError, match="FP64"):
        FaultInversion(Context(prob), 0.0, DT, [3], [39], 10)
    prob, F = problems.build(nx=40, nys=(20, 20), seed=1, law="SW")
    with pytest.raises(FdmError, match="RSL, RFL"):
        FaultInversion(Context(prob), 0.0, DT, [3], [39], 10)
    prob, F = problems.build(nx=40, nys=(20, 20), seed=1, medium=0.2)
    with pytest.raises(FdmError, match="heterogeneous"):
        FaultInversion(Context(prob), 0.0, DT, [3], [39], 10)
    prob, F = problems.build(nx=40, nys=(20, 20), seed=1)
    with pytest.raises(FdmError, match="receiver"):
        FaultInversion(Context(prob), 0.0, DT, [3], [40], 10)


def test_tiled_and_pointwise_bulk_adjoints_agree(gpu, monkeypatch):
    """The interior tile kernel (central stencil, LDS) and the per-point
    kernel (SBP matrix entries) sum the same terms in different orders."""
    out = []
    for mode, nt in (("tile", "0"), ("generic", "0"), ("tile", "1")):
        monkeypatch.setenv("FDM_ADJ_BULK", mode)
        monkeypatch.setenv("FDM_NT_STORE", nt)
        prob, F, ctx, inv = setup(nx=90, nys=(40, 46), nbx=1, seed=7, warp=0.04, background=BG,
                                  perturb=1e-3, law="RFL", nsteps=60)
        rng = np.random.default_rng(8)
        target = inv.run() + 1e-4 * rng.standard_normal((60, 8, 2))
        rec = inv.run()
        out.append(inv.adj.backward(DT * (rec - target), fields=True))
        inv.adj.close()
        ctx.close()
    for k in ("S0", "a", "b", "Psi0", "F0"):
        a, b = out[0][k], out[1][k]
        err = np.abs(a - b).max() / np.abs(b).max()
        print(k, "max rel diff", err, flush=True)
        assert err <= 1e-12, k
        assert np.array_equal(out[2][k], a), k  # non-temporal stores: same bits
