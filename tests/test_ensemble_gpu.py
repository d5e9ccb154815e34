"""Ensembles: every RK stage of all members is one batched launch per kernel
kind, and a step is one captured hipGraph launch (or the same launches made
directly).  Replayed in several chunks, with the program's own eager work and
synchronisation between the chunks, the ensemble must give the same bits as
stepping each member on its own (fdm_step), in FP64 and FP32-bulk, whatever
tile shape each member runs, with plastic members too."""
import numpy as np
import pytest

from fdmap_github_amd import ensemble, host
from fdmap_github_amd.device import Context, FdmError

pytestmark = pytest.mark.gpu


def build_members(tmp_path, precision, count=3):
    scen = ensemble.c4_scenarios(count=count, seed=5, nx_range=(72, 140), ny_range=(40, 70))
    pairs, times = [], []
    for sc in scen:
        path = ensemble.write_scenario(str(tmp_path), sc, 100)
        hp = host.HostProblem(path, precision=precision)
        eager, member = Context(hp.ptr), Context(hp.ptr)
        F = hp.fields()
        for c in (eager, member):
            c.set_fields(F)
            c.init_edges(hp.t)
        pairs.append((eager, member))
        times.append((hp.t, hp.dt))
    return pairs, np.array([t for t, _ in times]), np.array([d for _, d in times])


def snapshot(c):
    F, DF = c.get_fields()
    return F, DF, {f: c.get_iface(0, f) for f in ("D", "V", "S", "N", "Psi", "trup")}


@pytest.mark.parametrize("graph", ["1", "0"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_match_eager_steps(gpu, tmp_path, monkeypatch, precision, graph):
    monkeypatch.setenv("FDM_ENSEMBLE_GRAPH", graph)
    # the graph cases also run the non-temporal interior stores (members and
    # their eager twins alike)
    monkeypatch.setenv("FDM_NT_STORE", graph)
    pairs, t, dt = build_members(tmp_path, precision)
    for k, (_, member) in enumerate(pairs):
        ensemble.apply_tile(member, k % 3)
    ens = ensemble.Ensemble([m for _, m in pairs])
    n = 1
    for chunk in (3, 4, 5):
        print("precision", precision, "graph", graph, "chunk", chunk, flush=True)
        ens.step(t, dt, n, chunk)
        for k, (eager, _) in enumerate(pairs):
            eager.step(t[k], dt[k], n, chunk)
        n += chunk
        for eager, member in pairs:
            a, b = snapshot(eager), snapshot(member)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            for f in a[2]:
                assert np.array_equal(a[2][f], b[2][f]), f
    assert max(np.abs(snapshot(m)[2]["V"]).max() for _, m in pairs) > 0.0
    ens.close()


def test_members_with_point_sources_are_refused(gpu):
    import problems

    prob, F = problems.build(nx=40, nys=(20, 20), seed=1, source=True,
                             background=[0, 0, -120.0, 70.0, -120.0, -120.0])
    c = Context(prob)
    with pytest.raises(FdmError, match="point source"):
        ensemble.Ensemble([c])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_plastic_members_match_eager_steps(gpu, precision):
    """Members with Drucker-Prager plasticity (the last stage runs the full
    finishing variant for every member of the batch) next to an elastic one."""
    bg = [0.0, 0.0, -60.0, 44.0, -100.0, -80.0]
    plastic = dict(mu=0.6, beta=0.0, b=0.0, h=0.0, eta=0.0, s0=[0.0] * 6)
    specs = [dict(nx=150, nys=(44, 52), seed=41, warp=0.02, background=bg, law="RFL",
                  fault_overstress=20.0, plastic=plastic),
             dict(nx=97, nys=(40, 37), seed=43, warp=0.03, background=bg, law="RFL"),
             dict(nx=131, nys=(61, 45), seed=47, warp=0.01, background=bg, law="RFL",
                  fault_overstress=15.0, plastic=plastic)]
    import problems

    pairs = []
    for spec in specs:
        prob, F = problems.build(precision=precision, **spec)
        eager, member = Context(prob), Context(prob)
        for c in (eager, member):
            c.set_fields(F)
            c.init_edges(0.0)
        pairs.append((eager, member))
    dt = 0.5 * 0.025 / np.sqrt(3.464 ** 2 + 6.0 ** 2)
    t = np.zeros(len(pairs))
    dts = np.full(len(pairs), dt)
    ens = ensemble.Ensemble([m for _, m in pairs])
    ens.step(t, dts, 1, 7)
    for eager, _ in pairs:
        eager.step(0.0, dt, 1, 7)
    for k, (eager, member) in enumerate(pairs):
        print("member", k, precision, flush=True)
        a, b = snapshot(eager), snapshot(member)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        for f in a[2]:
            assert np.array_equal(a[2][f], b[2][f]), f
        pa, pb = eager.get_plastic(), member.get_plastic()
        assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])
    assert pairs[0][1].get_plastic()[0].max() > 0.0
    ens.close()


def test_members_of_mixed_precision_are_refused(gpu):
    import problems

    a = Context(problems.build(nx=40, nys=(20, 20), seed=1, precision="fp64")[0])
    b = Context(problems.build(nx=40, nys=(20, 20), seed=1, precision="fp32")[0])
    with pytest.raises(FdmError, match="precision"):
        ensemble.Ensemble([a, b])
    c = Context(problems.build(nx=40, nys=(20, 20), seed=1, medium=0.2)[0])
    with pytest.raises(FdmError, match="heterogeneous"):
        ensemble.Ensemble([a, c])
