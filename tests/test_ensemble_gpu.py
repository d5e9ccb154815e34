"""Ensembles: every RK step of all members is one captured hipGraph launch.
Replayed in several chunks, with the program's own eager work and
synchronisation between the chunks, the graphs must give the same bits as
stepping each member on its own (fdm_step), in FP64 and FP32-bulk, whatever
tile shape each member runs."""
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


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_match_eager_steps(gpu, tmp_path, precision):
    pairs, t, dt = build_members(tmp_path, precision)
    for k, (_, member) in enumerate(pairs):
        ensemble.apply_tile(member, k % 3)
    ens = ensemble.Ensemble([m for _, m in pairs])
    n = 1
    for chunk in (3, 4, 5):
        print("precision", precision, "chunk", chunk, flush=True)
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
