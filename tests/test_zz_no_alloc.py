"""fdm_max_slip_velocity_noalloc (the entry point that allocates nothing)
against fdm_max_slip_velocity through the C ABI: the same value bit for bit,
on two calls in a row, the second on a smaller problem, each also equal to
the maximum of the slip velocity the context returns; beyond its node limit
it refuses with an error."""
import numpy as np
import pytest

import problems
from fdmap_github_amd.device import Context, FdmError

pytestmark = pytest.mark.gpu

BG = [0.0, 0.0, -120.0, 70.0, -120.0, -120.0]
DT = 0.5 * 0.025 / np.sqrt(3.464 ** 2 + 6.0 ** 2)


def stepped(nx, nys, nbx, seed, nsteps):
    prob, F = problems.build(nx=nx, nys=nys, nbx=nbx, seed=seed, warp=0.02, background=BG,
                             perturb=1e-2, law="RFL")
    c = Context(prob)
    c.set_fields(F)
    c.init_edges(0.0)
    c.step(0.0, DT, 1, nsteps)
    return c


def test_noalloc_max_slip_velocity_matches_the_allocating_call(gpu):
    # a fault of 700 nodes (three workgroups of partial maxima), then a
    # smaller problem of 40 nodes
    for nx, nys, nbx, seed, nsteps in ((700, (30, 34), 1, 51, 30), (40, (20, 24), 1, 52, 12)):
        print("problem", nx, nys, nbx, flush=True)
        c = stepped(nx, nys, nbx, seed, nsteps)
        v_alloc = c.max_slip_velocity()
        v_noalloc = c.max_slip_velocity(noalloc=True)
        want = max(float(c.get_iface(k, "V").max()) for k in range(len(c.ifaces_n)))
        print("max V", v_alloc, v_noalloc, want, flush=True)
        assert v_noalloc == v_alloc == max(want, 0.0)
        assert v_alloc > 1e-6
        c.close()


def test_noalloc_refuses_beyond_its_node_limit(gpu):
    c = stepped(16400, (12, 12), 1, 53, 0)
    assert sum(c.ifaces_n) == 16400
    with pytest.raises(FdmError, match="exceed its limit of 16384"):
        c.max_slip_velocity(noalloc=True)
    v = c.max_slip_velocity()
    assert v == float(c.get_iface(0, "V").max())
    c.close()
