"""Problems built by the Fortran host from .in files: HIP library vs CPU
oracle, and the fdmap driver executable end to end."""
import os
import re
import subprocess

import numpy as np
import pytest

from fdmap_github_amd import capi, host, inputs
from fdmap_github_amd.device import Context
from oracle.binding import OracleContext

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


def run_pair(path, nsteps, precision="fp64"):
    hp = host.HostProblem(path, precision=precision)
    F0 = hp.fields()
    dev = Context(hp.ptr)
    dev.set_fields(F0)
    dev.init_edges(hp.t)
    dev.step(hp.t, hp.dt, 1, nsteps)
    hp2 = host.HostProblem(path, precision="fp64")
    orc = OracleContext(hp2.ptr, threads=4)
    orc.set_fields(F0)
    orc.init_edges(hp2.t)
    orc.step(hp2.t, hp2.dt, 1, nsteps)
    return hp, dev, orc


def test_c1_point_source_small_bit_exact(gpu, tmp_path):
    p = inputs.write(str(tmp_path / "c1.in"), inputs.c1_pointsource(name=str(tmp_path / "c1"), n=81,
                                                                     nsteps=150, outputs=False))
    hp, dev, orc = run_pair(p, 150)
    Fd, DFd = dev.get_fields()
    Fo, DFo = orc.get_fields()
    print("max |v|", np.abs(Fo[:2]).max(), "max |dF|", np.abs(Fd - Fo).max())
    assert np.abs(Fo[:2]).max() > 0  # the source radiated
    assert np.array_equal(Fd, Fo)
    assert np.array_equal(DFd, DFo)


def test_c2_planar_rupture_small(gpu, tmp_path):
    p = inputs.write(str(tmp_path / "c2.in"), inputs.c2_planar_rs(name=str(tmp_path / "c2"), nx=201,
                                                                   ny=101, nsteps=400, outputs=False,
                                                                   R=0.5))
    hp, dev, orc = run_pair(p, 400)
    Fd, _ = dev.get_fields()
    Fo, _ = orc.get_fields()
    e = rel_l2(Fd, Fo)
    print("fields", e)
    assert e <= 1e-12
    for f in ("D", "V", "S", "N", "Psi"):
        a, b = dev.get_iface(0, f), orc.get_iface(0, f)
        print(f, rel_l2(a, b), np.abs(b).max())
        assert rel_l2(a, b) <= 1e-12
    assert orc.get_iface(0, "V").max() > 1e-3  # rupture nucleated


def test_fdmap_executable_outputs(gpu, tmp_path):
    nsteps = 60
    p = inputs.write(str(tmp_path / "c2x.in"),
                     inputs.c2_planar_rs(name=str(tmp_path / "c2x"), nx=121, ny=61, nsteps=nsteps,
                                         outputs=True, R=0.5))
    r = subprocess.run([host.EXE_PATH, p], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "timing: 1" in r.stdout and "Finished problem" in r.stdout
    # stride nsteps//100 -> 1: V written at every step (stage-1 values)
    V = np.fromfile(str(tmp_path / "c2x_I_V.dat"), dtype=np.float32).reshape(-1, 121)
    t = np.fromfile(str(tmp_path / "c2x_I_V.t"), dtype=np.float32)
    assert V.shape[0] == nsteps + 1 and t.shape[0] == nsteps + 1
    # oracle: stage-1 V of step n is what the driver writes for step n
    hp = host.HostProblem(p)
    orc = OracleContext(hp.ptr, threads=4)
    orc.set_fields(hp.fields())
    orc.init_edges(hp.t)
    for n in (1, 25, nsteps):
        o2 = OracleContext(hp.ptr, threads=4)
        o2.set_fields(hp.fields())
        o2.init_edges(hp.t)
        if n > 1:
            o2.step(hp.t, hp.dt, 1, n - 1)
        o2.stage1(hp.t, hp.dt, n)
        Vo = o2.get_iface(0, "V", stage1=True).astype(np.float32)
        np.testing.assert_allclose(V[n - 1], Vo, rtol=1e-6, atol=1e-30)
        assert abs(t[n - 1] - np.float32(hp.t + (n - 1) * hp.dt)) <= 1e-6 * max(1.0, abs(t[n - 1]))
    m = open(str(tmp_path / "c2x.m")).read()
    assert "I_V.nx =" in m and "I_V.field = 'V';" in m


def test_antiplane_rupture_from_input_file(gpu, tmp_path):
    """an antiplane (mode 3) rate-and-state rupture built from its .in file
    after the reference's mode-3 example: vertical fault between two blocks,
    free surface, Gaussian load"""
    p = inputs.write(str(tmp_path / "m3.in"),
                     inputs.antiplane_rs(name=str(tmp_path / "m3"), nx=61, ny=61, nsteps=250,
                                         outputs=False))
    hp, dev, orc = run_pair(p, 250)
    Fd, _ = dev.get_fields()
    Fo, _ = orc.get_fields()
    assert Fd.shape == (3, 61, 122)
    e = rel_l2(Fd, Fo)
    print("fields", e)
    assert e <= 1e-12
    for f in ("D", "V", "S", "N", "Psi"):
        a, b = dev.get_iface(0, f), orc.get_iface(0, f)
        print(f, rel_l2(a, b))
        assert rel_l2(a, b) <= 1e-12
    assert orc.get_iface(0, "V").max() > 1e-3  # rupture nucleated


def test_antiplane_rupture_in_a_medium_from_input_file(gpu, tmp_path):
    """the antiplane rupture with a heterogeneous medium file of the whole
    grid (material.f90:226-321): the host hands rho, cs, cp of every point
    and the edge impedances at each node; HIP against the oracle"""
    nx, ny, h = 61, 61, 0.1
    rho, cs, cp = inputs.smooth_medium(2 * nx, ny, h, amp=0.15, seed=4)
    med = inputs.write_medium(str(tmp_path / "m3.medium"), rho, cs, cp)
    p = inputs.write(str(tmp_path / "m3h.in"),
                     inputs.antiplane_rs(name=str(tmp_path / "m3h"), nx=nx, ny=ny, h=h, nsteps=250,
                                         outputs=False, medium=med))
    hp, dev, orc = run_pair(p, 250)
    assert hp.ptr.contents.mat_rho
    e = rel_l2(dev.get_fields()[0], orc.get_fields()[0])
    print("fields", e)
    assert e <= 1e-12
    for f in ("D", "V", "S", "N", "Psi"):
        a, b = dev.get_iface(0, f), orc.get_iface(0, f)
        print(f, rel_l2(a, b))
        assert rel_l2(a, b) <= 1e-12
    assert orc.get_iface(0, "V").max() > 1e-3


def test_antiplane_executable_outputs(gpu, tmp_path):
    nsteps = 60
    p = inputs.write(str(tmp_path / "m3x.in"),
                     inputs.antiplane_rs(name=str(tmp_path / "m3x"), nx=41, ny=41, nsteps=nsteps))
    r = subprocess.run([host.EXE_PATH, p], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-800:])
    assert r.returncode == 0 and "Finished problem" in r.stdout
    vz = np.fromfile(str(tmp_path / "m3x_B_vz.dat"), dtype=np.float32).reshape(-1, 41, 82)
    sxz = np.fromfile(str(tmp_path / "m3x_B_sxz.dat"), dtype=np.float32).reshape(-1, 41, 82)
    V = np.fromfile(str(tmp_path / "m3x_I_V.dat"), dtype=np.float32).reshape(-1, 41)
    assert vz.shape[0] == nsteps // 6 + 1 and V.shape[0] == nsteps + 1
    assert np.isfinite(vz).all() and np.abs(sxz[0] - 30.0).max() == 0.0
    # the body record of step 6: stage-1 fields of step 7, i.e. after 6 steps
    hp = host.HostProblem(p)
    orc = OracleContext(hp.ptr, threads=4)
    orc.set_fields(hp.fields())
    orc.init_edges(hp.t)
    orc.step(hp.t, hp.dt, 1, 6)
    np.testing.assert_allclose(vz[1], orc.get_fields()[0][0].astype(np.float32), rtol=1e-6,
                               atol=1e-12)
    m = open(str(tmp_path / "m3x.m")).read()
    assert re.search(r"^mode = +3;$", m, re.M)


@pytest.mark.parametrize("fdmethod", [None, "SBP4"])
def test_gp229_forced_slip_weakening_rupture(gpu, tmp_path, fdmethod):
    """the reference's antiplane slip-weakening example with forced rupture
    expansion (gp229/gp229sw3ex1.in), at a quarter of its resolution; as the
    example is written (FDmethod left at SBP6) and with SBP4"""
    p = inputs.write(str(tmp_path / "gp.in"),
                     inputs.gp229_sw3(name=str(tmp_path / "gp"), nx=201, nsteps=150, outputs=False,
                                      fdmethod=fdmethod))
    hp, dev, orc = run_pair(p, 150)
    assert hp.ptr.contents.sbp == (6 if fdmethod is None else 4)
    e = rel_l2(dev.get_fields()[0], orc.get_fields()[0])
    print("fields", e)
    assert e <= 1e-12
    for f in ("D", "V", "S", "N"):
        a, b = dev.get_iface(0, f), orc.get_iface(0, f)
        print(f, rel_l2(a, b))
        assert rel_l2(a, b) <= 1e-12
    V = orc.get_iface(0, "V")
    assert V.max() > 0.1 and np.sum(V > 0) > 10  # the forced rupture has grown


@pytest.mark.parametrize("fdmethod", [None, "SBP4"])
def test_kali_fault_properties_from_files(gpu, tmp_path, fdmethod):
    """the reference's mode-3 aging-law example with a, b, the state and a
    fault prestress read from files (problems/kali-heterogeneous-input.in,
    files as problems/kali_input.m makes them), at 121 points a side (the
    example has 301).  Coarser grids and longer runs meet the aging law's
    stiffness: where the state has dropped, exp((f0 - Psi)/b) V0/L exceeds
    what LSRK(5,4) integrates stably at dt, the state blows up and the
    Newton solve fails, as the reference's would; 200 steps here stay well
    before that (the oracle alone runs to step 397 with SBP4).  As the
    example is written (FDmethod left at SBP6) and with SBP4."""
    p = inputs.write(str(tmp_path / "kali.in"),
                     inputs.kali_het(str(tmp_path), n=121, nsteps=200, outputs=False,
                                     fdmethod=fdmethod))
    hp, dev, orc = run_pair(p, 200)
    assert hp.ptr.contents.sbp == (6 if fdmethod is None else 4)
    e = rel_l2(dev.get_fields()[0], orc.get_fields()[0])
    print("fields", e)
    assert e <= 1e-12
    for f in ("D", "V", "S", "N", "Psi"):
        a, b = dev.get_iface(0, f), orc.get_iface(0, f)
        print(f, rel_l2(a, b))
        assert rel_l2(a, b) <= 1e-12
    # the properties along the fault are the files' (a varies, so does Psi)
    rs = np.ctypeslib.as_array(hp.ptr.contents.ifaces[0].rs, shape=(7, 121))
    assert np.ptp(rs[0]) > 0.009 and np.ptp(orc.get_iface(0, "Psi")) > 0.0
    assert np.isfinite(orc.get_iface(0, "V")).all()
