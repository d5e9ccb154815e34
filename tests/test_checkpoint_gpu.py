"""Checkpoint, abort and resume through the fdmap driver executable
(checkpoint.f90, main.f90:150-235): a resumed run continues bit for bit,
so its output files end identical to those of an uninterrupted run.

The problem: two 121 x 61 blocks, RFL fault, plastic blocks with a low
friction coefficient (so the plastic strain grows from the first step),
outputs on the fault every 10 steps and in the body every 7 steps, so the
driver batches the steps between outputs and must stop at checkpoints."""
import os
import subprocess

import numpy as np
import pytest

from fdmap_github_amd import host, inputs

pytestmark = pytest.mark.gpu

NSTEPS = 60
OUT = [("I", "iface1", "V S Psi D", 10), ("B", "body", "vx sxy gammap", 7)]


def problem(d, ckpt=""):
    name = os.path.join(str(d), "c")
    txt = inputs.c2_planar_rs(name=name, nx=121, ny=61, nsteps=NSTEPS, outputs=False, R=0.5,
                              plastic=True)
    assert "mu = 0.6d0" in txt
    txt = txt.replace("mu = 0.6d0", "mu = 0.45d0")
    items = "".join(f"{n}\n{loc}\n{f}\n0d0 1d10 {st}\n" for n, loc, f, st in OUT)
    txt = txt.replace("!---end:output_list---", items + "!---end:output_list---")
    if ckpt:
        txt += f"\n&checkpoint_list {ckpt} /\n"
    return inputs.write(name + ".in", txt), name


def run(path, expect_ok=True):
    print("running", path, flush=True)
    r = subprocess.run([host.EXE_PATH, path], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-800:], flush=True)
    if expect_ok:
        assert r.returncode == 0
    return r


def outputs(name):
    out = {}
    for n, _, flds, _ in OUT:
        for f in flds.split():
            for ext in (".dat", ".t"):
                p = f"{name}_{n}_{f}{ext}"
                out[os.path.basename(p)] = open(p, "rb").read()
    return out


def record_offset(begin, st):
    """records before the one a run resumed after step begin rewrites
    (output.f90:950-963, nmin = 0 here)"""
    return begin // st + (0 if begin % st == 0 else 1)


def corrupt_after(name, begin):
    """overwrite every record from the resume point on with 0xFF bytes"""
    for n, loc, flds, st in OUT:
        rec = {"iface1": 121, "body": 121 * 122}[loc] * 4
        k = record_offset(begin, st)
        for f in flds.split():
            for ext, size in ((".dat", rec), (".t", 4)):
                p = f"{name}_{n}_{f}{ext}"
                data = bytearray(open(p, "rb").read())
                assert len(data) > k * size
                data[k * size:] = b"\xff" * (len(data) - k * size)
                open(p, "wb").write(bytes(data))


def ckpt_files(d, number):
    return sorted(f for f in os.listdir(str(d)) if f.endswith(f".ckpt{number}"))


@pytest.fixture(scope="module")
def reference_run(tmp_path_factory):
    d = tmp_path_factory.mktemp("full")
    p, name = problem(d)
    run(p)
    out = outputs(name)
    g = np.frombuffer(out["c_B_gammap.dat"], dtype=np.float32)
    assert g.max() > 0.0  # plastic strain grows
    return out


def test_interval_checkpoints_and_resume(gpu, tmp_path, reference_run):
    p, name = problem(tmp_path, "ckpt = T, interval = 25, del_final = F")
    run(p)
    # checkpoints after steps 25 and 50; the first was deleted after the
    # second was written, the last is kept (del_final = F)
    assert int(open(name + ".ckpt").read().split()[0]) == 50
    assert ckpt_files(tmp_path, 25) == []
    assert len(ckpt_files(tmp_path, 50)) == 1 + 2 * 4 + 2
    assert outputs(name) == reference_run

    corrupt_after(name, 50)
    assert outputs(name) != reference_run
    p, name = problem(tmp_path, "ckpt = T, resume = T, interval = 25")
    r = run(p)
    assert "reading  checkpoint at time step 50" in r.stdout
    assert outputs(name) == reference_run
    # a finished run removes its last checkpoint and marks the problem done
    assert int(open(name + ".ckpt").read().split()[0]) == -1
    assert ckpt_files(tmp_path, 50) == []
    r = run(p, expect_ok=False)
    assert r.returncode != 0 and "successfully completed" in r.stdout


def test_abort_file_and_resume(gpu, tmp_path, reference_run):
    """<name>.abort holding T: the run stops after its first steps with a
    checkpoint even though ckpt is off (checkpoint.f90:194-196)"""
    p, name = problem(tmp_path)
    with open(name + ".abort", "w") as f:
        f.write("T\n")
    r = run(p)
    assert "Aborting now..." in r.stdout
    begin = int(open(name + ".ckpt").read().split()[0])
    assert 1 <= begin < NSTEPS
    assert len(ckpt_files(tmp_path, begin)) == 11
    os.remove(name + ".abort")
    p, name = problem(tmp_path, "ckpt = T, resume = T, interval = 1000")
    r = run(p)
    assert f"reading  checkpoint at time step {begin}" in r.stdout
    assert outputs(name) == reference_run
    assert ckpt_files(tmp_path, begin) == []


def test_pml_resume_continues_bit_for_bit(gpu, tmp_path):
    """a layered problem keeps Wx, Wy in its field checkpoint: a run resumed
    from it ends with output files identical to an uninterrupted run's"""
    layer = dict(pmlx=True, pmly=True, n=2.0, sigma=60.0, mx=-0.4, px=0.4, my=-0.3, py=0.3)
    nx, ny = 121, 61

    def pb(d, ckpt=""):
        name = os.path.join(str(d), "p")
        txt = inputs.c2_planar_rs(name=name, nx=nx, ny=ny, nsteps=NSTEPS, outputs=False, R=0.5, pml=layer)
        items = "B\nbody\nvx sxy\n0d0 1d10 7\n"
        txt = txt.replace("!---end:output_list---", items + "!---end:output_list---")
        if ckpt:
            txt += f"\n&checkpoint_list {ckpt} /\n"
        return inputs.write(name + ".in", txt), name

    def outs(name):
        return {f: open(f"{name}_B_{f}.dat", "rb").read() for f in ("vx", "sxy")}

    full = tmp_path / "full"
    full.mkdir()
    p, name = pb(full)
    run(p)
    ref = outs(name)
    res = tmp_path / "res"
    res.mkdir()
    p, name = pb(res, "ckpt = T, interval = 25, del_final = F")
    run(p)
    assert outs(name) == ref
    assert os.path.getsize(name + "F.ckpt50") == (12 + 12) * nx * (2 * ny) * 8
    W = np.fromfile(name + "F.ckpt50", dtype="<f8").reshape(24, 2 * ny, nx)[12:]
    assert np.abs(W).max() > 0.0
    rec = nx * 2 * ny * 4
    k = record_offset(50, 7)
    for f in ("vx", "sxy"):
        path = f"{name}_B_{f}.dat"
        data = bytearray(open(path, "rb").read())
        data[k * rec:] = b"\xff" * (len(data) - k * rec)
        open(path, "wb").write(bytes(data))
    assert outs(name) != ref
    p, name = pb(res, "ckpt = T, resume = T, interval = 25")
    r = run(p)
    assert "reading  checkpoint at time step 50" in r.stdout
    assert outs(name) == ref
