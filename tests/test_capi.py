"""The C-ABI boundary without a GPU: the HIP library loads, exports every
function include/fdmap.h declares, and the Python descriptors match the
header's structure layout.  Apart from fdm_create's argument checks, which
return before any device call, no entry point is called (each one needs a
device)."""
import ctypes as C
import os
import re

import pytest

from fdmap_github_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fdmap.h")
ENSEMBLE_HEADER = os.path.join(ROOT, "include", "fdmap_ensemble.h")
ADJOINT_HEADER = os.path.join(ROOT, "include", "fdmap_adjoint.h")


def declared_functions(path=HEADER):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(fdm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ("fdm_create", "fdm_destroy", "fdm_last_error", "fdm_set_fields",
                     "fdm_get_fields", "fdm_init_edges", "fdm_step", "fdm_stage1",
                     "fdm_get_iface"):
        assert required in names
    # the Python binding knows every declared entry point, and nothing else
    assert sorted(n for n, _, _ in capi.ENTRY_POINTS + capi.NOALLOC_ENTRY_POINTS
                  + capi.STREAM_ENTRY_POINTS) == names
    assert sorted(n for n, _, _ in capi.ENSEMBLE_ENTRY_POINTS) == declared_functions(ENSEMBLE_HEADER)
    assert sorted(n for n, _, _ in capi.ADJOINT_ENTRY_POINTS) == declared_functions(ADJOINT_HEADER)


def test_library_exports_every_declared_symbol():
    path = os.path.join(capi.HERE, "libfdmap_hip.so")
    if not os.path.exists(path):
        pytest.fail("libfdmap_hip.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    names = (declared_functions() + declared_functions(ENSEMBLE_HEADER) +
             declared_functions(ADJOINT_HEADER))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_descriptor_layout_matches_header():
    """Compile a probe against include/fdmap.h that prints sizeof/offsetof of
    the descriptors and compare with the ctypes mirrors."""
    import subprocess
    import tempfile

    fields = {
        "fdm_side": ["bc", "n", "normal", "M", "F0", "x", "y"],
        "fdm_block": ["mx", "py", "plastic", "mms", "id", "rho", "nu", "mu", "K", "s0", "side"],
        "fdm_iface": ["direction", "law", "rup_threshold", "ld_T", "nhat", "rs", "sw", "S0base",
                      "outside", "Psi0"],
        "fdm_source_place": ["block", "i0", "j0", "dx", "dy"],
        "fdm_source": ["nplace", "stf", "t_center", "Fx", "Myy", "place"],
        "fdm_problem": ["nx", "ny", "nblocks", "nifaces", "precision", "nstage", "rkA", "rkB",
                        "rkC", "xq", "Ji", "x", "y", "blocks", "ifaces", "source", "mat_rho",
                        "mat_cs", "mat_cp"],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fl in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fl:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for st, fl in fields.items():
        cls = getattr(capi, st)
        assert int(got[st]) == C.sizeof(cls), st
        for f in fl:
            assert int(got[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="libfdmap_hip.so not found"):
        capi.load_library(str(tmp_path / "libfdmap_hip.so"))


@pytest.mark.parametrize("change,msg", [
    (dict(sbp=8), "SBP2, SBP4 or SBP6"),
    (dict(sbp=6, Cdiss=300.0), "need FDmethod SBP4"),
    (dict(sbp=2, pml=dict(x=True)), "need FDmethod SBP4"),
    (dict(sbp=6, nys=(11,)), "SBP6 needs at least 12"),
])
def test_create_refuses_what_the_operator_cannot_serve(change, msg):
    """fdm_create validates the operator before it touches a device: the
    range of FDmethod, dissipation and PML (built on SBP4's tables in the
    reference, rates.f90:1397-2619) and the smallest block the closures fit"""
    import problems
    from fdmap_github_amd.device import Context, FdmError
    kw = dict(nx=24, nys=(20,), seed=1, coupling="locked")
    sbp = change.pop("sbp")
    kw.update(change)
    prob, _ = problems.build(**kw)
    prob.sbp = sbp
    with pytest.raises(FdmError, match=msg):
        Context(prob)
