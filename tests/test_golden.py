"""The oracle against its committed fixtures at a benchmark configuration's
full size (tests/golden/make_golden.py wrote them): the point source, one
401 x 401 block, 1000 steps from its .in file, bit for bit.  The fixtures
also anchor the GPU tests of the same configurations."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as mg  # noqa: E402


def test_c1_oracle_reproduces_its_fixture(tmp_path):
    g = np.load(os.path.join(HERE, "golden", "c1.npz"))
    o = mg.run_oracle(mg.c1_input(str(tmp_path)), mg.C1["nsteps"])
    F, DF = o.get_fields()
    s = mg.summary_c1(F, DF)
    assert np.array_equal(s["F_pts"], g["F_pts"]) and np.array_equal(s["DF_pts"], g["DF_pts"])
    assert s["F_sha"] == str(g["F_sha"]) and s["DF_sha"] == str(g["DF_sha"])
    # the source has radiated across the block
    assert np.abs(F[:2]).max() > 1e-3 and g["F_norm"][0] > 0.0
