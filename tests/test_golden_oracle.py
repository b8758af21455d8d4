"""The oracle reproduces its committed golden trajectories (a regression pin
of the oracle itself; the 961- and 4,585-camera trajectories are only checked
on the GPU)."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle
from rpba_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
sys_path_golden = os.path.join(HERE, "golden")
import sys  # noqa: E402

sys.path.insert(0, sys_path_golden)
import make_golden  # noqa: E402

SMALL = [os.path.basename(p)[:-5] for p in sorted(glob.glob(os.path.join(sys_path_golden, "*.json")))
         if os.path.basename(p)[:-5] in make_golden.CASES and "c2_" not in p]


@pytest.mark.parametrize("case", SMALL)
def test_oracle_reproduces_fixture(case):
    fix = json.load(open(os.path.join(sys_path_golden, case + ".json")))
    sc = make_golden.make_scene(make_golden.CASES[case]["scene"])
    rc, out, st, log = oracle.solve(sc, capi.make_flags(**fix["flags"]), fsites=0)
    assert rc == fix["status"] and st["itt"] == fix["itt"]
    assert np.array_equal(st["trace_accept"], fix["trace_accept"])
    assert np.array_equal(st["trace_nrobs"], fix["trace_nrobs"])
    assert np.allclose(st["trace_sig0"], fix["trace_sig0"], rtol=1e-12)
    assert make_golden.digest(out["obs_kept"]) == fix["obs_kept_sha256"]
    assert log == fix["log"]
