"""HIP path against the committed golden trajectories of the oracle, including
the full-size 961-camera scene (187,103 points, 1.68M observations), plus
size-independent properties of that solve."""
import json
import os
import sys

import numpy as np
import pytest

from rpba_amd import api, capi
from tolerance import check_trajectory

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = [c for c in make_golden.CASES if os.path.exists(os.path.join(GOLD, c + ".json"))]


# precision 3: the FP64 whitened-Jacobian records (the same products in
# another association), held to the same bar against the same fixtures
@pytest.mark.parametrize("precision", [0, 3])
@pytest.mark.parametrize("case", CASES)
def test_gpu_matches_golden(case, precision):
    fix = json.load(open(os.path.join(GOLD, case + ".json")))
    sc = make_golden.make_scene(make_golden.CASES[case]["scene"])
    print(f"[golden {case} precision={precision}] GPU solve n_obs={sc['n_obs']}", flush=True)
    fl = dict(fix["flags"], precision=precision, schur=2) if precision else fix["flags"]
    rc, out, st, log = api.calc_plsa(sc, capi.make_flags(**fl))
    assert rc == fix["status"]
    assert st["itt"] == fix["itt"] and st["passes_done"] == fix["passes_done"]
    check_trajectory(st["trace_sig0"], fix["trace_sig0"], st["trace_pcg"], fix["trace_pcg"],
                     st["trace_accept"], fix["trace_accept"], st["trace_nrobs"], fix["trace_nrobs"])
    assert make_golden.digest(out["obs_kept"]) == fix["obs_kept_sha256"]
    assert make_golden.digest(out["pt_kept"]) == fix["pt_kept_sha256"]
    for p in range(fix["passes_done"]):
        assert abs(st["sigma_out"][p] - fix["sigma_out"][p]) <= 1e-6 * fix["sigma_out"][p]
        assert st["redundancy"][p] == fix["redundancy"][p] and st["nobs"][p] == fix["nobs"][p]
    assert np.allclose(out["pts"].astype(np.float64).mean(axis=0), fix["pts_mean"], rtol=1e-5, atol=1e-6)
    assert np.allclose(out["cam_f"][:8], fix["cam_f_first"], rtol=1e-6)
    assert np.allclose(out["cam_P"][:4, :, 3].ravel(), fix["cam_t_first"], rtol=1e-5, atol=1e-6)
    assert len(log) == len(fix["log"])


def test_full_size_properties():
    """961 cameras at full size: converged to the 0.5 px noise, outliers
    found, bitwise reproducible, RMS of the inliers at the noise level."""
    sc = make_golden.make_scene(("c2_961", {}))
    print("[c2 properties] two GPU solves", flush=True)
    r1 = api.calc_plsa(sc, capi.make_flags(robust=1))
    r2 = api.calc_plsa(sc, capi.make_flags(robust=1))
    assert r1[0] == 0 and r1[2]["passes_done"] == 2
    for key in ("cam_P", "cam_f", "cam_d", "pts", "obs_kept", "pt_kept"):
        assert np.array_equal(r1[1][key], r2[1][key]), key
    st = r1[2]
    assert abs(st["sigma_out"][1] - 0.5) < 0.01
    assert abs(st["rms_px"] - 0.5) < 0.06
    truth = sc["truth"]["outlier"]
    kept = r1[1]["obs_kept"].astype(bool)
    assert (~kept & truth).sum() >= 0.97 * truth.sum()
    # the solve reduced sigma0 from the initial values to the noise level
    assert st["sigma_out"][1] < 0.2 * st["sigma_in"][0]
