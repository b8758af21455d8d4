"""The whitened-Jacobian forms against the oracle's matching arithmetic mode,
through the C ABI.

precision 2 stores the whitened Jacobians L^T Bc, L^T Bp (observation weight
W = L L^T) rounded to FP32 once and forms V, g, U, W, the Schur terms and the
PCG product from them in FP64; precision 3 stores the same records in FP64.
The oracle's modes 2 and 3 (oracle.arithmetic) form S block by block from the
same values in the reference's order, so the two differ only in summation
order and the comparison rule is that of the plain form (tolerance.py: first
iteration 1e-8, every iteration 1e-5, decisions, nrobs and labels exactly,
converged sigmas 1e-6, parameters 1e-6 of their scale)."""
import numpy as np
import pytest

from oracle import oracle
from rpba_amd import api, capi, scenes
from test_gpu_parity import _banded_scene, _cmp_solves, _say, _weighted_c1

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [2, 3])
@pytest.mark.parametrize("robust", [1, 0])
def test_whitened_form_matches_oracle(scene_c1, precision, robust):
    with oracle.arithmetic(precision):
        _cmp_solves(scene_c1, capi.make_flags(robust=robust, precision=precision, schur=2),
                    f"c1 precision={precision} robust={robust}")


def test_fp64_whitened_form_matches_reference_order_oracle(scene_c1, tiny_scene):
    """precision 3 against the oracle in the reference's own order (mode 0):
    the whitening is an exact identity (J^T W J = (L^T J)^T (L^T J)), so in
    FP64 the plain form's bar holds, on the scenes of the plain parity tests."""
    small = scenes.generate(None, seed=31, n_cams=12, n_pts=600, track="uniform2_8", outlier_frac=0.03)
    cases = [(scene_c1, dict(robust=1), "c1 robust"), (scene_c1, dict(robust=0), "c1 plain"),
             (small, dict(robust=1, k2=1, k4=0, f=1), "r1 k2 f"), (small, dict(robust=0, k2=0, k4=0, f=0), "r0"),
             (tiny_scene, dict(robust=1), "tiny"), (_banded_scene(), dict(robust=1), "banded"),
             (_weighted_c1(), dict(robust=1), "weighted c1")]
    for sc, kw, tag in cases:
        _cmp_solves(sc, capi.make_flags(precision=3, schur=1, **kw), f"precision 3 vs reference order: {tag}")


@pytest.mark.parametrize("precision", [2, 3])
def test_whitened_form_weighted_scene_matches_oracle(precision):
    """Correlated per-observation weights (wxy != 0), so L is not diagonal,
    and image sizes (ihw2, the residual clamp)."""
    with oracle.arithmetic(precision):
        _cmp_solves(_weighted_c1(), capi.make_flags(robust=1, precision=precision, schur=1),
                    f"weighted c1 precision={precision}")


def test_fp32_whitened_error_against_fp64():
    """The FP32 form's error against the plain FP64 solve on the 50-camera
    robust scene: final sigma0 within 1e-5 of the oracle's reference order,
    RMS within 1e-5 of the plain GPU form's (the oracle does not report it),
    labels within 0.1% of the observations, points within 1e-4 of the
    scene's scale."""
    sc = scenes.generate("c1_50")
    _say("[precision 2 vs plain] oracle plain solve")
    rc0, out0, st0, _ = oracle.solve(sc, capi.make_flags(robust=1), fsites=0)
    _say("[precision 2 vs plain] GPU plain and precision 2")
    rcp, outp, stp, _ = api.calc_plsa(sc, capi.make_flags(robust=1))
    rc1, out1, st1, _ = api.calc_plsa(sc, capi.make_flags(robust=1, precision=2, schur=2))
    assert rc0 == 0 and rcp == 0 and rc1 == 0
    assert abs(st1["sigma_out"][1] - st0["sigma_out"][1]) <= 1e-5 * st0["sigma_out"][1]
    assert abs(st1["rms_px"] - stp["rms_px"]) <= 1e-5 * stp["rms_px"]
    assert (out1["obs_kept"] != out0["obs_kept"]).sum() <= 1e-3 * sc["n_obs"]
    assert np.abs(out1["pts"].astype(np.float64) - out0["pts"]).max() <= 1e-4 * np.abs(out0["pts"]).max()


def test_whitened_form_is_deterministic(scene_c1):
    a = api.calc_plsa(scene_c1, capi.make_flags(robust=1, precision=2, schur=1))
    b = api.calc_plsa(scene_c1, capi.make_flags(robust=1, precision=2, schur=1))
    for key in ("cam_P", "pts", "obs_kept"):
        assert np.array_equal(a[1][key], b[1][key])
