"""The FP32-Jacobian form (precision = 1: FP32 Jacobians and W, Y = W V^-1
records, FP64 residuals, sums, S, PCG and cost) against plain FP64 on the
same scenes.  Tolerances: final sigma0 within 1e-6 relative of the plain
form, robust labels differing on fewer than 0.01% of the observations."""
import numpy as np
import pytest

from rpba_amd import api, capi, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["c1_50", "c2_961"])
def test_fp32_jacobian_form_reaches_plain_sigma(name):
    sc = scenes.generate(name)
    print(f"[precision 1 {name}] plain and FP32-Jacobian solves", flush=True)
    r0 = api.calc_plsa(sc, capi.make_flags(robust=1, precision=0))
    r1 = api.calc_plsa(sc, capi.make_flags(robust=1, precision=1))
    assert r0[0] == 0 and r1[0] == 0 and r1[2]["passes_done"] == 2
    s0, s1 = r0[2]["sigma_out"][1], r1[2]["sigma_out"][1]
    assert abs(s1 - s0) <= 1e-6 * s0, (s0, s1)
    assert abs(r1[2]["rms_px"] - r0[2]["rms_px"]) <= 1e-6 * r0[2]["rms_px"]
    assert (r1[1]["obs_kept"] != r0[1]["obs_kept"]).sum() <= 1e-4 * sc["n_obs"]
    # the first linearisation differs from plain FP64 only by FP32 rounding
    t0, t1 = r0[2]["trace_sig0"], r1[2]["trace_sig0"]
    assert abs(t1[0] - t0[0]) <= 1e-5 * t0[0]


def test_fp32_jacobian_form_is_deterministic(scene_c1):
    a = api.calc_plsa(scene_c1, capi.make_flags(robust=1, precision=1))
    b = api.calc_plsa(scene_c1, capi.make_flags(robust=1, precision=1))
    for key in ("cam_P", "pts", "obs_kept"):
        assert np.array_equal(a[1][key], b[1][key])


def test_matrix_free_form_full_size():
    """The 961-camera scene with the matrix-free Schur product: plain FP64
    records against the explicit S (same arithmetic, other summation order),
    and the chosen form (FP32 records + matrix-free) against plain."""
    sc = scenes.generate("c2_961")
    print("[matrix-free c2_961] explicit, matrix-free FP64, matrix-free FP32", flush=True)
    r0 = api.calc_plsa(sc, capi.make_flags(robust=1))
    r1 = api.calc_plsa(sc, capi.make_flags(robust=1, schur=1))
    r2 = api.calc_plsa(sc, capi.make_flags(robust=1, schur=1, precision=1))
    for r in (r1, r2):
        assert r[0] == 0 and r[2]["passes_done"] == 2
    s0 = r0[2]["sigma_out"][1]
    assert abs(r1[2]["sigma_out"][1] - s0) <= 1e-6 * s0
    assert abs(r1[2]["trace_sig0"][0] - r0[2]["trace_sig0"][0]) <= 1e-8 * r0[2]["trace_sig0"][0]
    assert np.array_equal(r1[1]["obs_kept"], r0[1]["obs_kept"])
    assert abs(r2[2]["sigma_out"][1] - s0) <= 1e-6 * s0
    assert (r2[1]["obs_kept"] != r0[1]["obs_kept"]).sum() <= 1e-4 * sc["n_obs"]
