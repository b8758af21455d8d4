"""HIP path vs the CPU oracle (plain FP64 mode) on the same seeded inputs,
through the C ABI.

Tolerances: the final sigma within 1e-6 relative (the FP64 bar); per-iteration
sig0 by the rule of tolerance.check_trajectory (first iteration 1e-8, all
iterations 1e-5); accept decisions, nrobs and the robust labels
(kept observations / points) exactly; parameters within 1e-6 of their scale.
The two implementations differ only in summation order."""
import sys

import numpy as np
import pytest

from oracle import oracle
from rpba_amd import api, capi, scenes
from tolerance import check_trajectory

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cmp_solves(sc, flags, tag):
    _say(f"[{tag}] oracle solve")
    rc0, out0, st0, log0 = oracle.solve(sc, flags, fsites=0)
    _say(f"[{tag}] GPU solve via rpba_calc_plsa n_cams={sc['n_cams']} n_obs={sc['n_obs']}")
    rc1, out1, st1, log1 = api.calc_plsa(sc, flags)
    assert rc1 == rc0
    assert st1["passes_done"] == st0["passes_done"]
    assert st1["itt"] == st0["itt"], (st0["trace_sig0"], st1["trace_sig0"])
    check_trajectory(st1["trace_sig0"], st0["trace_sig0"], st1["trace_pcg"], st0["trace_pcg"],
                     st1["trace_accept"], st0["trace_accept"], st1["trace_nrobs"], st0["trace_nrobs"])
    for p in range(st0["passes_done"]):
        for key in ("sigma_in", "sigma_out"):
            assert abs(st1[key][p] - st0[key][p]) <= 1e-6 * abs(st0[key][p])
        for key in ("redundancy", "nobs", "nunknowns"):
            assert st1[key][p] == st0[key][p]
    assert np.array_equal(out1["obs_kept"], out0["obs_kept"])
    assert np.array_equal(out1["pt_kept"], out0["pt_kept"])
    for key in ("cam_P", "cam_f", "cam_d", "pts"):
        a, b = out1[key].astype(np.float64), out0[key].astype(np.float64)
        scale = max(np.abs(b).max(), 1e-30)
        assert np.abs(a - b).max() <= 1e-6 * scale + 1e-6 * np.abs(b).max(), key
    assert len(log1) == len(log0)
    for l0, l1 in zip(log0, log1):
        t0, t1 = l0.split(), l1.split()
        assert t0[0] == t1[0] and t0[-4:] == t1[-4:]
    return st0, st1, out1


def test_exteval_matches_oracle(scene_c1):
    _say("[exteval c1]")
    fl = capi.make_flags(exteval=1)
    s0, n0, d0 = oracle.exteval(scene_c1, fl, fsites=0)
    rc, out, st, log = api.calc_plsa(scene_c1, fl)
    assert rc == 0
    assert st["exteval_nobs"] == n0 and st["exteval_dof"] == d0
    assert abs(st["exteval_sigma"] - s0) <= 1e-12 * s0
    assert log[0].startswith("ITTTT: ")
    # Exteval does not change the block
    assert np.array_equal(out["pts"], scene_c1["pts"])


def _weighted_c1():
    """The 50-camera scene with per-observation 2x2 weights (cross term
    included) and image sizes: ihw2 = 563^2 scales sigma and clamps residual
    components at 8.9 px, which the planted 50 px outliers reach."""
    return scenes.generate(None, seed=1, n_cams=50, n_pts=5000, track="uniform2_8", outlier_frac=0.02,
                           weights="random", img_size=(640, 480))


@pytest.mark.parametrize("robust", [1, 0])
def test_weighted_scene_with_image_sizes_matches_oracle(robust):
    sc = _weighted_c1()
    st0, st1, _ = _cmp_solves(sc, capi.make_flags(robust=robust), f"weighted c1 robust={robust}")
    assert st0["sigma_out"][1] > 100.0  # sigma in the ihw2 scale


def test_weighted_exteval_matches_oracle():
    sc = _weighted_c1()
    _say("[exteval weighted c1]")
    fl = capi.make_flags(exteval=1)
    s0, n0, d0 = oracle.exteval(sc, fl, fsites=0)
    rc, out, st, log = api.calc_plsa(sc, fl)
    assert rc == 0 and st["exteval_nobs"] == n0 and st["exteval_dof"] == d0
    assert abs(st["exteval_sigma"] - s0) <= 1e-12 * s0


def test_robust_solve_c1_matches_oracle(scene_c1):
    st0, st1, out1 = _cmp_solves(scene_c1, capi.make_flags(robust=1), "c1 robust")
    assert abs(st1["sigma_out"][1] - 0.5) < 0.02


@pytest.mark.parametrize("robust,k2,k4,f", [(0, 1, 1, 1), (1, 1, 0, 1), (0, 0, 0, 0), (1, 0, 0, 1)])
def test_solve_variants_match_oracle(robust, k2, k4, f):
    sc = scenes.generate(None, seed=31, n_cams=12, n_pts=600, track="uniform2_8", outlier_frac=0.03)
    _cmp_solves(sc, capi.make_flags(robust=robust, k2=k2, k4=k4, f=f), f"r{robust} k2{k2} k4{k4} f{f}")


def test_tiny_scene(tiny_scene):
    _cmp_solves(tiny_scene, capi.make_flags(robust=1), "tiny")


def test_deterministic(scene_c1):
    _say("[determinism c1]")
    r1 = api.calc_plsa(scene_c1, capi.make_flags(robust=1))
    r2 = api.calc_plsa(scene_c1, capi.make_flags(robust=1))
    for key in ("cam_P", "cam_f", "cam_d", "pts", "obs_kept"):
        assert np.array_equal(r1[1][key], r2[1][key])
    assert np.array_equal(r1[2]["trace_sig0"], r2[2]["trace_sig0"])


def test_context_chunks_and_reset(scene_c1):
    _say("[ctx chunks c1]")
    ref = api.calc_plsa(scene_c1, capi.make_flags(robust=1))
    ctx = api.Context(scene_c1, capi.make_flags(robust=1))
    total = 0
    while True:
        rc, done, fin = ctx.run(2)
        assert rc == 0
        total += done
        if fin:
            break
    rc, out, st = ctx.download()
    assert st["itt"] == ref[2]["itt"]
    assert np.array_equal(st["trace_sig0"], ref[2]["trace_sig0"])
    assert np.array_equal(out["pts"], ref[1]["pts"])
    assert ctx.reset() == 0
    rc, done, fin = ctx.run(-1)
    assert rc == 0 and fin
    rc, out2, st2 = ctx.download()
    assert np.array_equal(st2["trace_sig0"], ref[2]["trace_sig0"])
    ctx.close()


def test_error_no_observations_in_image():
    _say("[error empty image]")
    sc = scenes.generate(None, seed=8, n_cams=6, n_pts=80, track="uniform2_8", outlier_frac=0.0)
    m = sc["obs_cam"] != 5
    bad = dict(sc, obs_cam=sc["obs_cam"][m], obs_pt=sc["obs_pt"][m], obs_xy=sc["obs_xy"][m], n_obs=int(m.sum()))
    rc0, out0, _, log0 = oracle.solve(bad, capi.make_flags(robust=0))
    rc1, out1, _, log1 = api.calc_plsa(bad, capi.make_flags(robust=0))
    assert rc1 == rc0 == capi.RPBA_E_NO_OBS_IN_IMAGE
    assert log1 == log0
    assert np.array_equal(out1["pts"], out0["pts"])



def _banded_scene():
    # each point seen by cameras within a window of 10 consecutive ones: S is
    # banded (540 of 1,770 blocks), and the strip slabs S is stored in hold
    # every count of rows from 1 to 7, with blocks missing inside a strip
    return scenes.generate(None, seed=51, n_cams=60, n_pts=3000, track="uniform2_8", outlier_frac=0.02,
                           cam_window=10)


def test_banded_covisibility_matches_oracle():
    _cmp_solves(_banded_scene(), capi.make_flags(robust=1), "banded")


def test_banded_covisibility_matrix_free_matches_oracle():
    _cmp_solves(_banded_scene(), capi.make_flags(robust=1, schur=1), "banded-mf")


def test_matrix_free_schur_matches_oracle(scene_c1):
    """schur = 1 never forms the off-diagonal blocks; the oracle forms S block
    by block as the reference does.  The product is the same sum in another
    order, so the same bar applies."""
    _cmp_solves(scene_c1, capi.make_flags(robust=1, schur=1), "c1 matrix-free")


def test_matrix_free_schur_variants_match_oracle():
    sc = scenes.generate(None, seed=31, n_cams=12, n_pts=600, track="uniform2_8", outlier_frac=0.03)
    _cmp_solves(sc, capi.make_flags(robust=0, k2=1, k4=0, f=1, schur=1), "matrix-free r0 k2 f")


def test_pcg_graph_replay_same_bits(scene_c1):
    """The PCG replayed from a hipGraph gives the same bits as launches."""
    flags = capi.make_flags(robust=1)
    outs = []
    for on in (False, True, True):
        ctx = api.Context(scene_c1, flags)
        ctx.set_graph(on)
        rc, n, fin = ctx.run(-1)
        assert rc == 0 and fin
        outs.append(ctx.download())
        ctx.close()
    for rc, out, st in outs[1:]:
        assert np.array_equal(out["pts"], outs[0][1]["pts"]) and st["sig0_final"] == outs[0][2]["sig0_final"]
        assert np.array_equal(st["trace_sig0"], outs[0][2]["trace_sig0"])


@pytest.mark.parametrize("n_cams,n_pts", [(50, 5000), (300, 5000)])
def test_auto_product_is_the_explicit_one(n_cams, n_pts):
    """schur = 2 takes the product streaming fewer bytes per PCG iteration
    (S formed on the 50-camera scene, matrix-free on 300 cameras with few
    points per camera pair); its solve is bitwise the explicit form's."""
    sc = scenes.generate(None, seed=23, n_cams=n_cams, n_pts=n_pts, track="uniform2_8", outlier_frac=0.02)
    _say(f"[auto product] {n_cams} cameras")
    ctx = api.Context(sc, capi.make_flags(robust=1, schur=2))
    ctx.run(-1)
    rc, out, st = ctx.download()
    mf = ctx.sizes()["blocks"] == 0
    ctx.close()
    assert rc == 0
    assert mf == (n_cams == 300)
    rc1, out1, st1, _ = api.calc_plsa(sc, capi.make_flags(robust=1, schur=1 if mf else 0))
    assert rc1 == 0
    assert np.array_equal(np.asarray(st["trace_sig0"]), np.asarray(st1["trace_sig0"]))
    assert np.array_equal(out["obs_kept"], out1["obs_kept"]) and np.array_equal(out["pts"], out1["pts"])
