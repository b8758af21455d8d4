"""The 13,682-camera scene (4.46M points, 29M observations, about 71M upper
blocks of S): the S-formed solve against the matrix-free one.

The two forms compute the same S d in different orders (S blocks from the
covisibility pairs, or W and Y records without forming S), so their first LM
iterations agree to rounding.  The oracle cannot run at this size (its S alone
is 46 GB); this comparison is what pins the sizes the small parity scenes never
reach: more than 2^31 elements of S, record arrays above 4 GiB, and launches
whose one-workgroup-per-block grid would exceed the 32-bit work-item count of
a dispatch (the off-diagonal blocks past that count were once left unformed).
Tolerances: the rule of tolerance.check_trajectory over two LM iterations.

Also at this size: the whitened-Jacobian forms against plain FP64 through pass
0, and the learned preconditioner's whole solve by its properties."""
import sys

import numpy as np
import pytest

from rpba_amd import api, capi, scenes
from tolerance import check_trajectory

pytestmark = pytest.mark.gpu


def test_13682_cameras_formed_matches_matrix_free():
    sc = scenes.generate("c5_13682")
    traces = {}
    for schur in (0, 1):
        print(f"[c5_13682] schur={schur}: 2 LM iterations, n_obs={sc['n_obs']}", flush=True)
        sys.stdout.flush()
        ctx = api.Context(sc, capi.make_flags(robust=1, k2=1, k4=1, f=1, threads=1, schur=schur))
        rc, done, fin = ctx.run(2)
        assert rc == 0 and done == 2
        if schur == 0:
            sz = ctx.sizes()
            assert sz["blocks"] * 64 > 2**32, sz  # the size this test is for
        rc, res, tr = ctx.download()
        ctx.close()
        traces[schur] = tr
    a, b = traces[0], traces[1]
    check_trajectory(a["trace_sig0"], b["trace_sig0"], a["trace_pcg"], b["trace_pcg"],
                     a["trace_accept"], b["trace_accept"], a["trace_nrobs"], b["trace_nrobs"])
    assert np.all(np.isfinite(a["trace_sig0"]))


def _pcg_cap(sc):
    """The reference's PCG cap max(1000, nrioff) with k2, k4, f estimated
    (rbacore.cpp:1848-1849): 3 + 8 + 9 (N - 2) unknowns."""
    return max(1000, 3 + 8 + 9 * (sc["n_cams"] - 2))


@pytest.fixture(scope="module")
def scene_c5():
    return scenes.generate("c5_13682")


def _pass0(sc, flags, lp=None):
    """Pass 0 of the solve (the robust pass): stats after its RobustDelete."""
    ctx = api.Context(sc, flags)
    if lp is not None:
        ctx.set_lp(lp)
    while True:
        rc, done, fin = ctx.run(1)
        ps, itt, sg = ctx.progress()
        if rc != 0 or fin or ps >= 1:
            break
    assert rc == 0 and ps == 1
    rc, out, st = ctx.download()
    ctx.close()
    return out, st


def test_13682_cameras_whitened_forms_against_plain(scene_c5):
    """The whitened forms at full size against the plain FP64 form through
    pass 0 (lambda = 1e-3), where the forms converge alike: the FP64 records
    (the same products associated otherwise) within the trajectory rule of
    the parity tests, the FP32 records (consistently rounded, so S stays
    positive definite) with PCG counts within 2 per iteration, sig0 within
    1e-5 and the labels of all but 1e-5 of the observations alike; both
    converge in the PCG without reaching the reference's cap."""
    base = dict(robust=1, k2=1, k4=1, f=1, threads=1, schur=2)
    print("[c5_13682] pass 0: plain, precision 3, precision 2", flush=True)
    o0, s0 = _pass0(scene_c5, capi.make_flags(**base))
    o3, s3 = _pass0(scene_c5, capi.make_flags(**base, precision=3))
    o2, s2 = _pass0(scene_c5, capi.make_flags(**base, precision=2))
    print("[c5_13682] pcg plain", s0["trace_pcg"].tolist(), "p3", s3["trace_pcg"].tolist(), "p2",
          s2["trace_pcg"].tolist(), flush=True)
    check_trajectory(s3["trace_sig0"], s0["trace_sig0"], s3["trace_pcg"], s0["trace_pcg"],
                     s3["trace_accept"], s0["trace_accept"], s3["trace_nrobs"], s0["trace_nrobs"])
    assert np.array_equal(o3["obs_kept"], o0["obs_kept"])
    n = len(s0["trace_sig0"])
    assert len(s2["trace_sig0"]) == n
    rel = np.abs(s2["trace_sig0"] - s0["trace_sig0"]) / s0["trace_sig0"]
    print(f"[c5_13682] precision 2 sig0 rel {rel.tolist()}", flush=True)
    assert rel.max() < 1e-5
    assert np.abs(np.asarray(s2["trace_pcg"]) - np.asarray(s0["trace_pcg"])).max() <= 2
    assert (o2["obs_kept"] != o0["obs_kept"]).sum() <= 1e-5 * scene_c5["n_obs"]
    for s in (s0, s2, s3):
        assert max(s["trace_pcg"]) < _pcg_cap(scene_c5)


def test_13682_cameras_learned_form_properties(scene_c5):
    """The learned correction of block-Jacobi at full size, whole solve (S
    formed: 71M upper blocks): finite sigma0 in both passes, every PCG ends by
    the reference's stopping rule before the cap, the PCG counts reported; the
    trained network's weight stays in [0, 0.95)."""
    from rpba_amd import lp_train
    import os
    if not os.path.exists(lp_train.WEIGHTS):
        pytest.skip("no trained weights in the tree")
    lp = api.LpModel.load(lp_train.WEIGHTS)
    print("[c5_13682] learned preconditioner, whole solve", flush=True)
    ctx = api.Context(scene_c5, capi.make_flags(robust=1, k2=1, k4=1, f=1, threads=1, preconditioner=2))
    ctx.set_lp(lp)
    rc, done, fin = ctx.run(-1)
    w = ctx.last_omega()
    rc_d, out, st = ctx.download()
    ctx.close()
    print(f"[c5_13682 learned] pcg {st['trace_pcg'].tolist()} sigma {list(st['sigma_out'])} last weight {w}",
          flush=True)
    assert rc == 0 and fin and st["passes_done"] == 2
    assert np.all(np.isfinite(st["trace_sig0"])) and all(np.isfinite(st["sigma_out"]))
    assert max(st["trace_pcg"]) < _pcg_cap(scene_c5)
    assert 0.0 <= w < 0.95
