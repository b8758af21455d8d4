"""Dense form (solver = 1): the batched FP64-MFMA Cholesky against numpy on
random SPD systems of uneven size, the result-invariance of the tuned
variants, and the whole LM solve against the oracle's dense mode (the same
exact-solve algorithm in plain C) under the trajectory rule of the parity
tests."""
import sys

import numpy as np
import pytest

from oracle import oracle
from rpba_amd import api, capi, scenes
from test_gpu_parity import _cmp_solves

pytestmark = pytest.mark.gpu


def _spd(rng, ncam, cond=1e6):
    n = 9 * ncam
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    S = (Q * np.logspace(0, np.log10(cond), n)) @ Q.T
    return 0.5 * (S + S.T)


def test_batched_cholesky_against_numpy_and_variants():
    rng = np.random.default_rng(7)
    ncams = [1, 3, 7, 8, 15, 30, 2, 45]  # 9 .. 405 rows: 1 .. 7 tiles, uneven
    systems = [_spd(rng, k) for k in ncams]
    rhs = [rng.normal(size=s.shape[0]) for s in systems]
    print("[dense batch] autotuned solve of", ncams, flush=True)
    xs = api.dense_solve_batch(systems, rhs)
    for S, b, x in zip(systems, rhs, xs):
        xr = np.linalg.solve(S, b)
        # backward error of an FP64 Cholesky: ~ n eps cond
        assert np.abs(x - xr).max() <= 1e-8 * np.abs(xr).max()
        assert np.linalg.norm(S @ x - b) <= 1e-10 * np.linalg.norm(S) * np.linalg.norm(x)
    for variant, cap in [(0, 256), (0, 2048), (1, 512), (1, 64)]:
        print(f"[dense batch] variant {variant} cap {cap}", flush=True)
        xv = api.dense_solve_batch(systems, rhs, variant=variant, cap=cap)
        for a, b in zip(xs, xv):
            assert np.array_equal(a, b)


def test_dense_form_solve_matches_oracle(scene_c1):
    st0, st1, _ = _cmp_solves(scene_c1, capi.make_flags(robust=1, solver=1), "c1 dense")
    assert st1["pcg_iters"] == 0
    assert abs(st1["sigma_out"][1] - 0.5) < 0.02


@pytest.mark.parametrize("robust,k2,k4,f", [(0, 1, 1, 1), (1, 0, 0, 1), (0, 0, 0, 0)])
def test_dense_form_variants_match_oracle(robust, k2, k4, f):
    sc = scenes.generate(None, seed=31, n_cams=12, n_pts=600, track="uniform2_8", outlier_frac=0.03)
    _cmp_solves(sc, capi.make_flags(robust=robust, k2=k2, k4=k4, f=f, solver=1), f"dense r{robust} k2{k2} k4{k4} f{f}")


def test_dense_form_context_reports_tuning(scene_c1):
    ctx = api.Context(scene_c1, capi.make_flags(robust=1, solver=1))
    t = ctx.tuning()
    print("[dense tuning]", t, flush=True)
    assert t.startswith("chosen ")
    rc, n, fin = ctx.run(-1)
    assert rc == 0 and fin
    rc, out_a, st_a = ctx.download()
    ctx.reset()
    ctx.run(-1)
    rc, out_b, st_b = ctx.download()
    assert np.array_equal(out_a["pts"], out_b["pts"]) and st_a["sig0_final"] == st_b["sig0_final"]
    ctx.close()
