"""The batched-blocks workload at its real sizes: 64 independent blocks of
20-400 cameras (log-uniform; 8.97M observations in all) in one batched solve
(union launch sequence per LM iteration, hipGraph replay), every block against
the oracle's solve of that block alone with the same camera solve, under the
trajectory rule of the parity tests, labels bit for bit: the batched FP64
matrix-core Cholesky against the oracle's exact solve
(tests/golden/c4_blocks.json), and the block-segmented PCG against the
oracle's SparsePCG, the reference's camera solve (c4_blocks_pcg.json)."""
import json
import os
import sys

import numpy as np
import pytest

from rpba_amd import api, capi, scenes
from tolerance import check_trajectory

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
import make_golden  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fixture", ["c4_blocks.json", "c4_blocks_pcg.json"])
def test_c4_blocks_match_oracle_per_block(fixture):
    if not os.path.exists(os.path.join(GOLD, fixture)):
        pytest.skip("fixture not written")
    fix = json.load(open(os.path.join(GOLD, fixture)))
    print(f"[c4] generating 64 blocks; fixture {fixture}, solver {fix['flags']['solver']}", flush=True)
    blocks = scenes.generate_batch(**fix["config"])
    flags = capi.make_flags(**fix["flags"])
    print(f"[c4] batched solve of {len(blocks)} blocks, {sum(b['n_obs'] for b in blocks)} observations", flush=True)
    res = api.calc_plsa_batch(blocks, flags)
    assert len(res) == len(fix["blocks"])
    worst = 0.0
    bad = []
    for sc, (rc, out, st, log), ref in zip(blocks, res, fix["blocks"]):
        tag = f"block {ref['block']} ({ref['n_cams']} cameras)"
        try:
            assert sc["n_cams"] == ref["n_cams"] and sc["n_obs"] == ref["n_obs"], tag
            assert rc == ref["status"], tag
            assert st["itt"] == ref["itt"] and st["passes_done"] == ref["passes_done"], tag
            assert np.array_equal(st["trace_nrobs"], ref["trace_nrobs"]), (tag, "nrobs")
            # with the PCG, a converged LM iteration (sig0 moved < 1e-6) can
            # stop its PCG several iterations apart under rounding alone (the
            # oracle against itself: tests/golden/spread_c4_block6.json)
            worst = max(worst, check_trajectory(st["trace_sig0"], ref["trace_sig0"], st["trace_pcg"],
                                                ref["trace_pcg"], st["trace_accept"], ref["trace_accept"],
                                                st["trace_nrobs"], ref["trace_nrobs"],
                                                converged_rel=1e-6 if fix["flags"]["solver"] == 0 else None))
            for p in range(ref["passes_done"]):
                assert abs(st["sigma_out"][p] - ref["sigma_out"][p]) <= 1e-6 * ref["sigma_out"][p], tag
                for key in ("redundancy", "nobs", "nunknowns"):
                    assert st[key][p] == ref[key][p], (tag, key)
            assert make_golden.digest(out["obs_kept"]) == ref["obs_kept_sha256"], tag
            assert make_golden.digest(out["pt_kept"]) == ref["pt_kept_sha256"], tag
            # a PCG stopped apart on a converged iteration moves the solution
            # along its weakly determined directions: the oracle against
            # itself moves the focal lengths by 2.5e-6 and points by up to
            # 1.6e-4 of the scene's extent (spread_c4_block6.json)
            apart = np.abs(np.asarray(st["trace_pcg"]) - np.asarray(ref["trace_pcg"])).max(initial=0) > 1
            ext = float(np.abs(out["pts"]).max())
            assert np.allclose(out["cam_f"], ref["cam_f"], rtol=1e-5 if apart else 1e-6), tag
            assert np.allclose(out["pts"].astype(np.float64).mean(axis=0), ref["pts_mean"], rtol=1e-5,
                               atol=1e-3 * ext if apart else 1e-6), tag
            assert [l.split()[-4:] for l in log] == [l.split()[-4:] for l in ref["log"]], tag
        except AssertionError as e:
            n = min(len(st["trace_sig0"]), len(ref["trace_sig0"]))
            rel = np.abs(np.asarray(st["trace_sig0"][:n]) - ref["trace_sig0"][:n]) / np.asarray(ref["trace_sig0"][:n])
            print(f"[c4] {tag}: MISMATCH {str(e)[:200]}; itt {st['itt']} vs {ref['itt']}; nrobs "
                  f"{st['trace_nrobs'].tolist()} vs {ref['trace_nrobs']}; sig0 rel {rel.tolist()}", flush=True)
            bad.append(ref["block"])
    print(f"[c4] {len(res) - len(bad)} of {len(res)} blocks match; largest per-iteration sig0 difference of "
          f"those {worst:.3g}", flush=True)
    assert not bad, bad
