"""Pins of the oracle's LM control (LSA::Adjust, rbacore.cpp:333-616),
StoreResultsIteration (891-918), RobustDelete (922-995) and the output rule of
OutputLSAc (644-666), by known answers rather than by a second copy of the
code:

  * scripted sig0 sequences replace the computed cost values, and the
    expected accept decisions, iteration counts, iflag trace and Sigma0 lines
    are worked out by hand from the reference's rules (the rest of the
    numerics run for real);
  * per-iteration records of the solver state show which cameras, points and
    robust keys each step used (H4, H6);
  * RobustDelete runs on labels the test constructs;
  * each pin fails under the matching planted mutation of the oracle.

The scenes use f = 500, so lflflag holds and sig0frac = 1.01
(rbacore.cpp:213-220, 362-363); no image sizes are given, so ihw2 = 1."""
import math

import numpy as np
import pytest

from oracle import oracle
from rpba_amd import capi, scenes


def _tiny():
    return scenes.generate(None, seed=11, n_cams=6, n_pts=60, track="uniform2_8", outlier_frac=0.0)


def _robust_scene():
    return scenes.generate(None, seed=71, n_cams=12, n_pts=700, track="uniform2_8", outlier_frac=0.05)


def _run(sc, script, robust, rec_cap=0):
    h = oracle.Handle(sc, capi.make_flags(robust=robust))
    rc, st, log, rec = h.adjust(script=script, rec_cap=rec_cap)
    assert rc == 0
    return h, st, log, rec


# --- scripted sequences ----------------------------------------------------

def check_first_iteration_stored():
    """H5: the first trial of a pass is stored even when worse than the
    initial cost (rbacore.cpp:565).  Pass 0: initial 10, trial 12 is stored;
    10 < 12 * 1.01 ends the pass (iflag 1 -> 0).  Pass 1 starts from
    sig0old = 1e20 (rbacore.cpp:372): 7.0 stored, 6.98 stored and
    7.0 < 6.98 * 1.01 ends it."""
    _, st, log, _ = _run(_tiny(), [10.0, 12.0, 7.0, 6.98], robust=0)
    assert list(st["trace_accept"]) == [1, 1, 1]
    assert np.array_equal(st["trace_sig0"], [12.0, 7.0, 6.98])
    assert st["itt"] == 3 and st["passes_done"] == 2
    assert st["sigma_in"][0] == math.sqrt(10.0) and st["sigma_out"][0] == math.sqrt(12.0)
    assert st["sigma_out"][1] == math.sqrt(6.98)
    assert log[0].startswith("Sigma0: 3.16228 3.4641 ") and log[0].endswith("IT: 0  ITT: 1\n")
    assert log[1].startswith("Sigma0: 1e+10 2.64197 ") and log[1].endswith("IT: 1  ITT: 3\n")


def check_reject_and_stop():
    """A worse trial is not stored and, with the 1.01 fraction, ends the pass:
    pass 0: 8 stored (10 < 8.08 fails), 9 rejected (8 < 9.09: iflag 0);
    pass 1: 7.5, 7.4 stored (7.5 < 7.474 fails), 7.39 stored and ends it."""
    _, st, log, _ = _run(_tiny(), [10.0, 8.0, 9.0, 7.5, 7.4, 7.39], robust=0)
    assert list(st["trace_accept"]) == [1, 0, 1, 1, 1]
    assert st["itt"] == 5
    assert st["sigma_out"][0] == math.sqrt(8.0) and st["sigma_out"][1] == math.sqrt(7.39)
    assert log[0].endswith("IT: 0  ITT: 2\n") and log[1].endswith("IT: 1  ITT: 5\n")


def check_itt_cap():
    """The itt cap (rbacore.cpp:575-576): steadily improving trials never end
    a pass, so pass 0 runs while itt <= 200 (202 iterations, itt stays 201)
    and pass 1 stops after its first iteration (itt is already past 200)."""
    seq = [10.0] + [10.0 * 0.98 ** k for k in range(1, 210)]
    _, st, log, _ = _run(_tiny(), seq, robust=0)
    assert st["trace_accept"].size == 203 and st["trace_accept"].all()
    assert st["itt"] == 201
    assert log[0].endswith("IT: 0  ITT: 201\n") and log[1].endswith("IT: 1  ITT: 201\n")


# robust: iflag starts at 2 in both passes (rbacore.cpp:502-508).  Pass 0:
# 8 stored; 9 rejected (iflag 1); 7 and 6.9 stored; 7.5 rejected (iflag 0).
# Pass 1: 7.0 stored; 6.95 stored (iflag 1); 6.94 stored and ends it.
ROBUST_SCRIPT = [10.0, 8.0, 9.0, 7.0, 6.9, 7.5, 7.0, 6.95, 6.94]
ROBUST_ACCEPT = [1, 0, 1, 1, 0, 1, 1, 1]
ROBUST_IFLAG = [2, 1, 1, 1, 0, 2, 1, 0]


def check_linearisation_point():
    """H4: after a rejected step the next linearisation uses the rejected
    trial cameras (eRMnew / ePCnew / camerasnew are not restored) with the
    stored points, also across the pass boundary (rbacore.cpp:550-552,
    565-570, 1161, 1224-1265)."""
    _, st, _, rec = _run(_robust_scene(), ROBUST_SCRIPT, robust=1, rec_cap=32)
    assert list(st["trace_accept"]) == ROBUST_ACCEPT
    assert list(rec["iflag"]) == ROBUST_IFLAG
    n = len(ROBUST_ACCEPT)
    assert rec["lin_cam"].size == n
    for j in range(n - 1):
        assert rec["lin_cam"][j + 1] == rec["trial_cam"][j], j
        if ROBUST_ACCEPT[j]:
            assert rec["stored_cam"][j] == rec["trial_cam"][j], j
        else:
            assert rec["lin_cam"][j + 1] != rec["stored_cam"][j], j
        if rec["pass"][j + 1] == rec["pass"][j]:  # RobustDelete compacts the points between the passes
            assert rec["lin_pts"][j + 1] == rec["stored_pts"][j], j
            if not ROBUST_ACCEPT[j]:
                assert rec["lin_pts"][j + 1] != rec["trial_pts"][j], j


def check_labels_from_previous_trial():
    """H6: Reweight2D labels from the residuals of the last trial cost
    evaluation, accepted or not (rbacore.cpp:522, 552, 863), and ISflagnew
    takes ISflag only when a step is stored (rbacore.cpp:911)."""
    _, st, _, rec = _run(_robust_scene(), ROBUST_SCRIPT, robust=1, rec_cap=32)
    assert list(st["trace_accept"]) == ROBUST_ACCEPT
    passes = rec["pass"]
    assert rec["key_used"][0] == rec["init_key"]
    prev_new = 0.0
    for j in range(len(ROBUST_ACCEPT)):
        if passes[j] == 0:
            if j > 0:
                assert rec["key_used"][j] == rec["trial_key"][j - 1], j
            expect = rec["flag_fp"][j] if ROBUST_ACCEPT[j] else prev_new
            assert rec["flagnew_fp"][j] == expect, j
            prev_new = rec["flagnew_fp"][j]
        else:
            assert rec["key_used"][j] == 0.0  # no reweighting in pass 1
    # the scene is robust enough to label observations, and a rejected step
    # at iflag 1 had labels different from the stored ones
    assert rec["flag_fp"][4] != 0.0 and rec["flag_fp"][4] != rec["flagnew_fp"][3]


# --- RobustDelete on constructed labels ------------------------------------

def check_point_drop_rule():
    """RobustDelete (rbacore.cpp:922-995): labelled observations leave their
    track; a point with a label and fewer than minimagenrdel = 3 observations
    left is dropped whole (rbacore.cpp:73-76, 945-951); a point without labels
    is untouched."""
    sc = scenes.generate(None, seed=73, n_cams=12, n_pts=400, track="uniform2_8", outlier_frac=0.0)
    h = oracle.Handle(sc, capi.make_flags(robust=1))
    h.setup_pass0()
    order = np.argsort(sc["obs_pt"], kind="stable")  # track entries in point-major order
    t = np.bincount(sc["obs_pt"], minlength=sc["n_pts"])
    off = np.concatenate([[0], np.cumsum(t)])
    rng = np.random.default_rng(5)
    flags = np.zeros(sc["n_obs"], np.uint8)
    nf = np.zeros(sc["n_pts"], int)
    for p in range(sc["n_pts"]):
        # label 0, 1 or 2 entries; covers t - nf = 1, 2 (dropped) and 3+ (kept)
        k = int(rng.integers(0, 3))
        if k:
            pick = rng.choice(t[p], size=min(k, t[p]), replace=False)
            flags[off[p] + pick] = 1
            nf[p] = pick.size
    kept_pts, kept_obs = h.delete_with_flags(flags)
    expect_pts = np.array([p for p in range(sc["n_pts"]) if nf[p] == 0 or t[p] - nf[p] >= 3])
    assert np.array_equal(kept_pts, expect_pts)
    keep_p = np.zeros(sc["n_pts"], bool)
    keep_p[expect_pts] = True
    obs_keep = keep_p[sc["obs_pt"][order]] & (flags == 0)
    assert np.array_equal(kept_obs, order[obs_keep])
    # the construction reaches both sides of the rule
    assert ((nf > 0) & (t - nf == 2)).any() and ((nf > 0) & (t - nf == 3)).any()


def check_planted_outliers_removed():
    """End to end: with no other outliers, a 3-observation point with one
    planted 40 px error loses the observation and falls below 3, so it is
    dropped; points without planted errors keep every observation.  The error
    is vertical: every camera sits in the z = 0 plane, so no move of the point
    can absorb it into the other two views."""
    sc = scenes.generate(None, seed=74, n_cams=12, n_pts=800, track="uniform2_8", outlier_frac=0.0)
    rng = np.random.default_rng(6)
    t = np.bincount(sc["obs_pt"], minlength=sc["n_pts"])
    three = rng.choice(np.nonzero(t == 3)[0], 25, replace=False)
    bad = np.zeros(sc["n_pts"], bool)
    bad[three] = True
    xy = sc["obs_xy"].copy()
    for p in three:
        q = np.nonzero(sc["obs_pt"] == p)[0][int(rng.integers(0, 3))]
        xy[q, 1] += 40.0
    sc = dict(sc, obs_xy=xy)
    rc, out, st, _ = oracle.solve(sc, capi.make_flags(robust=1))
    assert rc == 0
    assert not out["pt_kept"][three].any()
    assert not out["obs_kept"][np.isin(sc["obs_pt"], three)].any()
    good = ~bad
    assert out["pt_kept"][good].all()
    assert out["obs_kept"][good[sc["obs_pt"]]].all()


# --- output rule -----------------------------------------------------------

def check_output_intrinsics_from_last_trial():
    """H19: OutputLSAc writes R|t and X of the stored state but the
    intrinsics of camerasnew, the last trial, here a rejected one
    (rbacore.cpp:655-663)."""
    h, st, _, _ = _run(_tiny(), [10.0, 8.0, 9.0, 7.0, 7.5], robust=0)
    assert list(st["trace_accept"]) == [1, 0, 1, 0]
    stored, trial = h.intrinsics(0), h.intrinsics(1)
    out = h.output()
    assert np.array_equal(out["cam_f"], trial[:, 0].astype(np.float32))
    assert np.array_equal(out["cam_d"][:, 0], trial[:, 1].astype(np.float32))
    differ = trial[:, 0].astype(np.float32) != stored[:, 0].astype(np.float32)
    assert differ.sum() >= out["cam_f"].size // 2


PINS = {
    "first_iteration_stored": check_first_iteration_stored,
    "reject_and_stop": check_reject_and_stop,
    "itt_cap": check_itt_cap,
    "linearisation_point": check_linearisation_point,
    "labels_from_previous_trial": check_labels_from_previous_trial,
    "point_drop_rule": check_point_drop_rule,
    "planted_outliers_removed": check_planted_outliers_removed,
    "output_intrinsics_from_last_trial": check_output_intrinsics_from_last_trial,
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(name):
    PINS[name]()


# mutation (see the head of oracle/rpba_oracle.c) -> the pin it must break
MUTATIONS = [
    (1, "linearisation_point"),
    (2, "first_iteration_stored"),
    (3, "labels_from_previous_trial"),
    (4, "point_drop_rule"),
    (5, "itt_cap"),
    (6, "output_intrinsics_from_last_trial"),
]


@pytest.mark.parametrize("mut,name", MUTATIONS)
def test_planted_mutation_fails_its_pin(mut, name):
    with oracle.mutation(mut):
        with pytest.raises(AssertionError):
            PINS[name]()
