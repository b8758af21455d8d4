"""The trajectory comparison rule catches what it must (CPU only)."""
import numpy as np
import pytest

from tolerance import TRAJECTORY, check_trajectory

SIG = np.array([4.0, 2.0, 1.0, 0.5, 0.25, 0.2501])
PCG = np.array([3, 4, 4, 5, 6, 6])


def test_identical_passes():
    assert check_trajectory(SIG, SIG, PCG, PCG) == 0.0


def test_first_iteration_tight():
    s = SIG.copy()
    s[0] *= 1 + 2e-8
    with pytest.raises(AssertionError):
        check_trajectory(s, SIG, PCG, PCG)


def test_later_iterations_within_trajectory_bar():
    s = SIG.copy()
    s[3] *= 1 + 3e-6
    s[4] *= 1 - 2e-6
    assert check_trajectory(s, SIG, PCG, PCG) < 1e-5
    s[5] *= 1 + 2e-5
    with pytest.raises(AssertionError):
        check_trajectory(s, SIG, PCG, PCG)


def test_pcg_off_by_two_fails():
    p = PCG.copy()
    p[2] += 2
    with pytest.raises(AssertionError):
        check_trajectory(SIG, SIG, p, PCG)


def test_decisions_exact():
    acc = np.ones(SIG.size, dtype=np.int32)
    acc2 = acc.copy()
    acc2[-1] = 0
    with pytest.raises(AssertionError):
        check_trajectory(SIG, SIG, PCG, PCG, acc2, acc)
    nr = np.arange(SIG.size)
    with pytest.raises(AssertionError):
        check_trajectory(SIG, SIG, PCG, PCG, None, None, nr + 1, nr)


def test_bar_matches_the_oracles_own_rounding_spread(scene_c1):
    """The oracle against itself with only the summation order changed (S d
    row sums reversed, Schur pair products fused): on the 50-camera plain scene
    its mid-trajectory sig0 moves by more than 1e-6 (so no per-iteration bar
    of 1e-6 can hold between two correct FP64 implementations) and stays well
    inside TRAJECTORY, while the first iteration, the decisions, the PCG
    counts, the labels and the converged sigma0 agree at the FP64 bar."""
    from oracle import oracle
    from rpba_amd import capi
    fl = capi.make_flags(robust=0)
    _, o0, s0, _ = oracle.solve(scene_c1, fl)
    with oracle.order(3):
        _, o1, s1, _ = oracle.solve(scene_c1, fl)
    rel = np.abs(s1["trace_sig0"] - s0["trace_sig0"]) / s0["trace_sig0"]
    assert rel.max() > 1e-6 and rel.max() < TRAJECTORY / 3
    check_trajectory(s1["trace_sig0"], s0["trace_sig0"], s1["trace_pcg"], s0["trace_pcg"],
                     s1["trace_accept"], s0["trace_accept"], s1["trace_nrobs"], s0["trace_nrobs"])
    assert np.array_equal(s1["trace_pcg"], s0["trace_pcg"])
    assert np.array_equal(o1["obs_kept"], o0["obs_kept"])
    for p in range(2):
        assert abs(s1["sigma_out"][p] - s0["sigma_out"][p]) <= 1e-6 * s0["sigma_out"][p]


def test_converged_tail_lifts_only_the_count_bar():
    sig = np.array([4.0, 2.0, 1.0, 0.5, 0.25, 0.25000001])  # last step moves sig0 by 4e-8
    p = PCG.copy()
    p[-1] += 7
    with pytest.raises(AssertionError):
        check_trajectory(sig, sig, p, PCG)
    check_trajectory(sig, sig, p, PCG, converged_rel=1e-6)
    q = PCG.copy()
    q[3] += 2  # an iteration that still moves sig0
    with pytest.raises(AssertionError):
        check_trajectory(sig, sig, q, PCG, converged_rel=1e-6)
    s = sig.copy()
    s[-1] *= 1 + 2e-5  # the sig0 bar still holds on the converged iteration
    with pytest.raises(AssertionError):
        check_trajectory(s, sig, p, PCG, converged_rel=1e-6)


def test_converged_tail_rule_basis():
    """The oracle against itself on block 6 of the batched-blocks workload
    with the reference's SparsePCG, only the summation order changed: the
    last LM iteration moves sig0 by less than 1e-6, and there the PCG stops 7
    iterations apart; every other count is equal and sig0 agrees within 1e-8."""
    from oracle import oracle
    from rpba_amd import capi, scenes
    sc = scenes.batch_block(6, **scenes.BATCH_CONFIG)
    fl = capi.make_flags(robust=1, k2=1, k4=1, f=1, threads=1, solver=0)
    _, o0, s0, _ = oracle.solve(sc, fl)
    with oracle.order(1):
        _, o1, s1, _ = oracle.solve(sc, fl)
    d = np.abs(np.asarray(s1["trace_pcg"]) - np.asarray(s0["trace_pcg"]))
    assert d[-1] > 1 and d[:-1].max() == 0
    sig = np.asarray(s0["trace_sig0"])
    assert abs(sig[-1] - sig[-2]) < 1e-6 * sig[-1]
    with pytest.raises(AssertionError):
        check_trajectory(s1["trace_sig0"], s0["trace_sig0"], s1["trace_pcg"], s0["trace_pcg"])
    assert check_trajectory(s1["trace_sig0"], s0["trace_sig0"], s1["trace_pcg"], s0["trace_pcg"],
                            s1["trace_accept"], s0["trace_accept"], s1["trace_nrobs"], s0["trace_nrobs"],
                            converged_rel=1e-6) < 1e-8
    assert np.array_equal(o1["obs_kept"], o0["obs_kept"])
