"""The trajectory comparison rule catches what it must (CPU only)."""
import numpy as np
import pytest

from tolerance import check_trajectory

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
