"""The sampling searches test a child box against the sample plane through per-axis
extremes of the plane products (tree_build.hip, openTerms / intersectChild) instead of
its eight corner sums (Nodes.cpp:490-517 intersect).  That is exact because FP64
round-to-nearest addition is monotone: the extremes added in the corner order give
exactly the largest (smallest) of the eight computed corner sums.  This checks that
claim, and the decision it feeds, on random and near-cancelling FP64 terms."""
import itertools

import numpy as np


def corner_sums(t):
    # t[a][j]: (plane_j - p_a) * n_a for the child's two planes j on axis a
    return np.array([(t[0][i] + t[1][j]) + t[2][k] for i, j, k in itertools.product((0, 1), repeat=3)])


def any_corner_on_side(t, positive):
    return any((v > 0.0) == positive for v in corner_sums(t))


def extreme_on_side(t, positive):
    ext = [max(a) if positive else min(a) for a in t]
    return ((ext[0] + ext[1]) + ext[2] > 0.0) == positive


def test_extreme_sum_is_extreme_corner_sum():
    rng = np.random.default_rng(7)
    for _ in range(20000):
        scale = 10.0 ** rng.integers(-12, 4, size=3)
        t = [list(rng.standard_normal(2) * s) for s in scale]
        sums = corner_sums(t)
        assert (max(t[0]) + max(t[1])) + max(t[2]) == sums.max()
        assert (min(t[0]) + min(t[1])) + min(t[2]) == sums.min()


def test_side_decision_matches_corner_loop_near_cancellation():
    rng = np.random.default_rng(11)
    for _ in range(20000):
        x = rng.standard_normal()
        y = rng.standard_normal()
        # terms that nearly cancel, so the corner sums straddle 0 by a few ulp
        t = [[x, np.nextafter(x, np.inf)], [y, np.nextafter(y, -np.inf)],
             [-(x + y), np.nextafter(-(x + y), rng.choice([np.inf, -np.inf]))]]
        for positive in (True, False):
            assert extreme_on_side(t, positive) == any_corner_on_side(t, positive)
