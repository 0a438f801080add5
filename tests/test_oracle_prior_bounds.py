"""The oracle's depth-capped prior bounds and the certification judge (CPU only).

* Uncapped, the bound traversal (whole-node subtree sums) gives the prior of the
  reference's leaf-by-leaf sphere query (Occupancy.cpp:530-568) to rounding: a wrong
  whole-node rule (depth limit, inner radius) would add or drop leaves.
* Capped, every interval holds the exact prior, and the intervals nest as the cap moves
  down the tree (opening a bounded node can only tighten it); swapped or dropped
  upper-bound terms break one of the two.
* The judge: known answers of the occupancy class on hand-made (E, S, M, p), and on two
  scenes every certified class equals the class of the exact prior.
"""
import numpy as np
import pytest

from oracle.oracle import EPS_F, Oracle, certified, occupancy_class
from tests.helpers import PARAMS, scene


@pytest.fixture(scope="module", params=[(3, 0.125, 5), (0, 0.5, 0)], ids=["cfg3", "cfg0"])
def field(request):
    cfg, res, seed = request.param
    s, c = scene(cfg, res, seed=seed)
    o = Oracle(s, c, PARAMS).build_tree().occupancy(fp64=True)
    f = o.field()
    L = len(f["occupancy"])
    E, S = f["kernel_sums"][:L], f["kernel_sums"][L:]
    q = np.flatnonzero(E + S >= PARAMS[1])
    return o, f, q, E[q], S[q]


def test_uncapped_bound_is_the_reference_prior(field):
    o, f, q, _, _ = field
    lo, hi = o.prior_bounds(q, -4)          # cap below the depth limit: nothing is bounded
    p = f["priors"][q]
    assert np.array_equal(lo, hi)
    assert np.allclose(lo, p, rtol=1e-12, atol=0)


def test_capped_intervals_hold_the_prior_and_nest(field):
    o, f, q, _, _ = field
    p = f["priors"][q]
    lo3, hi3 = o.prior_bounds(q, 3)
    lo2, hi2 = o.prior_bounds(q, 2)
    lo1, hi1 = o.prior_bounds(q, 1)
    tol = 1e-12
    for lo, hi in ((lo3, hi3), (lo2, hi2), (lo1, hi1)):
        assert np.all(lo <= p * (1 + tol)) and np.all(hi >= p * (1 - tol))
        assert np.all((lo >= 0) & (hi <= 1) & (lo <= hi))
    assert np.all(lo3 <= lo2 * (1 + tol)) and np.all(lo2 <= lo1 * (1 + tol))
    assert np.all(hi3 >= hi2 * (1 - tol)) and np.all(hi2 >= hi1 * (1 - tol))
    # the cap does bound something on these scenes
    assert np.mean(hi3 - lo3 > 1e-9) > 0.1


def test_occupancy_class_known_answers():
    M = 10.0
    # pE = p E / M, pS = (1 - p) S / M
    assert occupancy_class(0.5, 100.0, 50.0, M) == 0              # pE 5 >= pS 2.5
    assert occupancy_class(0.5, 50.0, 50.0, M) == -1             # pE == pS: the sign boundary
    assert occupancy_class(0.1, 10.0, 100.0, M) == 2             # v = 9 - 0.1 > EPS
    v = 0.5e-6                                                    # 0 < v < EPS: class 1
    assert occupancy_class(0.0, 0.0, v * M, M) == 1
    assert occupancy_class(0.0, 0.0, EPS_F * M, M) == -1         # v at EPS
    assert occupancy_class(0.0, 0.0, 1e-20 * M, M) == -1         # v below 1e-18
    assert occupancy_class(np.nan, 1.0, 1.0, M) == -1


def test_certified_classes_are_exact(field):
    o, f, q, E, S = field
    M = f["cone_count"]
    truth = occupancy_class(f["priors"][q], E, S, M)
    for cap in (3, 2):
        lo, hi = o.prior_bounds(q, cap)
        c = certified(lo, hi, E, S, M)
        ok = c >= 0
        assert np.all(c[ok] == truth[ok])
        assert ok.mean() > 0.5
    # an interval that straddles the sign boundary is never certified
    p = f["priors"][q]
    assert np.all(certified(np.zeros_like(p), np.ones_like(p), E, S, M)[(E > 0) & (S > 0)] == -1)
