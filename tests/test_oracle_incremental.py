"""The oracle's incremental updates (Occupancy::addSamples / eraseSamples) and the tree's
sample erasure (Nodes::eraseSamples), CPU only.

* Erasing a set of samples and adding it back gives the original field; erasing every
  sample leaves zero sums and M = 0; erasing A then B equals erasing A and B at once
  (the passes are linear in the chosen samples, so a wrong sign, a link of the wrong
  sample or a double count fails one of these).
* Node flags are only ever set (Occupancy.cpp:233): an erase leaves them as they were.
* erase_samples_in_nodes against a literal restatement of the reference's per-sample loop.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, erase_samples_in_nodes
from tests.helpers import PARAMS, scene


@pytest.fixture(scope="module")
def sphere():
    s, c = scene(0, 0.25)
    return s, c


def _field(o):
    f = o.field()
    return {k: np.array(v, copy=True) if isinstance(v, np.ndarray) else v for k, v in f.items()}


def _close(a, b, scale):
    return np.all(np.abs(a - b) <= 1e-9 * scale + 1e-12)


def test_erase_then_add_back_restores_the_field(sphere):
    s, c = sphere
    o = Oracle(s, c, PARAMS).build_tree().occupancy(fp64=True)
    f0 = _field(o)
    rng = np.random.default_rng(1)
    doomed = rng.choice(len(s["px"]), len(s["px"]) // 7, replace=False)
    o.update(True, doomed)
    f1 = _field(o)
    mag = np.maximum(np.abs(f0["kernel_sums"]), 1.0)
    assert not _close(f1["kernel_sums"], f0["kernel_sums"], mag)        # the erase changed something
    assert np.array_equal(f1["node_states"], f0["node_states"])          # flags are never cleared
    assert f1["cone_count"] < f0["cone_count"]
    o.update(False, doomed)
    f2 = _field(o)
    for k in ("kernel_sums", "cone_lengths"):
        assert _close(f2[k], f0[k], np.maximum(np.abs(f0[k]), 1.0)), k
    assert f2["cone_count"] == pytest.approx(f0["cone_count"], rel=1e-12)
    assert np.allclose(f2["priors"], f0["priors"], rtol=1e-9, atol=0, equal_nan=True)


def test_erasing_every_sample_leaves_nothing(sphere):
    s, c = sphere
    o = Oracle(s, c, PARAMS).build_tree().occupancy(fp64=True)
    f0 = _field(o)
    o.update(True, np.arange(len(s["px"])))
    f = _field(o)
    for k in ("kernel_sums", "cone_lengths"):
        assert np.all(np.abs(f[k]) <= 1e-9 * np.maximum(np.abs(f0[k]), 1.0)), k
    assert abs(f["cone_count"]) <= 1e-9 * f0["cone_count"]


def test_erasures_compose(sphere):
    s, c = sphere
    n = len(s["px"])
    rng = np.random.default_rng(2)
    perm = rng.permutation(n)
    a, b = perm[:n // 9], perm[n // 9:n // 4]
    o1 = Oracle(s, c, PARAMS).build_tree().occupancy(fp64=True)
    o1.update(True, a).update(True, b)
    o2 = Oracle(s, c, PARAMS).build_tree().occupancy(fp64=True)
    o2.update(True, np.concatenate([a, b]))
    f1, f2 = _field(o1), _field(o2)
    for k in ("kernel_sums", "cone_lengths"):
        assert _close(f1[k], f2[k], np.maximum(np.abs(f2[k]), 1.0)), k
    assert f1["cone_count"] == pytest.approx(f2["cone_count"], rel=1e-12)


def _erase_literal(counts, starts, inliers):
    # Scene.cpp:274-287 offsets, then Nodes.cpp:651-668 per node, as written
    n = len(inliers)
    offsets = [0] * (n + 1)
    for i in range(n):
        offsets[i + 1] = offsets[i] + (0 if inliers[i] else 1)
    c2, s2 = list(counts), list(starts)
    for nd in range(len(counts)):
        start, end = int(starts[nd]), int(starts[nd]) + int(counts[nd])
        deleted = sum(1 for i in range(start, end) if offsets[i] != offsets[i + 1])
        c2[nd] = int(counts[nd]) - deleted
        s2[nd] = start - offsets[start]
    return np.array(c2, np.uint32), np.array(s2, np.uint32)


def test_erase_in_nodes_matches_the_literal_loop(sphere):
    s, c = sphere
    t = Oracle(s, c, PARAMS).build_tree().tree()
    n = int(np.sum(t["sample_counts"]))
    rng = np.random.default_rng(3)
    inl = rng.random(n) > 0.2
    got = erase_samples_in_nodes(t["sample_counts"], t["sample_starts"], inl)
    ref = _erase_literal(t["sample_counts"], t["sample_starts"], inl)
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert int(got[0].sum()) == int(inl.sum())
    # the kept samples of a node stay one block, in node order
    nz = got[0] > 0
    order = np.argsort(got[1][nz], kind="stable")
    starts, counts = got[1][nz][order], got[0][nz][order]
    assert np.array_equal(starts[1:], (starts + counts)[:-1])
