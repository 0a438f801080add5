"""Incremental updates through the C ABI: Occupancy::addSamples / eraseSamples
(Occupancy.cpp:129-136,138-248) and Tree::eraseSamples with Samples::compact
(Nodes.cpp:644-671, Scene.cpp:270-296).

* Erase: kernel sums and cone lengths within 1e-5 of the pre-erase magnitude of the
  oracle's erase (subtracting FP32 sums leaves their rounding, as the reference's float
  sums do), M within 1e-9, node flags unchanged, and the occupancy class of every leaf
  equal to the oracle's except where the oracle's own value sits within 1e-4 of a class
  boundary (relative to pS + pE).
* Adding the samples back gives the original field within the same bound.
* The tree erase equals the restatement of Nodes::eraseSamples, the kept samples are the
  original node-ordered samples in order, the input-to-node map follows, and a fresh field
  on the erased tree has the sampleness sums of the incremental erase (within 1e-5 of the
  pre-erase magnitude).
"""
import numpy as np
import pytest

import tsr_amd
from oracle.oracle import Oracle, erase_samples_in_nodes
from tests.helpers import PARAMS, scene
from tsr_amd import _lib as L

pytestmark = pytest.mark.gpu

EPS = np.float64(np.float32(1e-6))


def _cls(v):
    return np.where(v < 0, 0, np.where(v <= EPS, 1, 2))


def _near_boundary(f):
    Lc = len(f["occupancy"])
    E, S = f["kernel_sums"][:Lc], f["kernel_sums"][Lc:]
    p, M = f["priors"], f["cone_count"]
    pE, pS = p * E / M, (1 - p) * S / M
    v = pS - pE
    tol = 1e-4 * (pE + pS)
    return (np.abs(v) <= tol) | (np.abs(v - EPS) <= tol)


@pytest.mark.parametrize("cfg,res,seed", [(3, 0.125, 5), (0, 0.5, 0)], ids=lambda c: str(c))
def test_erase_and_add_back_match_the_oracle(gpu_ctx, cfg, res, seed):
    s, c = scene(cfg, res, seed=seed)
    print(f"\n[incremental] config={cfg} res={res} seed={seed} samples={len(s['px'])}", flush=True)
    sc = tsr_amd.Scene(gpu_ctx, s, c)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    occ = tsr_amd.Occupancy(gpu_ctx, tree, sc, flags=L.TSR_PRIORS_ALL_LEAVES)
    f0 = occ.download()
    t = tree.download()
    Lc = tree.leaf_count
    n = tree.sample_count
    rng = np.random.default_rng(seed + 11)
    doomed = np.sort(rng.choice(n, n // 8, replace=False)).astype(np.uint32)   # node order
    inv = np.empty(n, np.uint32)
    inv[t["sample_target"]] = np.arange(n, dtype=np.uint32)                   # node -> original
    o = Oracle(s, c, PARAMS).build_tree().occupancy(fp64=True)
    of0 = o.field()
    mag = {k: np.maximum(np.abs(np.asarray(of0[k], np.float64)), 1e-3) for k in ("kernel_sums", "cone_lengths")}
    # erase
    occ.erase_samples(tree, sc, doomed)
    f1 = occ.download()
    of1 = o.update(True, inv[doomed]).field()
    for k in ("kernel_sums", "cone_lengths"):
        err = np.abs(np.asarray(f1[k], np.float64) - of1[k]) / mag[k]
        print(f"erase {k}: max err / pre-erase magnitude {err.max():.2e}", flush=True)
        assert err.max() <= 1e-5, k
    assert f1["cone_count"] == pytest.approx(of1["cone_count"], rel=1e-9)
    assert np.array_equal(f1["node_states"], f0["node_states"])
    diff = (_cls(f1["occupancy"]) != _cls(of1["occupancy"])) & ~_near_boundary(of1)
    print(f"erase: {int(np.sum(_cls(f1['occupancy']) != _cls(of1['occupancy'])))} class differences, "
          f"{int(diff.sum())} away from a boundary", flush=True)
    assert diff.sum() == 0
    m = tsr_amd.DualMarchingCells(gpu_ctx, tree, occ).download()
    assert len(m["triangles"]) > 0
    # add back
    occ.add_samples(tree, sc, doomed)
    f2 = occ.download()
    for k in ("kernel_sums", "cone_lengths"):
        err = np.abs(np.asarray(f2[k], np.float64) - np.asarray(f0[k], np.float64)) / mag[k]
        assert err.max() <= 1e-5, k
    assert f2["cone_count"] == pytest.approx(f0["cone_count"], rel=1e-9)
    assert np.array_equal(f2["node_states"], f0["node_states"])
    occ.close()
    tree.close()
    sc.close()


def test_tree_erase_samples(gpu_ctx):
    s, c = scene(3, 0.125, seed=5)
    sc = tsr_amd.Scene(gpu_ctx, s, c)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    occ = tsr_amd.Occupancy(gpu_ctx, tree, sc)
    t0 = tree.download()
    s0 = tree.download_samples()
    f0 = occ.download()
    n = tree.sample_count
    rng = np.random.default_rng(4)
    inl = rng.random(n) > 0.15
    doomed = np.flatnonzero(~inl).astype(np.uint32)
    occ.erase_samples(tree, sc, doomed)          # Scene::eraseSamples: the field first
    f1 = occ.download()
    tree.erase_samples(inl)
    t1 = tree.download()
    counts, starts = erase_samples_in_nodes(t0["sample_counts"], t0["sample_starts"], inl)
    assert np.array_equal(t1["sample_counts"], counts) and np.array_equal(t1["sample_starts"], starts)
    for k in ("children", "parents", "node_to_leaf", "leaf_nodes", "cells"):
        assert np.array_equal(t1[k], t0[k]), k
    s1 = tree.download_samples()
    assert tree.sample_count == int(inl.sum())
    for k in ("px", "py", "pz", "nx", "ny", "nz", "scale", "confidence"):
        assert np.array_equal(s1[k], s0[k][inl]), k
    mc = tree.max_cams
    assert np.array_equal(s1["parent_cameras"], s0["parent_cameras"].reshape(-1, mc)[inl].reshape(-1))
    new_index = np.cumsum(inl) - 1
    expect = np.where(inl[t0["sample_target"]], new_index[t0["sample_target"]], 0xFFFFFFFF).astype(np.uint32)
    assert np.array_equal(t1["sample_target"], expect)
    # a fresh field on the erased tree: the sampleness pass reads the remaining samples only
    fresh = tsr_amd.Occupancy(gpu_ctx, tree, sc)
    f2 = fresh.download()
    Lc = tree.leaf_count
    # (the incremental sums carry the FP32 rounding of the pre-erase sums they were subtracted from)
    mag = np.maximum(np.abs(np.asarray(f0["kernel_sums"][Lc:], np.float64)), 1e-3)
    err = np.abs(np.asarray(f2["kernel_sums"][Lc:], np.float64) - f1["kernel_sums"][Lc:]) / mag
    assert err.max() <= 1e-5
    m = tsr_amd.DualMarchingCells(gpu_ctx, tree, fresh).download()
    assert len(m["triangles"]) > 0
    for x in (fresh, occ, tree, sc):
        x.close()


def test_update_argument_checks(gpu_ctx):
    s, c = scene(0, 0.25)
    sc = tsr_amd.Scene(gpu_ctx, s, c)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    occ = tsr_amd.Occupancy(gpu_ctx, tree, sc)
    with pytest.raises(tsr_amd.TsrError) as e:
        occ.erase_samples(tree, sc, [tree.sample_count])        # index past the samples
    assert e.value.status_name == "TSR_ERR_INVALID_ARGUMENT"
    occ.add_samples(tree, sc, [])                              # nothing: priors recomputed, no change
    for x in (occ, tree, sc):
        x.close()
