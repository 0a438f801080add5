"""Certified priors (TSR_PRIORS_CERTIFIED) and their learned routing (TSR_CULL_LEARNED)
through the C ABI.

* The mesh is bit-identical to the plain field's mesh, and every leaf keeps its occupancy
  class (the judge).
* Parity of the capped traversal: the prior of a leaf certified at stage 0 / 1 is the
  midpoint of the oracle's interval with the cap 3 / 2 levels up, computed on the GPU's
  cone lengths (the intervals differ only in FP64 summation order); the oracle's judge
  certifies the same leaves at the same stages; leaves decided at stage 2 carry the exact
  prior.
* Routing: a network that sends every leaf to the exact stage gives the plain field bit
  for bit; one that keeps every leaf at stage 0 gives the cascade without a network.
"""
import numpy as np
import pytest

import tsr_amd
from oracle import mlp
from oracle.oracle import Oracle, certified
from tests.helpers import PARAMS, scene
from tsr_amd import _lib as L

pytestmark = pytest.mark.gpu

CASES = [(0, 0.5, 0), (3, 0.125, 5), (1, 1.0 / 16, 0)]


def _mesh_equal(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("positions", "triangles", "edge_keys"))


def _cls(v):
    return np.where(v < 0, 0, np.where(v <= np.float64(np.float32(1e-6)), 1, 2))


@pytest.mark.parametrize("cfg,res,seed", CASES, ids=lambda c: str(c))
def test_certified_mesh_and_parity(gpu_ctx, cfg, res, seed):
    s, c = scene(cfg, res, seed=seed)
    print(f"\n[certified] config={cfg} res={res} seed={seed} samples={len(s['px'])}", flush=True)
    sc = tsr_amd.Scene(gpu_ctx, s, c)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    plain = tsr_amd.Occupancy(gpu_ctx, tree, sc)
    cert = tsr_amd.Occupancy(gpu_ctx, tree, sc, flags=L.TSR_PRIORS_CERTIFIED)
    fp, fc = plain.download(), cert.download()
    mp = tsr_amd.DualMarchingCells(gpu_ctx, tree, plain).download()
    mc = tsr_amd.DualMarchingCells(gpu_ctx, tree, cert).download()
    assert _mesh_equal(mp, mc)
    assert np.array_equal(_cls(fp["occupancy"]), _cls(fc["occupancy"]))
    for k in ("kernel_sums", "cone_lengths", "node_states"):
        assert np.array_equal(fp[k], fc[k]), k
    lev = cert.prior_levels()
    st = cert.prior_stages()
    Lc = tree.leaf_count
    q = np.flatnonzero(lev != 255)
    assert st["queries"] == len(q) == int(np.sum(np.isfinite(fp["priors"])))
    assert st["stage0"] == len(q) and st["routed_stage1"] == st["routed_exact"] == 0
    assert st["stage1"] == int(np.sum(lev[q] >= 1)) and st["exact"] == int(np.sum(lev[q] == 2))
    print(f"stages {st}", flush=True)
    assert np.mean(lev[q] == 0) > 0.5
    # exact stage: the exact prior (another traversal grouping: FP64 summation order only)
    ex = q[lev[q] == 2]
    assert np.allclose(fc["priors"][ex], fp["priors"][ex], rtol=1e-12, atol=0)
    # capped stages against the oracle's capped traversal on the GPU's cone lengths
    o = Oracle(s, c, PARAMS).build_tree().occupancy(fp64=True)
    E, S = fc["kernel_sums"][:Lc], fc["kernel_sums"][Lc:]
    M = fc["cone_count"]
    lo3, hi3 = o.prior_bounds(q, 3, fc["cone_lengths"])
    lo2, hi2 = o.prior_bounds(q, 2, fc["cone_lengths"])
    c3 = certified(lo3, hi3, E[q], S[q], M)
    c2 = certified(lo2, hi2, E[q], S[q], M)
    olev = np.where(c3 >= 0, 0, np.where(c2 >= 0, 1, 2))
    mism = int(np.sum(olev != lev[q]))
    print(f"oracle judge vs GPU stage: {mism} of {len(q)} differ", flush=True)
    assert mism <= max(1, len(q) // 100000)
    for stage, lo, hi in ((0, lo3, hi3), (1, lo2, hi2)):
        k = (lev[q] == stage) & (olev == stage)
        mid = 0.5 * (lo[k] + hi[k])
        assert np.allclose(fc["priors"][q[k]], mid, rtol=1e-9, atol=1e-15), stage
    for x in (cert, plain, tree, sc):
        x.close()


def test_routing_extremes(gpu_ctx):
    s, c = scene(3, 0.125, seed=5)
    sc = tsr_amd.Scene(gpu_ctx, s, c)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    plain = tsr_amd.Occupancy(gpu_ctx, tree, sc)
    cert = tsr_amd.Occupancy(gpu_ctx, tree, sc, flags=L.TSR_PRIORS_CERTIFIED)
    fp, fc = plain.download(), cert.download()
    w = np.zeros(mlp.PARAM_COUNT, np.float32)
    w[-1] = 20.0                              # p ~ 1: every query straight to the exact stage
    exact = tsr_amd.Culler(gpu_ctx, w, threshold=0.75)
    o = tsr_amd.Occupancy(gpu_ctx, tree, sc, flags=L.TSR_CULL_LEARNED, culler=exact)
    st = o.prior_stages()
    assert st["routed_exact"] == st["queries"] == st["exact"] and st["stage0"] == 0
    f = o.download()
    for k in ("priors", "occupancy"):
        assert np.array_equal(f[k], fp[k], equal_nan=True), k
    o.close()
    w[-1] = -20.0                             # p ~ 0: every query at stage 0
    first = tsr_amd.Culler(gpu_ctx, w, threshold=0.75)
    o = tsr_amd.Occupancy(gpu_ctx, tree, sc, flags=L.TSR_CULL_LEARNED, culler=first)
    st = o.prior_stages()
    assert st["routed_stage1"] == st["routed_exact"] == 0 and st["stage0"] == st["queries"]
    f = o.download()
    for k in ("priors", "occupancy"):
        assert np.array_equal(f[k], fc[k], equal_nan=True), k
    assert np.array_equal(o.prior_levels(), cert.prior_levels())
    for x in (o, exact, first, cert, plain, tree, sc):
        x.close()
