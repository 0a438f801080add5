"""GPU parity: HIP tree / occupancy / DMC through the C ABI vs the CPU oracle.

Tree tables are compared bit for bit; kernel sums, cone lengths and priors within
1e-5 relative; occupancy within 1e-5 of (pS + pE); the DMC mesh (vertex keys,
triangles) bit for bit where no occupancy sign differs, positions within 1e-9
of the leaf size scale.
"""
import os
import sys

import numpy as np
import pytest

from tests.helpers import PARAMS, occupancy_rel_err, rel_err, scene

pytestmark = pytest.mark.gpu

CASES = [
    (0, 0.5),     # sphere, 64x48
    (0, 1.0),     # config 0 as specified (128x96)
    (1, 0.0625),  # thin wedge + sphere, 40x30 x 64 cameras, outliers
    (3, 0.125),   # random primitive union
]


def _oracle(samples, cameras):
    from oracle.oracle import Oracle
    o = Oracle(samples, cameras, PARAMS)
    o.build_tree().occupancy().dmc()
    return o


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"cfg{c[0]}x{c[1]}")
def case(request, gpu_ctx):
    from tsr_amd import DualMarchingCells, Occupancy, Scene, Tree
    from tsr_amd._lib import TSR_PRIORS_ALL_LEAVES
    cfg, res = request.param
    samples, cameras = scene(cfg, res)
    print(f"\n[case] config={cfg} res={res} samples={len(samples['px'])}", flush=True)
    orc = _oracle(samples, cameras)
    sc = Scene(gpu_ctx, samples, cameras)
    tree = Tree(gpu_ctx, sc)
    occ = Occupancy(gpu_ctx, tree, sc, flags=TSR_PRIORS_ALL_LEAVES)
    mesh = DualMarchingCells(gpu_ctx, tree, occ)
    return {"orc": orc, "tree": tree.download(), "field": occ.download(), "mesh": mesh.download(),
            "gpu_tree": tree, "gpu_occ": occ, "samples": samples, "scene": sc}


def test_tree_tables_bit_exact(case):
    t, o = case["tree"], case["orc"].tree()
    for k in ("children", "parents", "sample_counts", "sample_starts", "node_to_leaf", "leaf_nodes",
              "sample_target", "cells"):
        assert t[k].shape == o[k].shape, f"{k}: shape {t[k].shape} vs {o[k].shape}"
        bad = np.flatnonzero(t[k] != o[k])
        assert bad.size == 0, f"{k}: {bad.size} mismatches, first at {bad[:5]}"
    assert np.array_equal(t["leaf_scopes"], o["leaf_scopes"])


def test_samples_reordered(case):
    t = case["tree"]
    s = case["samples"]
    d = case["gpu_tree"].download_samples()
    inv = np.empty_like(t["sample_target"])
    inv[t["sample_target"]] = np.arange(len(inv), dtype=np.uint32)
    for k in ("px", "py", "pz", "nx", "scale"):
        assert np.array_equal(d[k], s[k][inv])
    orig = case["scene"].download()
    for k in ("px", "py", "pz", "nx", "scale"):
        assert np.array_equal(orig[k], s[k])


def test_field_parity(case):
    g, o = case["field"], case["orc"].field()
    L = len(o["occupancy"])
    assert np.array_equal(g["node_states"], o["node_states"]), "node state flags differ"
    for k in ("kernel_sums", "cone_lengths"):
        r = rel_err(g[k], o[k], floor=1e-30)
        assert np.max(r) <= 1e-5, f"{k}: max rel err {np.max(r):.3e}"
    finite = np.isfinite(o["priors"])
    assert np.array_equal(np.isfinite(g["priors"]), finite)
    r = rel_err(g["priors"][finite], o["priors"][finite], floor=1e-30)
    assert np.max(r, initial=0) <= 1e-5, f"priors: max rel err {np.max(r):.3e}"
    assert abs(g["cone_count"] - o["cone_count"]) <= 1e-9 * o["cone_count"]
    err, flips = occupancy_rel_err(g["occupancy"], o)
    print(f"occupancy: leaves={L} max rel err={err.max():.3e} p99={np.percentile(err, 99):.3e} flips={flips}")
    assert err.max() <= 1e-5
    assert flips == 0


def test_mesh_parity(case):
    g, o = case["mesh"], case["orc"].mesh()
    assert o["conflicts"] == 0
    assert np.array_equal(g["edge_keys"], o["edge_keys"])
    assert np.array_equal(g["triangles"], o["triangles"])
    scale = float(case["orc"].tree()["root_size"])
    assert np.max(np.abs(g["positions"] - o["positions"]), initial=0) <= 1e-6 * scale
    assert np.max(np.abs(g["scales"] - o["scales"]), initial=0) <= 1e-6 * scale


def test_deterministic_rerun(case, gpu_ctx):
    from tsr_amd import Occupancy
    from tsr_amd._lib import TSR_PRIORS_ALL_LEAVES
    occ2 = Occupancy(gpu_ctx, case["gpu_tree"], case["scene"], flags=TSR_PRIORS_ALL_LEAVES).download()
    g = case["field"]
    for k in ("kernel_sums", "cone_lengths", "priors", "occupancy", "node_states"):
        assert np.array_equal(occ2[k], g[k]), k


def test_default_priors_only_where_used(case, gpu_ctx):
    # the default mode computes priors only for leaves whose occupancy reads them;
    # the occupancy field is identical
    from tsr_amd import Occupancy
    occ = Occupancy(gpu_ctx, case["gpu_tree"], case["scene"]).download()
    g = case["field"]
    assert np.array_equal(occ["occupancy"], g["occupancy"])
    L = len(g["occupancy"])
    used = g["kernel_sums"][:L].astype(np.float64) + g["kernel_sums"][L:] >= 100.0
    assert np.array_equal(occ["priors"][used], g["priors"][used])
    assert np.all(np.isnan(occ["priors"][~used]))


def test_link_chunks_match_single_pass(gpu_ctx, monkeypatch):
    """Cone passes split into link chunks (folded in chunk order with two-sum) give
    the single-pass field within the parity bound, the same hit counts and mesh."""
    from tsr_amd import DualMarchingCells, Occupancy, Scene, Tree
    samples, cameras = scene(1, 0.0625)
    sc = Scene(gpu_ctx, samples, cameras)
    tree = Tree(gpu_ctx, sc)
    one = Occupancy(gpu_ctx, tree, sc)
    f1, h1 = one.download(), one.hit_counts()
    m1 = DualMarchingCells(gpu_ctx, tree, one).download()
    monkeypatch.setenv("TSR_CONE_CHUNK", "7000")   # ~10 chunks on this scene
    many = Occupancy(gpu_ctx, tree, sc)
    f2, h2 = many.download(), many.hit_counts()
    m2 = DualMarchingCells(gpu_ctx, tree, many).download()
    monkeypatch.delenv("TSR_CONE_CHUNK")
    assert len(samples["px"]) > 5 * 7000
    assert np.array_equal(h1, h2)
    assert np.array_equal(f1["node_states"], f2["node_states"])
    for k in ("kernel_sums", "cone_lengths"):
        assert np.max(rel_err(f2[k], f1[k], floor=1e-30)) <= 1e-5, k
    assert f1["cone_count"] == f2["cone_count"]
    assert np.sum((f1["occupancy"] > 0) != (f2["occupancy"] > 0)) == 0
    assert np.array_equal(m1["triangles"], m2["triangles"])
    for h in (many, one, tree, sc):
        h.close()


@pytest.mark.parametrize("radius", ["0", "0.3"])
def test_device_lookups_of_the_visit_index(gpu_ctx, monkeypatch, radius):
    """TSR_HOT_RADIUS shrinks the part of the sampling visit index sent to the host, so the
    in-order replay asks the device for the visits of the leaves outside it: the tree is the
    oracle's, bit for bit, and the device lookups did run."""
    import tsr_amd
    from oracle.oracle import Oracle
    s, c = scene(3, 0.125, seed=5)
    print(f"\n[device lookups] config=3 res=0.125 seed=5 TSR_HOT_RADIUS={radius}", flush=True)
    monkeypatch.setenv("TSR_HOT_RADIUS", radius)
    sc = tsr_amd.Scene(gpu_ctx, s, c)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    monkeypatch.delenv("TSR_HOT_RADIUS")
    st = tree.build_stats()
    print(f"build stats {st}", flush=True)
    assert st["events"] > 0 and st["device_lookups"] > 0
    t, o = tree.download(), Oracle(s, c, PARAMS).build_tree().tree()
    for k in ("children", "parents", "sample_counts", "sample_starts", "node_to_leaf", "leaf_nodes",
              "sample_target", "cells"):
        assert np.array_equal(t[k], o[k]), k
    # the default radius answers every lookup from the transferred part
    tree2 = tsr_amd.Tree(gpu_ctx, sc)
    assert tree2.build_stats()["device_lookups"] == 0
    assert np.array_equal(tree2.download()["children"], t["children"])
    tree2.close()
    tree.close()
    sc.close()
