"""The large single-GPU scene: 30 random primitives including thin sheets in [-4,4]^3,
120 cameras at 3..20, 1024x768 depth maps, 2 % outliers (36.5M samples).

* Reduced resolution (1/16 of the image side, ~142k samples): parity with the FP64
  oracle -- tree tables bit for bit, kernel sums within 1e-5 relative, occupancy
  within 1e-5 of (pS + pE) with no sign flip, the same mesh.
* Full size (and the 9.2M-sample thin wedge + sphere scene at full size too): properties
  that do not need the oracle -- a rerun gives bit-identical
  tables, field and mesh; node ids are in DFS child-block pre-order (Nodes.cpp:536-554);
  the container rule (Nodes.cpp:86-99) and 2:1 face balance
  (Nodes.cpp:339-395) on a sample of samples and leaves; every mesh edge is shared by
  at most two triangles and no triangle is degenerate (FlexibleMesh.cpp:1227-1260).
"""
import hashlib

import numpy as np
import pytest

from tests.helpers import PARAMS, occupancy_rel_err, rel_err, scene

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


def _digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_config4_reduced_resolution_matches_oracle(gpu_ctx):
    from oracle.oracle import Oracle
    from tsr_amd import DualMarchingCells, Occupancy, Scene, Tree
    from tsr_amd._lib import TSR_PRIORS_ALL_LEAVES
    samples, cameras = scene(4, 0.0625)
    print(f"\n[config4 1/16] samples={len(samples['px'])}", flush=True)
    orc = Oracle(samples, cameras, PARAMS).build_tree().occupancy(threads=8, fp64=True).dmc()
    sc = Scene(gpu_ctx, samples, cameras)
    tree = Tree(gpu_ctx, sc)
    occ = Occupancy(gpu_ctx, tree, sc, flags=TSR_PRIORS_ALL_LEAVES)
    mesh = DualMarchingCells(gpu_ctx, tree, occ)
    t, o = tree.download(), orc.tree()
    for k in ("children", "parents", "sample_counts", "sample_starts", "node_to_leaf", "leaf_nodes",
              "sample_target", "cells", "leaf_scopes"):
        assert np.array_equal(t[k], o[k]), k
    g, of = occ.download(), orc.field()
    assert np.array_equal(g["node_states"], of["node_states"])
    for k in ("kernel_sums", "cone_lengths"):
        assert np.max(rel_err(g[k], of[k], floor=1e-30)) <= 1e-5, k
    err, flips = occupancy_rel_err(g["occupancy"], of)
    assert flips == 0 and err.max() <= 1e-5, (flips, err.max())
    m, om = mesh.download(), orc.mesh()
    assert np.array_equal(m["triangles"], om["triangles"])
    assert np.array_equal(m["edge_keys"], om["edge_keys"])
    assert np.allclose(m["positions"], om["positions"], rtol=0, atol=1e-9 * float(t["leaf_scopes"][3::4].max()))
    for x in (mesh, occ, tree, sc):
        x.close()


def _node_coords(node, depth, parents):
    """integer grid coordinates of nodes at their own depth, from the DFS block numbering
    (slot of node n = (n - 1) mod 8, Nodes.cpp:536-554)"""
    k = np.zeros((len(node), 3), np.int64)
    cur = node.astype(np.int64).copy()
    d = depth.astype(np.int64)
    for lvl in range(int(d.max()) if len(d) else 0):
        act = lvl < d
        slot = (cur - 1) & 7
        for a in range(3):
            k[:, a] |= np.where(act, ((slot >> a) & 1) << lvl, 0)
        cur = np.where(act, parents[np.where(act, cur, 0)].astype(np.int64), cur)
    return k


def _preorder_violations(children):
    """Number of inner nodes whose child block is out of DFS pre-order (Nodes.cpp:536-554):
    sorting inner nodes by child-block rank must sort them by (Morton code padded to the
    deepest level, depth)."""
    N = len(children)
    loc = np.zeros(N, np.uint64)
    loc[0] = 1
    depth = np.zeros(N, np.int64)
    front = np.array([0], np.int64)
    while len(front):
        inner = front[children[front] != INVALID]
        c0 = children[inner].astype(np.int64)
        nxt = []
        for s in range(8):
            loc[c0 + s] = (loc[inner] << np.uint64(3)) | np.uint64(s)
            depth[c0 + s] = depth[inner] + 1
            nxt.append(c0 + s)
        front = np.concatenate(nxt) if nxt else np.zeros(0, np.int64)
    ids = np.flatnonzero(children != INVALID)
    d = depth[ids].astype(np.uint64)
    top = np.uint64(int(depth.max()))
    m = loc[ids] ^ (np.uint64(1) << (np.uint64(3) * d))
    key = (m << (np.uint64(3) * (top - d) + np.uint64(5))) | d
    k = key[np.argsort((children[ids].astype(np.int64) - 1) // 8)]
    return int(np.sum(k[1:] <= k[:-1]))


def _leaf_depth_at(q, d, children, depth):
    """depth of the leaf holding the same-depth cell q (descending from the root,
    stopping at depth d)"""
    node = np.zeros(len(q), np.int64)
    for lvl in range(int(d.max()) if len(d) else 0):
        c = children[node].astype(np.int64)
        go = (lvl < d) & (c != INVALID)
        sh = (d - 1 - lvl).clip(min=0)
        slot = ((q[:, 0] >> sh) & 1) | (((q[:, 1] >> sh) & 1) << 1) | (((q[:, 2] >> sh) & 1) << 2)
        node = np.where(go, c + slot, node)
    return depth[node]


@pytest.mark.parametrize("config", [4, 1])
def test_full_size_properties(gpu_ctx, config):
    """config 4 (36.5M samples) and the thin wedge + sphere scene (config 1, 9.2M samples)"""
    from tsr_amd import DualMarchingCells, Occupancy, Scene, Tree
    samples, cameras = scene(config, 1.0)
    n = len(samples["px"])
    print(f"\n[config{config} full] samples={n}", flush=True)
    sc = Scene(gpu_ctx, samples, cameras)
    digests = []
    keep = None
    for run in range(2):
        tree = Tree(gpu_ctx, sc)
        occ = Occupancy(gpu_ctx, tree, sc)
        mesh = DualMarchingCells(gpu_ctx, tree, occ)
        t = tree.download(("children", "parents", "sample_counts", "sample_target", "leaf_nodes", "node_depth",
                           "cells"))
        f = occ.download()
        m = mesh.download()
        digests.append((_digest(t["children"], t["cells"], t["sample_target"]),
                        _digest(f["kernel_sums"], f["cone_lengths"], f["occupancy"], f["node_states"]),
                        _digest(m["triangles"], m["positions"], m["edge_keys"])))
        print(f"[config{config} full] run {run}: nodes={tree.node_count} leaves={tree.leaf_count} "
              f"cells={tree.cell_count} triangles={mesh.triangle_count}", flush=True)
        if run == 0:
            keep = (t, m, tree.root_min.astype(np.float64), float(tree.root_size))
        for x in (mesh, occ, tree):
            x.close()
    sc.close()
    assert digests[0] == digests[1], "rerun differs"
    t, m, root_min, root_size = keep
    assert _preorder_violations(t["children"]) == 0, "node ids are not in DFS block pre-order"
    rng = np.random.default_rng(4)
    children, parents, depth = t["children"], t["parents"], t["node_depth"].astype(np.int64)
    # container rule on 200k samples: node size <= msd < 2 * node size, sample inside the node
    idx = rng.choice(n, size=min(n, 200_000), replace=False)
    ends = np.cumsum(t["sample_counts"].astype(np.int64))
    node_of = np.searchsorted(ends, t["sample_target"][idx].astype(np.int64), side="right")
    h = root_size * 2.0 ** -depth[node_of]
    msd = samples["scale"][idx].astype(np.float32) * np.float32(PARAMS[4])
    assert np.all(h <= msd) and np.all(msd < 2 * h)
    k = _node_coords(node_of, depth[node_of], parents)
    p = np.stack([samples["px"][idx], samples["py"][idx], samples["pz"][idx]], 1).astype(np.float64)
    lo = root_min + k * h[:, None]
    assert np.all(p >= lo) and np.all(p < lo + h[:, None] * (1 + 1e-12))
    # 2:1 face balance on 200k leaves
    leaves = t["leaf_nodes"]
    sel = leaves[rng.choice(len(leaves), size=min(len(leaves), 200_000), replace=False)].astype(np.int64)
    d = depth[sel]
    kl = _node_coords(sel, d, parents)
    for a in range(3):
        for sgn in (-1, 1):
            q = kl.copy()
            q[:, a] += sgn
            ok = (q[:, a] >= 0) & (q[:, a] < (1 << d))
            nd = _leaf_depth_at(q[ok], d[ok], children, depth)
            assert np.all(nd >= d[ok] - 1), "2:1 balance violated"
    # mesh: no degenerate triangle, every edge in at most two triangles, unique vertex keys
    tri = m["triangles"].reshape(-1, 3).astype(np.int64)
    assert np.all((tri[:, 0] != tri[:, 1]) & (tri[:, 1] != tri[:, 2]) & (tri[:, 0] != tri[:, 2]))
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    e.sort(axis=1)
    _, cnt = np.unique(e[:, 0] * (1 << 32) + e[:, 1], return_counts=True)
    assert cnt.max() <= 2, "non-manifold edge"
    keys = m["edge_keys"].reshape(-1, 2).astype(np.int64)
    assert len(np.unique(keys[:, 0] * (1 << 32) + keys[:, 1])) == len(keys)
