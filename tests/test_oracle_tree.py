"""The oracle's octree, dual cells, priors and occupancy against invariants, brute
force and closed forms (independent Python restatements on small scenes)."""
import numpy as np
import pytest

from oracle.oracle import Oracle, leaf_centres, mc_case, uniform
from tests.helpers import PARAMS, scene

INVALID = 0xFFFFFFFF


@pytest.fixture(scope="module", params=[(0, 0.25), (3, 0.0625)], ids=["sphere", "primitives"])
def built(request):
    cfg, res = request.param
    s, c = scene(cfg, res)
    o = Oracle(s, c, PARAMS).build_tree()
    return s, c, o, o.tree()


def _node_geometry(t):
    """depth and integer coordinates of every node, from the child links"""
    ch = t["children"]
    n = len(ch)
    depth = np.zeros(n, np.int64)
    k = np.zeros((n, 3), np.int64)
    for nd in range(n):   # parents precede children in DFS block order
        c = ch[nd]
        if c == INVALID:
            continue
        for s in range(8):
            depth[c + s] = depth[nd] + 1
            k[c + s] = 2 * k[nd] + [(s >> 0) & 1, (s >> 1) & 1, (s >> 2) & 1]
    return depth, k


def test_dfs_block_numbering(built):
    _, _, _, t = built
    ch, par = t["children"], t["parents"]
    # Nodes.cpp:536-554: children of the r-th inner node in pre-order occupy 1 + 8r .. 8 + 8r
    order = []
    stack = [0]
    while stack:
        nd = stack.pop()
        if ch[nd] != INVALID:
            order.append(nd)
            stack.extend(ch[nd] + 7 - i for i in range(8))
    for r, nd in enumerate(order):
        assert ch[nd] == 1 + 8 * r
        assert all(par[ch[nd] + i] == nd for i in range(8))
    assert par[0] == INVALID and len(ch) == 1 + 8 * len(order)


def test_samples_contiguous_and_in_their_container(built):
    s, _, o, t = built
    counts, starts = t["sample_counts"].astype(np.int64), t["sample_starts"].astype(np.int64)
    assert np.array_equal(starts, np.concatenate([[0], np.cumsum(counts)[:-1]]))
    assert counts.sum() == len(s["px"])
    depth, k = _node_geometry(t)
    owner = np.repeat(np.arange(len(counts)), counts)      # new sample index -> node
    node_of = owner[t["sample_target"]]                     # original sample -> node
    root_min = t["root_min"].astype(np.float64)
    size = float(t["root_size"])
    msd = s["scale"].astype(np.float32) * np.float32(2.0)
    h = size * 2.0 ** -depth[node_of]
    # container rule (Nodes.cpp:86-99): node size <= msd < 2 * node size
    assert np.all(h <= msd) and np.all(msd < 2 * h)
    p = np.stack([s["px"], s["py"], s["pz"]], 1).astype(np.float64)
    lo = root_min + k[node_of] * h[:, None]
    assert np.all(p >= lo) and np.all(p < lo + h[:, None] * (1 + 1e-12))


def test_leaves_in_morton_order_and_balanced(built):
    _, _, o, t = built
    depth, k = _node_geometry(t)
    leaves = t["leaf_nodes"]
    assert np.array_equal(t["node_to_leaf"][leaves], np.arange(len(leaves)))
    D = 19
    padded = [tuple(int(x) << (D - depth[n]) for x in k[n]) for n in leaves]

    def morton(v):
        m = 0
        for b in range(D):
            for a in range(3):
                m |= ((v[a] >> b) & 1) << (3 * b + a)
        return m
    keys = [morton(v) for v in padded]
    assert keys == sorted(keys)
    # 2:1 face balance (Nodes.cpp:339-395): the leaf holding the centre of a
    # same-size face neighbour is at most one level coarser
    cell = {}
    for n in leaves:
        cell[(int(depth[n]),) + tuple(int(x) for x in k[n])] = int(depth[n])

    def leaf_depth_at(d, q):
        for dd in range(d, -1, -1):
            key = (dd,) + tuple(int(x) >> (d - dd) for x in q)
            if key in cell:
                return cell[key]
        return None
    for n in leaves:
        d = int(depth[n])
        for a in range(3):
            for sgn in (-1, 1):
                q = k[n].copy()
                q[a] += sgn
                if q[a] < 0 or q[a] >= (1 << d):
                    continue
                nd = leaf_depth_at(d, q)
                if nd is not None:
                    assert nd >= d - 1


def test_dual_cells_are_interior_leaf_corners(built):
    _, _, o, t = built
    depth, k = _node_geometry(t)
    D = 19
    leaves = t["leaf_nodes"]
    L = len(leaves)
    cells = t["cells"].reshape(-1, 8)
    # characterisation: one cell per interior corner of any leaf; entry s = leaf
    # containing the octant s around the corner (Schaefer-Warren, DualCells.cpp:49-357)
    lut = {}
    for li, n in enumerate(leaves):
        lut[(int(depth[n]),) + tuple(int(x) for x in k[n])] = li

    def leaf_at(c):
        for d in range(D, -1, -1):
            key = (d,) + tuple(int(x) >> (D - d) for x in c)
            if key in lut:
                return lut[key]
        raise AssertionError("no leaf")
    corners = set()
    for n in leaves:
        sh = D - int(depth[n])
        for s in range(8):
            v = tuple((int(k[n][a]) + ((s >> a) & 1)) << sh for a in range(3))
            if all(0 < x < (1 << D) for x in v):
                corners.add(v)
    expect = set()
    for v in corners:
        expect.add(tuple(leaf_at([v[a] - 1 + ((s >> a) & 1) for a in range(3)]) for s in range(8)))
    got = set(tuple(int(x) for x in c) for c in cells)
    assert len(got) == len(cells) == len(expect)
    assert got == expect
    assert cells.max() < L


def test_uniform_tree_dual_cells_and_sphere_mesh():
    depth = 4
    o = uniform(depth)
    n = 2 ** depth
    assert o.count("leaves") == n ** 3
    assert o.count("cells") == (n - 1) ** 3          # closed form for a full tree
    c = leaf_centres(o)
    sdf = 0.3 - np.linalg.norm(c - 0.5, axis=1)      # > 0 inside a sphere of radius 0.3
    o.dmc(sdf)
    m = o.mesh()
    tri = m["triangles"].reshape(-1, 3)
    V = len(m["positions"]) // 3
    edges = {}
    for t in tri:
        for i in range(3):
            e = tuple(sorted((int(t[i]), int(t[(i + 1) % 3]))))
            edges[e] = edges.get(e, 0) + 1
    assert set(edges.values()) == {2}                # closed 2-manifold
    assert V - len(edges) + len(tri) == 2            # sphere topology
    # independent marching cubes on the leaf-centre grid with the same table
    g = sdf.reshape(-1)
    idx = {}
    for li in range(len(c)):
        idx[tuple(np.round(c[li] * n - 0.5).astype(int))] = li
    count = 0
    corner_offsets = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    for x in range(n - 1):
        for y in range(n - 1):
            for z in range(n - 1):
                cfg = 0
                for ci, off in enumerate(corner_offsets):
                    if g[idx[(x + off[0], y + off[1], z + off[2])]] > 0:
                        cfg |= 1 << ci
                count += len(mc_case(cfg)) // 3
    assert count == len(tri)
    r = np.linalg.norm(m["positions"].reshape(-1, 3) - 0.5, axis=1)
    assert np.all(np.abs(r - 0.3) < 1.0 / n)         # vertices near the sphere


def test_traversal_culling_matches_brute_force():
    s, c = scene(0, 0.125)
    a = Oracle(s, c, PARAMS).build_tree().occupancy()
    b = Oracle(s, c, PARAMS).build_tree().occupancy(brute_force=True)
    fa, fb = a.field(), b.field()
    for key in ("kernel_sums", "cone_lengths", "priors", "occupancy", "node_states"):
        assert np.array_equal(fa[key], fb[key], equal_nan=True), key


def test_priors_brute_force_and_occupancy_closed_form():
    s, c = scene(0, 0.125)
    o = Oracle(s, c, PARAMS).build_tree().occupancy()
    t, f = o.tree(), o.field()
    depth, k = _node_geometry(t)
    leaves = t["leaf_nodes"]
    L = len(leaves)
    size = float(t["root_size"])
    rmin = t["root_min"].astype(np.float64)
    ld = depth[leaves]
    h = size * 2.0 ** -ld
    lo = rmin + k[leaves] * h[:, None]
    hi = lo + h[:, None]
    ctr = lo + h[:, None] / 2
    sflag = (f["node_states"][leaves] & 1) != 0
    E, S = f["cone_lengths"][:L], f["cone_lengths"][L:]
    for li in range(0, L, max(1, L // 300)):
        r = h[li] * 25.0
        d2 = np.sum(np.maximum(0, np.maximum(lo - ctr[li], ctr[li] - hi)) ** 2, axis=1)
        sel = (d2 <= r * r) & (ld <= ld[li] + 3) & sflag
        e, sm = E[sel].sum(), S[sel].sum()
        if e + sm == 0:
            continue
        assert f["priors"][li] == pytest.approx(e / (e + sm), rel=1e-12, abs=1e-15)
    # Occupancy.cpp:570-587
    kE, kS = f["kernel_sums"][:L], f["kernel_sums"][L:]
    M = f["cone_count"]
    pr = f["priors"]
    with np.errstate(invalid="ignore"):
        pE, pS = pr * (kE / M), (1 - pr) * (kS / M)
        occ = np.where((kE + kS < 100.0) | (pE >= pS), -np.float64(np.float32(1e-6)), pS - pE)
    assert np.array_equal(occ, f["occupancy"])
