"""Known answers and independent checks that pin the oracle's tree functions which the
closed forms and invariants of test_oracle_tree.py leave open (CPU only).

* Dual-cell order (DualCells.cpp:49-357): the 8 cells of two depth-2 trees, derived by
  hand from the reference's recursion -- root split plus child 0 split, and root split
  plus child 7 split.  A swapped child-slot table or a reordered face/edge/vertex step
  changes the lists.
* Root scope (Tree.cpp:70-101, Samples.cpp:290-298,442-449): a two-sample known answer
  and the float32 closed form on a random scene.
* Sample depth (Tree.cpp:86-91 / Nodes.cpp:250): known answers.
* Sampling refinement (Nodes.cpp:271-337,433-517): a two-sample scene whose nodes were
  derived by hand (below), compared after 2:1 balancing with the balance closure computed
  here; and, on small scenes, the post-condition that every sample has a reliable
  sampling leaf on both sides, checked by brute force over all leaves.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, split_tree
from tests.helpers import PARAMS, scene

INVALID = 0xFFFFFFFF


def _cells(o):
    return o.tree()["cells"].reshape(-1, 8).tolist()


def test_dual_cell_order_known_answer_child0_split():
    # inner nodes: root, root child 0 (location code 0o10).  DFS numbering: root 0,
    # children 1..8, children of node 1 are 9..16; leaves 0..7 = nodes 9..16, leaves
    # 8..14 = nodes 2..8.  Hand trace of createDualCellsByNode(0):
    #   byNode(1) -> byVertex(9..16); facesYZ By2AlongX(1,2); facesXZ By2AlongY(3,1);
    #   facesXY By2AlongZ(1,5); edgesZ By4InXY(1,2,3,4); edgesY By4InXZ(1,2,5,6);
    #   edgesX By4InYZ(1,3,5,7); byVertex(1..8) -> byVertex(16,2,...,8)
    o = split_tree([0o10])
    assert _cells(o) == [
        [0, 1, 2, 3, 4, 5, 6, 7],
        [1, 8, 3, 8, 5, 8, 7, 8],
        [2, 3, 9, 9, 6, 7, 9, 9],
        [4, 5, 6, 7, 11, 11, 11, 11],
        [3, 8, 9, 10, 7, 8, 9, 10],
        [5, 8, 7, 8, 11, 12, 11, 12],
        [6, 7, 9, 9, 11, 11, 13, 13],
        [7, 8, 9, 10, 11, 12, 13, 14],
    ]


def test_dual_cell_order_known_answer_child7_split():
    # inner nodes: root, root child 7 (0o17).  Leaves 0..6 = nodes 1..7, leaves 7..14 =
    # nodes 9..16 (children of node 8).  Hand trace: byNode(8) -> byVertex(9..16);
    # By2AlongX(7,8); By2AlongY(8,6); By2AlongZ(4,8); By4InXY(5,6,7,8); By4InXZ(3,4,7,8);
    # By4InYZ(2,4,6,8); byVertex(1..8) -> byVertex(1,...,7,9)
    o = split_tree([0o17])
    assert _cells(o) == [
        [7, 8, 9, 10, 11, 12, 13, 14],
        [6, 7, 6, 9, 6, 11, 6, 13],
        [5, 5, 7, 8, 5, 5, 11, 12],
        [3, 3, 3, 3, 7, 8, 9, 10],
        [4, 5, 6, 7, 4, 5, 6, 11],
        [2, 3, 2, 3, 6, 7, 6, 9],
        [1, 1, 3, 3, 5, 5, 7, 8],
        [0, 1, 2, 3, 4, 5, 6, 7],
    ]


def _scene(points, normals, scales, cam=(5.0, 0.0, 0.0)):
    p = np.asarray(points, np.float32).reshape(-1, 3)
    nrm = np.asarray(normals, np.float32).reshape(-1, 3)
    n = len(p)
    samples = {"px": p[:, 0].copy(), "py": p[:, 1].copy(), "pz": p[:, 2].copy(),
               "nx": nrm[:, 0].copy(), "ny": nrm[:, 1].copy(), "nz": nrm[:, 2].copy(),
               "scale": np.asarray(scales, np.float32), "confidence": np.ones(n, np.float32),
               "parent_cameras": np.zeros(n, np.uint32)}
    cameras = {"x": np.array([cam[0]], np.float32), "y": np.array([cam[1]], np.float32),
               "z": np.array([cam[2]], np.float32)}
    return samples, cameras


def test_root_scope_and_sample_depth_known_answer():
    # samples (0,0,0) and (1,0,0), scale 0.25: msd = 0.5, inflation r = 0.5 + 1e-6.
    # AABB x in [-r, 1 + r], y, z in [-r, r]: side = 1 + 2r (the x extent), centre
    # (0.5, 0, 0), so the root's min corner is (0.5 - side/2, -side/2, -side/2).
    s, c = _scene([[0, 0, 0], [1, 0, 0]], [[0, 0, 1], [0, 0, 1]], [0.25, 0.25])
    t = Oracle(s, c, PARAMS).build_tree().tree()
    r = np.float32(0.5) + np.float32(1e-6)
    side = np.float32(np.float32(1) + r) - np.float32(-r)
    assert t["root_size"] == side
    half = side * np.float32(0.5)
    assert np.array_equal(t["root_min"], np.array([np.float32(0.5) - half, -half, -half], np.float32))
    # sample depth: smallest d with side * 2^-d <= 0.5 -> side = 2.000002: d = 3
    assert t["sample_depth"].tolist() == [3, 3]


def test_root_scope_closed_form_random_scene():
    rng = np.random.default_rng(11)
    n = 400
    p = rng.uniform(-3, 2, (n, 3)).astype(np.float32)
    p[:, 1] *= np.float32(0.3)                     # a flat cloud: the side is the x/z extent
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    sc = rng.uniform(0.01, 0.2, n).astype(np.float32)
    s, c = _scene(p, nrm, sc)
    t = Oracle(s, c, PARAMS).build_tree().tree()
    r = sc * np.float32(PARAMS[4]) + np.float32(1e-6)
    mn = (p - r[:, None]).min(0)
    mx = (p + r[:, None]).max(0)
    ctr = (mn + mx) * np.float32(0.5)
    side = (mx - mn).max()
    assert t["root_size"] == side
    assert np.array_equal(t["root_min"], ctr - side * np.float32(0.5))
    # every inflated sample box lies in the root cube, which touches the extreme boxes
    lo = t["root_min"].astype(np.float64)
    hi = lo + float(t["root_size"])
    assert np.all(p - r[:, None] >= lo - 1e-6) and np.all(p + r[:, None] <= hi + 1e-6)
    # sample depth: the smallest d with rootSize * 2^-d <= msd (Nodes.cpp:250)
    msd = (sc * np.float32(PARAMS[4])).astype(np.float64)
    d = t["sample_depth"].astype(np.int64)
    h = float(t["root_size"]) * 2.0 ** -d
    assert np.all(h <= msd) and np.all(2 * h > msd)


def _inner_codes(children):
    """location codes of the inner nodes, from the DFS-ordered child table"""
    loc = np.zeros(len(children), np.uint64)
    loc[0] = 1
    codes = []
    stack = [0]
    while stack:
        nd = stack.pop()
        c = children[nd]
        if c == INVALID:
            continue
        codes.append(int(loc[nd]))
        for s in range(8):
            loc[c + s] = (int(loc[nd]) << 3) | s
            stack.append(c + s)
    return set(codes)


def _balance_closure(inner):
    """smallest 2:1 face-balanced tree containing the given inner nodes (Nodes.cpp:339-395:
    a leaf's same-size face neighbour cell may lie in a leaf at most one level coarser)"""
    inner = set(inner)

    def dk(code):
        d = (code.bit_length() - 1) // 3
        m = code ^ (1 << 3 * d)
        k = [0, 0, 0]
        for b in range(d):
            for a in range(3):
                k[a] |= ((m >> (3 * b + a)) & 1) << b
        return d, k

    def code_of(d, k):
        m = 0
        for b in range(d):
            for a in range(3):
                m |= ((k[a] >> b) & 1) << (3 * b + a)
        return (1 << 3 * d) | m

    changed = True
    while changed:
        changed = False
        leaves = [(c << 3) | s for c in inner for s in range(8) if ((c << 3) | s) not in inner]
        for leaf in leaves:
            d, k = dk(leaf)
            for a in range(3):
                for sg in (-1, 1):
                    q = list(k)
                    q[a] += sg
                    if not 0 <= q[a] < (1 << d):
                        continue
                    nd = 0   # descend from the root while the node is inner
                    node = 1
                    while node in inner and nd < d:
                        nd += 1
                        node = code_of(nd, [x >> (d - nd) for x in q])
                    if node not in inner and nd < d - 1:
                        inner.add(node)
                        changed = True
    return inner


def test_sampling_refinement_known_answer():
    # Sample A: origin, normal +x, scale 0.05 (msd 0.1); sample B: origin, normal +x,
    # scale 0.5 (msd 1).  Root: centre 0, side 2(1 + 1e-6); node sizes s_d = 2.000002 / 2^d.
    # Containment: both samples sit on the root centre, so child slots >= centre give the
    # chain root -> 7 -> 0 -> 0 -> 0 -> 0: A's container at depth 5 (s_5 = 0.0625 <= 0.1 <
    # s_4), B's at depth 2.  Inner nodes 1, 0o17, 0o170, 0o1700, 0o17000.
    # (A, +): the leaf 0o170000 = [0, s_5]^3 has its centre 0.054 from A with x > 0, and
    #   its child 0 centre 0.027 with x > 0: reliable, nothing created.
    # (A, -): every leaf within 0.1 of A has x >= 0 (on the plane counts as positive only
    #   if > 0, so their centres are positive): none is reliable.  Descent along the
    #   segment A -> A - 0.1 x: at the root the four x- children overlap it fully (y = z = 0
    #   on their closed boundaries); child 0 comes first and 1 >= 1 - EPS stops the scan.
    #   Child 0 = [-1, 0]^3 is split; only its child 7 meets A's ball and its centre is
    #   0.43 away.  Child 7 (the only full overlap) is split, likewise its child 7 (centre
    #   0.217 away) and that one's child 7 (0.108 > 0.1).  After this fourth split the leaf
    #   [-s_5, 0]^3 has its centre 0.054 away with x < 0 and its child 7 0.027 away:
    #   reliable.  Created inner nodes 0o10, 0o107, 0o1077, 0o10777.
    # (B, +), (B, -): root child 1 (centre 0.87 from B, x > 0; its child 6 0.43 away) and
    #   the depth-2 leaf 0o103 (centre 0.83 away, x < 0; its child 7 0.65 away) are
    #   reliable: nothing created.
    s, c = _scene([[0, 0, 0], [0, 0, 0]], [[1, 0, 0], [1, 0, 0]], [0.05, 0.5])
    o = Oracle(s, c, PARAMS).build_tree()
    t = o.tree()
    assert t["sample_depth"].tolist() == [5, 2]
    assert o.count("sampling_nodes_created") == 4 * 8
    containment = {1, 0o17, 0o170, 0o1700, 0o17000}
    sampling = {0o10, 0o107, 0o1077, 0o10777}
    assert _inner_codes(t["children"]) == _balance_closure(containment | sampling)


def test_sampling_refinement_sample_order_and_side_sign():
    # (1) The scene above with the samples' order swapped: B's units now run first, while
    #   root child 0 is still a leaf; it is reliable for B's negative side by itself (centre
    #   0.87 from B with x < 0, its child 7 0.43 away), so B creates nothing and A's units
    #   then create the same four nodes: the same tree.
    # (2) A's normal flipped to -x: its positive side is now x < 0, so the positive unit
    #   needs the same four splits (same segment A -> (-0.1, 0, 0)) and the negative unit
    #   finds [0, s_5]^3 (centre x > 0, i.e. on the negative side of the flipped plane):
    #   the same tree again.  Swapped sides in the refinement would move the splits into
    #   root child 7.
    sA, cA = _scene([[0, 0, 0], [0, 0, 0]], [[1, 0, 0], [1, 0, 0]], [0.05, 0.5])
    sB, cB = _scene([[0, 0, 0], [0, 0, 0]], [[1, 0, 0], [1, 0, 0]], [0.5, 0.05])
    ta = _inner_codes(Oracle(sA, cA, PARAMS).build_tree().tree()["children"])
    tb = _inner_codes(Oracle(sB, cB, PARAMS).build_tree().tree()["children"])
    assert ta == tb
    sC, cC = _scene([[0, 0, 0], [0, 0, 0]], [[-1, 0, 0], [1, 0, 0]], [0.05, 0.5])
    oc = Oracle(sC, cC, PARAMS).build_tree()
    assert oc.count("sampling_nodes_created") == 4 * 8
    assert _inner_codes(oc.tree()["children"]) == ta


def _node_boxes(t):
    """min corner and size of every node (FP64 closed form B(k,d)) and depths"""
    ch = t["children"]
    n = len(ch)
    depth = np.zeros(n, np.int64)
    k = np.zeros((n, 3), np.int64)
    for nd in range(n):
        c = ch[nd]
        if c == INVALID:
            continue
        for s in range(8):
            depth[c + s] = depth[nd] + 1
            k[c + s] = 2 * k[nd] + [(s >> 0) & 1, (s >> 1) & 1, (s >> 2) & 1]
    size = float(t["root_size"]) * 2.0 ** -depth
    lo = t["root_min"].astype(np.float64) + k * size[:, None]
    return lo, size


@pytest.mark.parametrize("cfg,res", [(0, 0.25), (3, 0.0625)], ids=["sphere", "primitives"])
def test_every_sample_has_reliable_sampling_leaves(cfg, res):
    # Nodes.cpp:433-488: after the sampling pass each sample has, on both sides of its
    # tangent plane, a leaf whose centre is within msd on that side and one of whose
    # child centres is too.  Brute force over all leaves of the final tree, in FP64 with
    # the plane dot(x - p, n) > 0 as the positive side.
    s, c = scene(cfg, res)
    t = Oracle(s, c, PARAMS).build_tree().tree()
    lo, size = _node_boxes(t)
    leaves = t["leaf_nodes"]
    ctr = lo[leaves] + size[leaves, None] / 2
    hs = size[leaves] / 4
    p = np.stack([s["px"], s["py"], s["pz"]], 1).astype(np.float64)
    nrm = np.stack([s["nx"], s["ny"], s["nz"]], 1).astype(np.float64)
    msd = (s["scale"].astype(np.float32) * np.float32(PARAMS[4])).astype(np.float64)
    offs = np.array([[(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], np.float64) * 2 - 1
    for i in range(len(p)):
        d2 = np.sum((ctr - p[i]) ** 2, axis=1)
        near = np.flatnonzero(d2 <= msd[i] ** 2)
        side = (ctr[near] - p[i]) @ nrm[i] > 0
        ok = {True: False, False: False}
        for positive in (True, False):
            cand = near[side == positive]
            if len(cand) == 0:
                continue
            kids = ctr[cand][:, None, :] + offs[None] * hs[cand][:, None, None]
            kd2 = np.sum((kids - p[i]) ** 2, axis=2)
            kside = (kids - p[i]) @ nrm[i] > 0
            ok[positive] = bool(np.any((kd2 <= msd[i] ** 2) & (kside == positive)))
        assert ok[True] and ok[False], f"sample {i}: reliable leaf missing ({ok})"


def _expected_neighbour_csr(t):
    """Leaves.cpp:82-180 derived from geometry alone: leaves A, B are neighbours iff their
    boxes share part of a face.  The pair is inserted by the finer leaf (the same-size
    one with the smaller node index when equal) at its loop step (leaf index, side), and
    each leaf's list is in the order of those steps."""
    ch = t["children"]
    leaves = t["leaf_nodes"].astype(np.int64)
    n = len(ch)
    depth = np.zeros(n, np.int64)
    k = np.zeros((n, 3), np.int64)
    for nd in range(n):
        c = ch[nd]
        if c == INVALID:
            continue
        for s in range(8):
            depth[c + s] = depth[nd] + 1
            k[c + s] = 2 * k[nd] + [(s >> 0) & 1, (s >> 1) & 1, (s >> 2) & 1]
    D = int(depth.max())
    d = depth[leaves]
    lo = k[leaves] << (D - d)[:, None]
    sz = (1 << (D - d))
    hi = lo + sz[:, None]
    L = len(leaves)
    recs = {i: [] for i in range(L)}
    for a in range(L):
        for axis in range(3):
            o1, o2 = [x for x in range(3) if x != axis]
            for side_sign in (-1, 1):
                touch = (hi[:, axis] == lo[a, axis]) if side_sign < 0 else (lo[:, axis] == hi[a, axis])
                ov = (np.minimum(hi[:, o1], hi[a, o1]) > np.maximum(lo[:, o1], lo[a, o1])) & \
                     (np.minimum(hi[:, o2], hi[a, o2]) > np.maximum(lo[:, o2], lo[a, o2]))
                for b in np.flatnonzero(touch & ov):
                    # a is responsible if it is finer, or the same size with the smaller node index
                    if sz[a] < sz[b] or (sz[a] == sz[b] and leaves[a] < leaves[b]):
                        side = 2 * axis + (1 if side_sign > 0 else 0)
                        key = a * 6 + side
                        recs[a].append((key, b))
                        recs[b].append((key, a))
    offs = [0]
    nb = []
    for i in range(L):
        nb.extend(b for _, b in sorted(recs[i]))
        offs.append(len(nb))
    return np.array(offs, np.uint32), np.array(nb, np.uint32)


@pytest.mark.parametrize("which", ["sphere", "two-samples"])
def test_leaf_neighbour_csr_matches_geometry(which):
    if which == "sphere":
        s, c = scene(0, 0.125)
    else:
        s, c = _scene([[0, 0, 0], [0, 0, 0]], [[1, 0, 0], [1, 0, 0]], [0.05, 0.5])
    o = Oracle(s, c, PARAMS).build_tree()
    offs, nb = o.leaf_neighbours()
    eo, en = _expected_neighbour_csr(o.tree())
    assert np.array_equal(offs, eo)
    assert np.array_equal(nb, en)
