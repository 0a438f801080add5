"""The oracle's crust post-processing (oracle/crust.py) against independent checks (CPU only).

* isles: the kept vertices are exactly those of scipy's connected components with at
  least the minimum vertex count, in their old order.
* holes: a pyramid over a hexagon without its base (a 6-edge hole) and an octahedron with
  one face removed (a 3-edge hole) come back closed -- every edge in two triangles, Euler
  characteristic 2, the fill oriented like the reference's (opposite to the mesh along
  the former border) -- with the vertex and triangle counts of
  FlexibleMesh::reserveForHoleFilling (FlexibleMesh.cpp:1933-1975) and the inner-ring
  positions (f0 / 3)(p_i + p_i+1 + p_i+2) + f1 centre of createHoleFillingRingVertices,
  the ring starting at the first border edge of the edge creation order.
* normals: on an origin-centred regular octahedron the angle-weighted vertex normals
  are the unit position vectors.
"""
import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from oracle.crust import delete_isles, fill_holes, normals

OCTA = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float64)
OCTA_T = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]],
                  np.uint32)


def _closed(tris, border_len):
    """every undirected edge in exactly two triangles; the fill patch walks each former
    border edge in the direction of the existing triangle there (FlexibleMesh::fillHoles
    orders border edges by that triangle's winding and adds fill triangles in ring order),
    so exactly the border edges are traversed twice in the same direction; returns E"""
    t = tris.reshape(-1, 3)
    directed = {}
    for tri in t:
        for k in range(3):
            e = (int(tri[k]), int(tri[(k + 1) % 3]))
            directed[e] = directed.get(e, 0) + 1
    undirected = {}
    for (a, b), c in directed.items():
        undirected[(min(a, b), max(a, b))] = undirected.get((min(a, b), max(a, b)), 0) + c
    assert set(undirected.values()) == {2}
    assert sum(1 for c in directed.values() if c == 2) == border_len
    return len(undirected)


def test_isles_match_connected_components():
    rng = np.random.default_rng(3)
    # three strips of 30, 8 and 3 triangles, vertices interleaved at random
    tris, off = [], 0
    for n in (30, 8, 3):
        for i in range(n):
            tris.append([off + i, off + i + 1, off + i + 2])
        off += n + 2
    V = off
    perm = rng.permutation(V)
    tris = perm[np.array(tris)].astype(np.uint32)
    pos = rng.normal(size=(V, 3)).reshape(-1)
    scales = rng.random(V)
    keys = np.arange(2 * V, dtype=np.uint32)
    p, s, k, t = delete_isles(pos, scales, keys, tris.reshape(-1), 12)
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]]])
    n, lab = connected_components(coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(V, V)), directed=False)
    keep = np.bincount(lab)[lab] >= 12
    assert np.array_equal(s, scales[keep])
    assert np.array_equal(p.reshape(-1, 3), pos.reshape(-1, 3)[keep])
    assert np.array_equal(k.reshape(-1, 2), keys.reshape(-1, 2)[keep])
    new = np.cumsum(keep) - 1
    keep_t = keep[tris].all(1)
    assert np.array_equal(t.reshape(-1, 3), new[tris[keep_t]])
    assert len(s) == 32


def test_three_edge_hole_closes_the_octahedron():
    tris = OCTA_T[1:].reshape(-1)
    p, s, k, t = fill_holes(OCTA.reshape(-1), np.ones(6), np.zeros(12, np.uint32), tris)
    assert len(s) == 6 and len(t) == 3 * 8
    E = _closed(t, 3)
    assert 6 - E + 8 == 2


def test_six_edge_hole_known_answer():
    ring = np.array([[np.cos(a), np.sin(a), 0.0] for a in np.arange(6) * np.pi / 3])
    pos = np.vstack([[0, 0, 1.0], ring])              # apex 0, ring 1..6
    tris = np.array([[0, 1 + i, 1 + (i + 1) % 6] for i in range(6)], np.uint32).reshape(-1)
    p, s, k, t = fill_holes(pos.reshape(-1), np.ones(7), np.zeros(14, np.uint32), tris)
    # reserveForHoleFilling: ring 6 -> 3 new vertices, 6 + 3 triangles; ring 3 -> 1 triangle
    assert len(s) == 7 + 3 and len(t) // 3 == 6 + 10
    E = _closed(t, 6)
    assert len(s) - E + len(t) // 3 == 2
    assert np.all(s[7:] == 0.0) and np.all(k[14:] == 0xFFFFFFFF)
    # the ring starts at the first border edge in creation order, triangle 0's edge (1, 2),
    # taken in that triangle's winding (1 -> 2), and follows the windings: ring 1, 2, ..., 6
    r = [1, 2, 3, 4, 5, 6]
    P = p.reshape(-1, 3)
    centre = ring.mean(0)
    f0, f1 = 0.5, 0.5                                 # floor(log2 6) = 2 rings
    for j, (a, b, c) in enumerate([(0, 1, 2), (2, 3, 4), (4, 5, 0)]):
        expect = (pos[r[a]] + pos[r[b]] + pos[r[c]]) * (f0 / 3.0) + centre * f1
        assert np.allclose(P[7 + j], expect, rtol=0, atol=1e-15)


def test_octahedron_normals_are_radial():
    n = normals(OCTA.reshape(-1), OCTA_T.reshape(-1)).reshape(-1, 3)
    assert np.allclose(n, OCTA, atol=1e-15)
