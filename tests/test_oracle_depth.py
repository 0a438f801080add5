"""The oracle's depth-map sample preparation (oracle/depth.py) against hand-derived answers
(CPU only).

* A 3 x 3 depth map on a plane tilted so that every block takes the diagonal 1-2: the vertex ids
  follow the pixels' first use in block order, derived by hand, and so do the 8 triangles.
* A step of many pixel footprints splits the mesh: no triangle crosses it; a pixel without depth
  leaves its blocks one triangle (mask cases 7, 11, 13, 14).
* The vertex at the principal point lies at centre + R^T (0, 0, 1) * depth; the normals of a
  planar depth map all equal the plane normal facing the camera; scales are the mean distance to
  the hand-derived neighbour sets; parent cameras come from the views image, or the view itself.
"""
import numpy as np

from oracle.depth import depth_mesh, samples_from_meshes, triangulate

IDENT = np.eye(3)


def _view(depths, **kw):
    v = {"depths": np.asarray(depths, np.float32), "position": np.zeros(3), "rotation": IDENT,
         "focal_length": 1.0, "principal_point": (0.5, 0.5), "aspect_ratio": 1.0, "camera_index": 7}
    v.update(kw)
    return v


def test_first_use_order_known_answer():
    # depth grows by 0.01 per column and 0.02 per row: |d0 - d3| = 0.03 > |d1 - d2| = 0.01 in
    # every block, so every block takes triangles 0 {0, 2, 1} and 3 {1, 2, 3}
    d = np.array([[10 + 0.01 * x + 0.02 * y for x in range(3)] for y in range(3)])
    p2v, tris, pos = triangulate(_view(d))
    assert p2v.tolist() == [0, 2, 4, 1, 3, 5, 6, 7, 8]
    assert tris.reshape(-1, 3).tolist() == [[0, 1, 2], [2, 1, 3], [2, 3, 4], [4, 3, 5],
                                            [1, 6, 3], [3, 6, 7], [3, 7, 5], [5, 7, 8]]
    # the other diagonal: depth falling along x
    d2 = np.array([[10 - 0.01 * x + 0.02 * y for x in range(3)] for y in range(3)])
    p2v2, tris2, _ = triangulate(_view(d2))
    # block (0, 0) takes triangles 1 {0, 3, 1} and 2 {0, 2, 3}: pixels 0, 4, 1 then 3
    assert p2v2[[0, 4, 1, 3]].tolist() == [0, 1, 2, 3]
    assert tris2.reshape(-1, 3)[:2].tolist() == [[0, 1, 2], [0, 3, 1]]


def test_depth_discontinuity_and_missing_pixels():
    d = np.full((4, 6), 10.0)
    d[:, 3:] = 20.0                       # a step of 10 at about 60 pixel footprints per unit depth
    p2v, tris, _ = triangulate(_view(d))
    t = tris.reshape(-1, 3)
    col = np.flatnonzero(p2v != 0xFFFFFFFF) % 6
    side = {int(v): int(c >= 3) for v, c in zip(p2v[p2v != 0xFFFFFFFF], col)}
    assert all(len({side[int(v)] for v in tri}) == 1 for tri in t)   # no triangle crosses the step
    # blocks inside each side keep both triangles (3 rows x (2 + 2) columns of blocks), the column
    # of blocks across the step none
    assert len(t) == 2 * 3 * 4
    # one pixel without depth: its 4 blocks keep one triangle each
    d = np.full((3, 3), 10.0)
    d[1, 1] = 0.0
    _, tris, _ = triangulate(_view(d))
    assert len(tris) // 3 == 4


def test_positions_normals_scales_cameras():
    W = H = 5
    f = 0.8
    # a fronto-parallel plane at z = 4: depth along the ray = 4 |ray|
    fx = f * W
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    ray = np.stack([(xs + 0.5 - 2.5) / fx, (ys + 0.5 - 2.5) / fx, np.ones_like(xs, float)], -1)
    d = 4.0 * np.linalg.norm(ray, axis=-1)
    R = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], float)   # world-to-view
    c = np.array([1.0, 2.0, 3.0])
    views = np.arange(W * H * 2, dtype=np.uint32).reshape(H, W, 2)
    m = depth_mesh(_view(d, rotation=R, position=c, focal_length=f, views=views, views_channels=2))
    p2v = m["pixel_to_vertex"]
    centre = 2 * W + 2                                        # pixel (2, 2): its ray is (0, 0, 1)
    assert np.allclose(m["positions"][p2v[centre]], c + R.T @ np.array([0, 0, 1.0]) * 4.0, atol=1e-12)
    # the plane z = 4 (view) with its normal facing the camera, -z in view space, R^T (0, 0, -1) in
    # world space (depths are float32: the plane is flat to 1e-7 relative)
    n = m["normals"]
    assert np.allclose(n, np.tile(R.T @ np.array([0, 0, -1.0]), (len(n), 1)), atol=1e-6)
    # scales: mean distance to the distinct neighbours of the triangle list
    P, t = m["positions"], m["triangles"].reshape(-1, 3)
    for v in range(len(P)):
        nb = set(t[(t == v).any(1)].reshape(-1).tolist()) - {v}
        assert np.isclose(m["scales"][v], np.mean([np.linalg.norm(P[b] - P[v]) for b in nb]), rtol=1e-12)
    # parent cameras: the views image at each vertex's pixel
    assert m["cameras_per_sample"] == 2
    inv = np.full(len(P), -1)
    inv[p2v[p2v != 0xFFFFFFFF]] = np.flatnonzero(p2v != 0xFFFFFFFF)
    assert np.array_equal(m["parent_cameras"].reshape(-1, 2), views.reshape(-1, 2)[inv])
    # without a views image: the view's own camera (Scene.cpp:561-571)
    m1 = depth_mesh(_view(d))
    assert m1["cameras_per_sample"] == 1 and np.all(m1["parent_cameras"] == 7)
    s, mc = samples_from_meshes([m1, m])
    assert mc == 2 and len(s["px"]) == len(m1["positions"]) + len(P)
    pc = s["parent_cameras"].reshape(-1, 2)
    assert np.all(pc[:len(m1["positions"]), 0] == 7) and np.all(pc[:len(m1["positions"]), 1] == 0xFFFFFFFF)
    assert np.all(s["confidence"] == 1.0)
