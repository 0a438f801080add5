"""Sample preparation from depth maps on the GPU (tsr_depth_mesh) against the oracle
restatement (oracle/depth.py): pixel-to-vertex ids and triangles bit for bit, positions,
normals and scales to FP64 rounding, parent cameras exactly; then the depth-map samples run
through tree, occupancy and DMC.

Depth maps: the hand-derived 3 x 3 tilted plane, and a unit sphere seen by cameras on a ring,
rendered here by ray casting (depth along the ray, 0 where a ray misses), one view with a views
image and one without.
"""
import numpy as np
import pytest

import tsr_amd
from oracle.depth import depth_mesh

pytestmark = pytest.mark.gpu


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0)):
    """world-to-view rotation of a camera at eye looking at target (view: x right, y down, z forward)"""
    z = np.asarray(target, float) - eye
    z /= np.linalg.norm(z)
    x = np.cross(z, up)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return np.stack([x, y, z])


def sphere_view(eye, W, H, f, index, noise=0.0, seed=0):
    R = look_at(np.asarray(eye, float))
    fx = f * max(W, H)
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    ray = np.stack([(xs + 0.5 - 0.5 * W) / fx, (ys + 0.5 - 0.5 * H) / fx, np.ones((H, W))], -1)
    w = ray @ R                                   # R^T r for row vectors
    w /= np.linalg.norm(w, axis=-1, keepdims=True)
    e = np.asarray(eye, float)
    b = w @ e
    disc = b * b - (e @ e - 1.0)
    t = -b - np.sqrt(np.maximum(disc, 0.0))
    depth = np.where(disc > 0, t, 0.0)
    if noise:
        depth = np.where(depth > 0, depth * (1 + noise * np.random.default_rng(seed).normal(size=depth.shape)), 0.0)
    # the camera as the C ABI carries it (float32), so that the oracle sees the same camera
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
    return {"depths": depth.astype(np.float32), "position": f32(e), "rotation": f32(R), "focal_length": f32(f),
            "principal_point": (0.5, 0.5), "aspect_ratio": 1.0, "camera_index": index}


def _compare(ctx, view):
    g = tsr_amd.DepthMesh(ctx, view)
    got = g.download()
    g.close()
    ref = depth_mesh(view)
    assert np.array_equal(got["pixel_to_vertex"], ref["pixel_to_vertex"])
    assert np.array_equal(got["triangles"], ref["triangles"])
    P = ref["positions"].reshape(-1)
    assert np.allclose(got["positions"], P, rtol=1e-13, atol=1e-13)
    assert np.allclose(got["normals"], ref["normals"].reshape(-1), rtol=0, atol=1e-9)
    assert np.allclose(got["scales"], ref["scales"], rtol=1e-12, atol=0)
    assert got["cameras_per_sample"] == ref["cameras_per_sample"]
    assert np.array_equal(got["parent_cameras"], ref["parent_cameras"])
    return got


def test_depth_meshes_match_the_oracle(gpu_ctx):
    d = np.array([[10 + 0.01 * x + 0.02 * y for x in range(3)] for y in range(3)], np.float32)
    plane = {"depths": d, "position": np.zeros(3), "rotation": np.eye(3), "focal_length": 1.0,
             "principal_point": (0.5, 0.5), "aspect_ratio": 1.0, "camera_index": 0}
    got = _compare(gpu_ctx, plane)
    assert got["pixel_to_vertex"].tolist() == [0, 2, 4, 1, 3, 5, 6, 7, 8]
    for i, eye in enumerate([(3.0, 0.0, 0.4), (0.0, 2.5, -0.3)]):
        v = sphere_view(eye, 64, 48, 0.9, i, noise=0.002, seed=i)
        if i == 0:
            H, W = v["depths"].shape
            v["views"] = (np.arange(H * W * 3, dtype=np.uint32).reshape(H, W, 3) % 5).astype(np.uint32)
            v["views"][..., 2] = 0xFFFFFFFF
        g = _compare(gpu_ctx, v)
        print(f"\n[depth] view {i}: {len(g['scales'])} vertices, {len(g['triangles']) // 3} triangles", flush=True)
        assert len(g["triangles"]) > 0


def test_depth_samples_reconstruct(gpu_ctx):
    eyes = [(3 * np.cos(a), 3 * np.sin(a), 0.3 * np.sin(3 * a)) for a in np.linspace(0, 2 * np.pi, 8, endpoint=False)]
    views = [sphere_view(e, 128, 96, 0.9, i, noise=0.002, seed=i) for i, e in enumerate(eyes)]
    samples, mc = tsr_amd.samples_from_depth_maps(gpu_ctx, views)
    cams = {"x": np.array([e[0] for e in eyes], np.float32), "y": np.array([e[1] for e in eyes], np.float32),
            "z": np.array([e[2] for e in eyes], np.float32)}
    assert np.all(samples["scale"] > 0)
    r = np.sqrt(samples["px"].astype(np.float64) ** 2 + samples["py"] ** 2 + samples["pz"] ** 2)
    assert np.all(np.abs(r - 1.0) < 0.05)
    # the normals face the cameras: outwards on the sphere
    nd = samples["nx"] * samples["px"] + samples["ny"] * samples["py"] + samples["nz"] * samples["pz"]
    assert np.mean(nd > 0.5) > 0.99
    sc = tsr_amd.Scene(gpu_ctx, samples, cams, max_cams=mc)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    occ = tsr_amd.Occupancy(gpu_ctx, tree, sc)
    m = tsr_amd.DualMarchingCells(gpu_ctx, tree, occ).download()
    P = m["positions"].reshape(-1, 3)
    dev = np.abs(np.linalg.norm(P, axis=1) - 1.0)
    print(f"\n[depth] {len(samples['px'])} samples -> {len(P)} mesh vertices; distance to the sphere "
          f"p50 {np.median(dev):.4f} p99 {np.percentile(dev, 99):.4f} max {dev.max():.4f}", flush=True)
    assert len(P) > 1000
    # the crust lies within a few leaf sizes (about 2 bandwidths of the sample scale, ~0.02 here)
    assert np.percentile(dev, 99) < 0.1 and dev.max() < 0.3
    for x in (occ, tree, sc):
        x.close()
