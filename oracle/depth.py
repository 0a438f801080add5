"""CPU restatement of the sample preparation from depth maps -- TEST INFRASTRUCTURE ONLY
(the product never imports it).

Follows (paths relative to SurfaceReconstruction/SurfaceReconstruction/):
  triangulate       DepthImage::triangulate / triangulateBlock / isDepthDiscontinuity /
                    setVertexPositions, Image/DepthImage.cpp:191-438 (DEPTH_DIFFERENCE_FACTOR 5,
                    DepthImage.cpp:25)
  vertex_scales     FlexibleMesh::findVertexNeighbors (Geometry/FlexibleMesh.cpp:137-165) and
                    Mesh::computeVertexScales / computeAverageDistance (Geometry/Mesh.cpp:63-124)
  vertex_cameras    Scene::loadCameraIndices, Scene/Scene.cpp:526-573 (views image, else the
                    reference view alone)
  samples_from_meshes  Samples::addSamplesViaMeshes / addSamplesViaMesh, Scene/Samples.cpp:224-273
Normals are oracle.crust.normals (Mesh::computeNormalsWeightedByAngles).  The pixel ray is
this build's camera model (the reference's comes from BaseProject's PinholeCamera, absent):
((x + 0.5 - ppx W) / fx, (y + 0.5 - ppy H) / fy, 1) in view space, fx = f max(W, H),
fy = fx * aspect, rotated to world space by R^T.  FP64 throughout; pure Python loops over the
blocks, so small depth maps only.
"""
import math

import numpy as np

from oracle.crust import normals as angle_weighted_normals

INVALID = 0xFFFFFFFF
BLOCK_TRIANGLES = ((0, 2, 1), (0, 3, 1), (0, 2, 3), (1, 2, 3))   # DepthImage.cpp:227-234


def _intrinsics(view, W, H):
    fx = float(view["focal_length"]) * max(W, H)
    return fx, fx * float(view["aspect_ratio"]), float(view["principal_point"][0]) * W, \
        float(view["principal_point"][1]) * H


def _ray(x, y, fx, fy, cx, cy):
    return np.array([(x + 0.5 - cx) / fx, (y + 0.5 - cy) / fy, 1.0])


def triangulate(view):
    """(pixel_to_vertex, triangles, positions) of one depth map"""
    d = np.asarray(view["depths"], np.float64)
    H, W = d.shape
    fx, fy, cx, cy = _intrinsics(view, W, H)
    p2v = np.full(W * H, INVALID, np.int64)
    tris = []
    vcount = 0
    flat = d.reshape(-1)
    for y in range(H - 1):
        for x in range(W - 1):
            p = y * W + x
            bd = [flat[p], flat[p + 1], flat[p + W], flat[p + W + 1]]
            mask, avail = 0, 0
            for k in range(4):
                if bd[k] > 0.0:
                    mask |= 1 << k
                    avail += 1
            if avail < 3:
                continue
            t = [None, None]
            if mask == 7:
                t[0] = 0
            elif mask == 11:
                t[0] = 1
            elif mask == 13:
                t[0] = 2
            elif mask == 14:
                t[0] = 3
            elif mask == 15:
                if abs(bd[0] - bd[3]) < abs(bd[1] - bd[2]):
                    t = [1, 2]
                else:
                    t = [0, 3]
            else:
                continue
            fp = [0.0] * 4
            for k in range(4):
                if bd[k] > 0.0:
                    v = _ray(x + k % 2, y + k // 2, fx, fy, cx, cy)
                    fp[k] = bd[k] / math.sqrt(v @ v) / fx
            for s in range(2):
                if t[s] is None:
                    continue
                tr = BLOCK_TRIANGLES[t[s]]
                for e in range(3):
                    v0, v1 = tr[e], tr[(e + 1) % 3]
                    closer, further = (v0, v1) if bd[v0] < bd[v1] else (v1, v0)
                    dd = 5.0 * (math.sqrt(2.0) if v0 + v1 == 3 else 1.0)
                    if bd[further] - bd[closer] > fp[closer] * dd:
                        t[s] = None
                        break
            for s in range(2):
                if t[s] is None:
                    continue
                for corner in BLOCK_TRIANGLES[t[s]]:
                    q = p + (corner // 2) * W + corner % 2
                    if p2v[q] == INVALID:
                        p2v[q] = vcount
                        vcount += 1
                    tris.append(int(p2v[q]))
    R = np.asarray(view["rotation"], np.float64).reshape(3, 3)
    c = np.asarray(view["position"], np.float64)
    pos = np.zeros((vcount, 3))
    for q in np.flatnonzero(p2v != INVALID):
        w = R.T @ _ray(q % W, q // W, fx, fy, cx, cy)
        pos[p2v[q]] = c + w / np.linalg.norm(w) * flat[q]
    return p2v.astype(np.uint32), np.array(tris, np.uint32), pos


def vertex_scales(tris, pos):
    """mean distance to the distinct mesh neighbours (-REAL_MAX without any)"""
    nb = [[] for _ in range(len(pos))]
    t = np.asarray(tris).reshape(-1, 3)
    for tri in t:
        e = [int(tri[0]), int(tri[1])]
        for k in range(3):
            for a, b in ((e[0], e[1]), (e[1], e[0])):
                if b not in nb[a]:
                    nb[a].append(b)
            e = [e[1], int(tri[(k + 2) % 3])]
    out = np.empty(len(pos))
    for v in range(len(pos)):
        ds = [np.linalg.norm(pos[b] - pos[v]) for b in nb[v] if b != v]
        out[v] = sum(ds) / len(ds) if ds else -3.4028234663852886e38
    return out


def vertex_cameras(view, p2v, vcount):
    """(cameras per sample, parent cameras of every vertex)"""
    if view.get("views") is None:
        return 1, np.full(vcount, view["camera_index"], np.uint32)
    v = np.asarray(view["views"], np.uint32)
    ch = int(view.get("views_channels") or (v.shape[2] if v.ndim == 3 else 1))
    views = v.reshape(-1, ch)
    out = np.full((vcount, ch), INVALID, np.uint32)
    for q in np.flatnonzero(p2v != INVALID):
        out[p2v[q]] = views[q]
    return ch, out.reshape(-1)


def depth_mesh(view):
    p2v, tris, pos = triangulate(view)
    nrm = angle_weighted_normals(pos.reshape(-1), tris).reshape(-1, 3) if len(tris) else np.zeros_like(pos)
    cps, cams = vertex_cameras(view, p2v, len(pos))
    return {"pixel_to_vertex": p2v, "triangles": tris, "positions": pos, "normals": nrm,
            "scales": vertex_scales(tris, pos), "cameras_per_sample": cps, "parent_cameras": cams}


def samples_from_meshes(meshes):
    """Samples::addSamplesViaMeshes: one sample per vertex, confidence 1, parent cameras padded
    with INVALID to the largest cameras-per-sample of the meshes"""
    mc = max([m["cameras_per_sample"] for m in meshes] + [1])
    out = {k: [] for k in ("px", "py", "pz", "nx", "ny", "nz", "scale", "confidence", "parent_cameras")}
    for m in meshes:
        P, N = m["positions"], m["normals"]
        n = len(P)
        for a, k in enumerate(("px", "py", "pz")):
            out[k].append(P[:, a])
        for a, k in enumerate(("nx", "ny", "nz")):
            out[k].append(N[:, a])
        out["scale"].append(m["scales"])
        out["confidence"].append(np.ones(n))
        cams = np.full((n, mc), INVALID, np.uint32)
        cams[:, :m["cameras_per_sample"]] = np.asarray(m["parent_cameras"]).reshape(n, -1)
        out["parent_cameras"].append(cams.reshape(-1))
    res = {k: np.concatenate(v).astype(np.float32) if k != "parent_cameras" else np.concatenate(v).astype(np.uint32)
           for k, v in out.items()}
    return res, mc
