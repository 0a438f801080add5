"""Mesh and field comparison measures reported next to every result.

hausdorff(): distance between two crust meshes as a fraction of a bounding-box
diagonal.  When both meshes have the same triangles over the same vertex keys
(EdgeVertexIndex pairs, EdgeVertexIndex.cpp:15-33), corresponding points of
corresponding triangles are the same barycentric blend of the vertices, so the
largest vertex displacement bounds the Hausdorff distance from above and is
reported exactly.  Otherwise the symmetric vertex-to-nearest-vertex distance is
reported (an upper bound of the vertex-sampled point-to-surface distance).
"""
import numpy as np


def _same_topology(a, b):
    return (len(a["triangles"]) == len(b["triangles"]) and len(a["edge_keys"]) == len(b["edge_keys"])
            and np.array_equal(a["triangles"], b["triangles"]) and np.array_equal(a["edge_keys"], b["edge_keys"]))


def hausdorff(a, b, diagonal):
    """(distance / diagonal, method) for two meshes given as download() dicts."""
    pa = np.asarray(a["positions"], np.float64).reshape(-1, 3)
    pb = np.asarray(b["positions"], np.float64).reshape(-1, 3)
    if len(pa) == 0 and len(pb) == 0:
        return 0.0, "empty"
    if len(pa) == 0 or len(pb) == 0:
        return float("inf"), "one mesh empty"
    if _same_topology(a, b):
        d = float(np.max(np.linalg.norm(pa - pb, axis=1)))
        return d / diagonal, "same topology: max vertex displacement"
    from scipy.spatial import cKDTree
    dab = cKDTree(pb).query(pa)[0].max()
    dba = cKDTree(pa).query(pb)[0].max()
    return float(max(dab, dba)) / diagonal, "vertex sets: symmetric nearest-vertex distance"


def bbox_diagonal(samples):
    p = np.stack([samples["px"], samples["py"], samples["pz"]], 1).astype(np.float64)
    return float(np.linalg.norm(p.max(0) - p.min(0)))


def field_error(occ, kernel_sums, ref_field):
    """Per-leaf relative errors against an FP64 field: occupancy as |occ - occ64| / (pS + pE)
    of the FP64 field (Occupancy.cpp:570-587), kernel sums as |k - k64| / |k64|; and the
    number of leaves whose sign (occupancy > 0) differs."""
    f = ref_field
    L = len(f["occupancy"])
    E, S = f["kernel_sums"][:L], f["kernel_sums"][L:]
    M = f["cone_count"]
    pr = f["priors"]
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = np.where(np.isfinite(pr), pr * E / M + (1 - pr) * S / M, 0.0)
    o = f["occupancy"]
    err = np.abs(np.asarray(occ, np.float64) - o) / np.maximum(scale, 1e-300)
    err[(o < 0) & (np.asarray(occ) == o)] = 0.0
    ks = np.asarray(f["kernel_sums"], np.float64)
    kr = np.abs(np.asarray(kernel_sums, np.float64) - ks) / np.maximum(np.abs(ks), 1e-30)
    kr[(ks == 0) & (np.asarray(kernel_sums) == 0)] = 0.0
    flips = int(np.sum((np.asarray(occ) > 0) != (o > 0)))
    return {"occupancy_rel_err_max": float(err.max()) if L else 0.0,
            "occupancy_rel_err_p99": float(np.percentile(err, 99)) if L else 0.0,
            "kernel_sum_rel_err_max": float(kr.max()) if len(kr) else 0.0,
            "kernel_sum_rel_err_p99": float(np.percentile(kr, 99)) if len(kr) else 0.0,
            "sign_flips": flips}
