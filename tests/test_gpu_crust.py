"""Crust post-processing (DualMarchingCells.cpp:26-39 after extractMesh) through the C ABI
against the oracle's restatement (oracle/crust.py) applied to the same raw mesh: isle
removal (FlexibleMesh.cpp:1411-1426), hole filling (FlexibleMesh.cpp:1580-1932) and
angle-weighted normals (Mesh.cpp:145-202).  Triangles, edge keys and scales equal; positions
equal (the same FP64 operations in the same order); normals within 1e-12 (zero-length sums
stay zero: the reference's Vector3::normalize is not available to pin that case)."""
import numpy as np
import pytest

from tests.helpers import scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,res,min_isle", [(3, 0.25, 20), (1, 0.0625, 20), (3, 0.25, 1000)],
                         ids=["primitives-20", "wedge-20", "primitives-1000"])
def test_crust_postprocess_matches_oracle(gpu_ctx, cfg, res, min_isle):
    from oracle.crust import postprocess
    from tsr_amd import DualMarchingCells, Occupancy, Scene, Tree
    samples, cameras = scene(cfg, res)
    print(f"\n[crust] config={cfg} res={res} min_isle={min_isle} samples={len(samples['px'])}", flush=True)
    sc = Scene(gpu_ctx, samples, cameras)
    tree = Tree(gpu_ctx, sc)
    occ = Occupancy(gpu_ctx, tree, sc)
    mesh = DualMarchingCells(gpu_ctx, tree, occ)
    raw = mesh.download()
    want = postprocess(raw, min_isle)
    mesh.postprocess(min_isle)
    got = mesh.download()
    got["normals"] = mesh.normals()
    assert len(got["scales"]) == len(want["scales"]) and len(got["triangles"]) == len(want["triangles"])
    for k in ("triangles", "edge_keys", "scales", "positions"):
        assert np.array_equal(got[k], want[k]), k
    assert np.allclose(got["normals"], want["normals"], rtol=0, atol=1e-12)
    # unit normals, or zero at vertices whose triangles all have zero area (DMC snaps a
    # vertex onto the negative leaf's centre when |occupancy| <= EPSILON there,
    # DualMarchingCell.cpp:51-71, so vertices sharing that leaf coincide)
    ln = np.linalg.norm(got["normals"].reshape(-1, 3), axis=1)
    assert np.all(np.isclose(ln, 1.0, rtol=0, atol=1e-12) | (ln == 0.0))
    assert np.mean(ln == 0.0) < 0.2
    # the post-processed crust is closed: every edge in two triangles
    t = got["triangles"].reshape(-1, 3).astype(np.int64)
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e[:, 0] * (1 << 32) + e[:, 1], return_counts=True)
    assert set(cnt.tolist()) == {2}
    for x in (mesh, occ, tree, sc):
        x.close()
