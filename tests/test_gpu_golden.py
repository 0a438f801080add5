"""The GPU path through the C ABI on the committed golden fixtures: tree tables
bit for bit, field within the parity bound, the identical DMC mesh."""
import glob
import os

import numpy as np
import pytest

from tests.helpers import occupancy_rel_err, rel_err
from tests.test_golden import FIXTURES, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_gpu_matches_golden(gpu_ctx, path):
    from tsr_amd import DualMarchingCells, Occupancy, Scene, Tree
    from tsr_amd._lib import TSR_PRIORS_ALL_LEAVES
    z, samples, cameras = load(path)
    sc = Scene(gpu_ctx, samples, cameras)
    tree = Tree(gpu_ctx, sc)
    occ = Occupancy(gpu_ctx, tree, sc, flags=TSR_PRIORS_ALL_LEAVES)
    mesh = DualMarchingCells(gpu_ctx, tree, occ)
    t, f, m = tree.download(), occ.download(), mesh.download()
    for k in ("children", "parents", "sample_counts", "sample_starts", "node_to_leaf", "leaf_nodes",
              "leaf_scopes", "cells", "sample_target"):
        assert np.array_equal(t[k], z["tree_" + k]), k
    assert np.array_equal(f["node_states"], z["field_node_states"])
    for k in ("kernel_sums", "cone_lengths"):
        assert np.max(rel_err(f[k], z["field_" + k], floor=1e-30), initial=0) <= 1e-5, k
    of = {k[6:]: z[k] for k in z.files if k.startswith("field_")}
    err, flips = occupancy_rel_err(f["occupancy"], of)
    assert np.max(err, initial=0) <= 1e-5 and flips == 0
    assert np.array_equal(m["triangles"], z["mesh_triangles"])
    assert np.array_equal(m["edge_keys"], z["mesh_edge_keys"])
    assert np.allclose(m["positions"], z["mesh_positions"], rtol=0, atol=1e-9)
    for h in (mesh, occ, tree, sc):
        h.close()
