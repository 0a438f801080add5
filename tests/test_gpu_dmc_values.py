"""tsr_dmc_extract_values: DMC over an explicit per-leaf occupancy.  With the GPU field's
own occupancy it gives the mesh of tsr_dmc_extract bit for bit; with the FP64 oracle's
occupancy it gives the oracle's mesh (DualMarchingCells.cpp:41-169)."""
import numpy as np
import pytest

from tests.helpers import PARAMS, scene

pytestmark = pytest.mark.gpu


def test_dmc_over_given_values(gpu_ctx):
    from oracle.oracle import Oracle
    from tsr_amd import DualMarchingCells, Occupancy, Scene, Tree
    from tsr_amd.api import dmc_from_values
    samples, cameras = scene(3, 0.125)
    print(f"\n[dmc values] samples={len(samples['px'])}", flush=True)
    sc = Scene(gpu_ctx, samples, cameras)
    tree = Tree(gpu_ctx, sc)
    occ = Occupancy(gpu_ctx, tree, sc)
    ref = DualMarchingCells(gpu_ctx, tree, occ).download()
    mine = dmc_from_values(gpu_ctx, tree, occ.download()["occupancy"]).download()
    for k in ("positions", "scales", "edge_keys", "triangles"):
        assert np.array_equal(mine[k], ref[k]), k
    orc = Oracle(samples, cameras, PARAMS).build_tree().occupancy()
    of = orc.field()
    orc.dmc()
    om = orc.mesh()
    m2 = dmc_from_values(gpu_ctx, tree, of["occupancy"]).download()
    assert np.array_equal(m2["triangles"], om["triangles"])
    assert np.array_equal(m2["edge_keys"], om["edge_keys"])
    assert np.array_equal(m2["positions"], om["positions"])
    for x in (occ, tree, sc):
        x.close()
