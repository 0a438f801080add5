"""What the trained culler (tsr_amd/culler_weights.npz) gains over the plain
method on a held-out scene (config 3, seed 1000, 32 cameras at 320x240): the
emptiness work it removes, what it costs in mesh change, and the wall-clock
ratio (printed, not asserted: the feature pass costs more than the culled work
on these scenes)."""
import time

import numpy as np
import pytest

import tsr_amd
from tsr_amd import _lib as L
from tsr_amd.culling import hausdorff, load_culler

pytestmark = pytest.mark.gpu


@pytest.mark.xfail(reason="the node culler moves the mesh by more than 0.1% of the bbox diagonal "
                          "on held-out scenes (see tsr_amd/culler_report.json)", strict=False)
def test_trained_culler_on_held_out_scene(gpu_ctx):
    syn = tsr_amd.SyntheticScene(3, seed=1000, resolution_scale=1.0)
    sc = tsr_amd.Scene(gpu_ctx, syn.samples, syn.cameras)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    culler = load_culler(gpu_ctx)
    runs = {}
    for mode in ("plain", "learned", "plain", "learned"):   # second round timed warm
        gpu_ctx.synchronize()
        t0 = time.perf_counter()
        if mode == "plain":
            occ = tsr_amd.Occupancy(gpu_ctx, tree, sc)
        else:
            occ = tsr_amd.Occupancy(gpu_ctx, tree, sc, flags=L.TSR_CULL_LEARNED, culler=culler)
        mesh = tsr_amd.DualMarchingCells(gpu_ctx, tree, occ)
        gpu_ctx.synchronize()
        dt = time.perf_counter() - t0
        if mode in runs:
            runs[mode][0].close()
        runs[mode] = (occ, mesh.download(), dt, occ.hit_counts())
        mesh.close()
    (po, pm, pdt, ph), (lo, lm, ldt, lh) = runs["plain"], runs["learned"]
    labels = tsr_amd.node_labels(gpu_ctx, tree, po)
    leaves = tree.download()["leaf_nodes"]
    culled = lo.culled_leaves() == 1
    fp, fl = po.download(), lo.download()
    missed = int(np.sum(culled & (labels[leaves] == 1)))
    flips = int(np.sum((fp["occupancy"] > 0) != (fl["occupancy"] > 0)))
    P = pm["positions"].reshape(-1, 3)
    diag = float(np.linalg.norm(P.max(0) - P.min(0)))
    h = hausdorff(pm, lm) / diag
    print(f"\nculled leaves {culled.sum()} of {tree.leaf_count}, emptiness hits {lh[0]} vs {ph[0]} "
          f"({1 - lh[0] / ph[0]:.4%} removed), missed surface leaves {missed}, sign flips {flips}, "
          f"Hausdorff {h:.2e} of the diagonal, plain {1e3 * pdt:.1f} ms, learned {1e3 * ldt:.1f} ms "
          f"(speed ratio {pdt / ldt:.2f})")
    assert culled.sum() > 0 and lh[0] < ph[0]
    # the culled leaves are (almost) all far-empty; the few sign flips come from priors
    # reading the emptiness cone lengths that culled leaves no longer receive
    assert missed <= 0.001 * np.sum(labels[leaves] == 1)
    assert flips <= 1e-4 * tree.leaf_count
    assert h <= 1e-3   # 0.1 % of the bbox diagonal
    for o in (po, lo):
        o.close()
    culler.close()
    tree.close()
    sc.close()
