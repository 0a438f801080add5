"""What the trained routing network (tsr_amd/culler_weights.npz) gains on a held-out scene
(config 3, seed 1000, 32 cameras at 320x240), against the certified cascade without it and
against the plain evaluation: the stages it routes and the wall-clock ratios (printed; the
test asserts the learned form is no slower than plain -- on these scenes the network's
routing gains nothing over the cascade without it, see profiles/culler_report.json).  The mesh of every form is
asserted equal to the plain one, so the Hausdorff distance is 0, within the 0.1% bound."""
import numpy as np
import pytest

import tsr_amd
from tsr_amd.culling import load_culler, run_modes
from tsr_amd.metrics import bbox_diagonal, hausdorff

pytestmark = pytest.mark.gpu


def test_trained_router_on_held_out_scene(gpu_ctx):
    syn = tsr_amd.SyntheticScene(3, seed=1000, resolution_scale=1.0)
    sc = tsr_amd.Scene(gpu_ctx, syn.samples, syn.cameras)
    tree = tsr_amd.Tree(gpu_ctx, sc)
    culler = load_culler(gpu_ctx)
    r = run_modes(gpu_ctx, tree, sc, culler, reps=3)
    diag = bbox_diagonal(syn.samples)
    pm = r["plain"][1]
    for mode in ("certified", "learned"):
        m = r[mode][1]
        for k in ("positions", "triangles", "edge_keys"):
            assert np.array_equal(m[k], pm[k]), (mode, k)
        assert hausdorff(m, pm, diag)[0] <= 1e-3
    st = r["learned"][4]
    pri = {k: r[k][3][3] for k in r}
    ms = {k: 1e3 * r[k][2] for k in r}
    print(f"\nstages with the network {st}, without {r['certified'][4]}; priors ms {pri}; "
          f"occupancy + DMC ms {ms}; learned vs plain {ms['plain'] / ms['learned']:.2f}x, "
          f"network vs none (priors) {pri['certified'] / pri['learned']:.2f}x")
    assert st["stage0"] == st["queries"] - st["routed_stage1"] - st["routed_exact"]
    assert pri["learned"] < pri["plain"] and pri["certified"] < pri["plain"]
    assert ms["learned"] <= ms["plain"] * 1.02
    culler.close()
    tree.close()
    sc.close()
