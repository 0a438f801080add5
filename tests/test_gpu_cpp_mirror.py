"""The C++ class mirror (include/tsr_gpu.hpp) driven like Scene::reconstruct
(Scene.cpp:169-197) by tests/cpp/pipeline_check: its own accessor checks pass and
its counts and checksums equal those of the same run through the Python binding."""
import json
import os
import subprocess

import numpy as np
import pytest

import tsr_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "pipeline_check")


def _fnv(a):
    h = 1469598103934665603
    for b in np.ascontiguousarray(a).view(np.uint8).tobytes():
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_cpp_mirror_compiles_and_links():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(BIN)


@pytest.mark.gpu
@pytest.mark.parametrize("config,res", [(0, 0.5), (3, 0.125)])
def test_cpp_mirror_matches_python_binding(gpu_ctx, config, res, tmp_path):
    r = subprocess.run([BIN, str(config), str(res), "0", str(tmp_path / "scene")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    assert got["failures"] == 0
    syn = tsr_amd.SyntheticScene(config, seed=0, resolution_scale=res)
    scene, tree, occ, mesh = tsr_amd.reconstruct(gpu_ctx, syn.samples, syn.cameras,
                                                 flags=tsr_amd._lib.TSR_PRIORS_ALL_LEAVES)
    t = tree.download()
    f = occ.download()
    m = mesh.postprocess(1000).download()   # DualMarchingCells(occupancy, 1000)
    owner = np.repeat(np.arange(tree.node_count), t["sample_counts"])
    depth_sum = int(t["node_depth"][owner[t["sample_target"]]].sum())
    pos = f["occupancy"][f["occupancy"] > 0]
    expect = {
        "samples": syn.count, "nodes": tree.node_count, "leaves": tree.leaf_count, "cells": tree.cell_count,
        "max_depth": tree.max_depth, "sample_depth_sum": depth_sum, "positive_leaves": int(len(pos)),
        "cone_count": f["cone_count"], "vertices": mesh.vertex_count, "triangles": mesh.triangle_count,
        "cells_hash": _fnv(t["cells"]), "triangles_hash": _fnv(m["triangles"]),
        "positions_hash": _fnv(m["positions"]), "finite_priors": int(np.isfinite(f["priors"]).sum()),
        "neighbour_entries": int(len(tree.neighbors()[1])),
    }
    for k, v in expect.items():
        assert got[k] == v, k
    s = 0.0
    for v in pos:        # the C++ loop's order
        s += float(v)
    assert got["positive_sum"] == s
    for h in (mesh, occ, tree, scene):
        h.close()
