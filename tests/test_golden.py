"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py from
the oracle): the oracle still reproduces them exactly, and the synthetic scene
generator still produces their inputs."""
import glob
import os

import numpy as np
import pytest

from tests.helpers import PARAMS, scene

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
CASES = {"sphere_r025": (0, 0.25, 0), "primitives_r005_seed7": (3, 0.05, 7)}


def load(path):
    z = np.load(path, allow_pickle=False)
    samples = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    cameras = {k[4:]: z[k] for k in z.files if k.startswith("cam_")}
    return z, samples, cameras


def test_fixtures_present():
    assert {os.path.basename(p)[:-4] for p in FIXTURES} == set(CASES)


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_generator_reproduces_inputs(path):
    z, samples, cameras = load(path)
    cfg, res, seed = CASES[os.path.basename(path)[:-4]]
    s, c = scene(cfg, res, seed)
    for k, v in samples.items():
        assert np.array_equal(np.asarray(s[k]), v), k
    for k, v in cameras.items():
        assert np.array_equal(np.asarray(c[k]), v), k


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_fixture(path):
    from oracle.oracle import Oracle
    z, samples, cameras = load(path)
    o = Oracle(samples, cameras, PARAMS)
    o.build_tree().occupancy().dmc()
    t, f, m = o.tree(), o.field(), o.mesh()
    for k in ("children", "parents", "sample_counts", "sample_starts", "node_to_leaf", "leaf_nodes",
              "leaf_scopes", "cells", "sample_target"):
        assert np.array_equal(t[k], z["tree_" + k]), k
    for k in ("kernel_sums", "cone_lengths", "priors", "node_states", "occupancy"):
        assert np.array_equal(f[k], z["field_" + k], equal_nan=True), k
    assert f["cone_count"] == z["field_cone_count"]
    for k in ("positions", "scales", "edge_keys", "triangles"):
        assert np.array_equal(m[k], z["mesh_" + k]), k
