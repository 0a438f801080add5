"""Writes the golden fixtures of the oracle: small seeded scenes (inputs) with the
oracle's tree tables, FP64 field and DMC mesh (expected outputs).

    python -m tests.golden.make_golden

The reference ships no fixtures for this path and cannot be built here, so the
fixtures come from the oracle, which the CPU tests pin independently (closed
forms, brute force, invariants).  Committed so that the GPU path and later
oracle versions are checked against the same fixed data.
"""
import os

import numpy as np

from oracle.oracle import Oracle
from tests.helpers import PARAMS, scene

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {"sphere_r025": (0, 0.25, 0), "primitives_r005_seed7": (3, 0.05, 7)}


def make(name, config, res, seed):
    s, c = scene(config, res, seed)
    o = Oracle(s, c, PARAMS)
    o.build_tree().occupancy().dmc()
    t, f, m = o.tree(), o.field(), o.mesh()
    out = {f"in_{k}": np.asarray(v) for k, v in s.items()}
    out.update({f"cam_{k}": np.asarray(v) for k, v in c.items()})
    for k in ("children", "parents", "sample_counts", "sample_starts", "node_to_leaf", "leaf_nodes",
              "leaf_scopes", "cells", "sample_target"):
        out[f"tree_{k}"] = t[k]
    out["tree_root_min"] = t["root_min"]
    out["tree_root_size"] = np.float32(t["root_size"])
    for k in ("kernel_sums", "cone_lengths", "priors", "node_states", "occupancy"):
        out[f"field_{k}"] = f[k]
    out["field_cone_count"] = np.float64(f["cone_count"])
    for k in ("positions", "scales", "edge_keys", "triangles"):
        out[f"mesh_{k}"] = m[k]
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return path


if __name__ == "__main__":
    for name, (cfg, res, seed) in CASES.items():
        p = make(name, cfg, res, seed)
        print(p, os.path.getsize(p))
