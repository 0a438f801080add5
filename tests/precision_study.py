"""Precision study on the thin wedge + sphere scene: the GPU field with FP32
Kahan sums in a deterministic (sorted) order, the GPU field with plain FP32
atomics, and the FP64 oracle on the same samples.

Reported per accumulation mode: per-leaf relative occupancy error against the
oracle (|occ - occ64| / (pS + pE) of the oracle, max and p99), the number of
sign flips, and the symmetric Hausdorff distance between the GPU mesh and the
oracle mesh as a fraction of the oracle mesh's bounding-box diagonal.

    python -m tests.precision_study --res 0.25 --out precision.json

Test infrastructure: it runs the oracle, so it lives with the tests.
"""
import argparse
import json
import time

import numpy as np

import tsr_amd
from tsr_amd import _lib as L
from tsr_amd.metrics import hausdorff
from tests.helpers import PARAMS, occupancy_rel_err, scene


def study(ctx, res, seed=0, config=1):
    from oracle.oracle import Oracle
    samples, cameras = scene(config, res, seed)
    t0 = time.perf_counter()
    orc = Oracle(samples, cameras, PARAMS)
    orc.build_tree().occupancy().dmc()
    oracle_s = time.perf_counter() - t0
    of, om = orc.field(), orc.mesh()
    P = om["positions"].reshape(-1, 3)
    diag = float(np.linalg.norm(P.max(0) - P.min(0))) if len(P) else 1.0
    sc = tsr_amd.Scene(ctx, samples, cameras)
    tree = tsr_amd.Tree(ctx, sc)
    out = {"config": config, "resolution_scale": res, "seed": seed, "samples": int(len(samples["px"])),
           "leaves": tree.leaf_count, "oracle_seconds": oracle_s, "oracle_triangles": int(len(om["triangles"]) // 3),
           "bbox_diagonal": diag, "modes": {}}
    for name, flags in (("fp32_kahan_sorted", L.TSR_ACCUM_KAHAN_SEGMENTED | L.TSR_PRIORS_ALL_LEAVES),
                        ("fp32_atomic", L.TSR_ACCUM_FP32_ATOMIC | L.TSR_PRIORS_ALL_LEAVES)):
        occ = tsr_amd.Occupancy(ctx, tree, sc, flags=flags)
        f = occ.download()
        mesh = tsr_amd.DualMarchingCells(ctx, tree, occ)
        m = mesh.download()
        err, flips = occupancy_rel_err(f["occupancy"], of)
        Lc = tree.leaf_count
        ks = f["kernel_sums"].astype(np.float64)
        oks = of["kernel_sums"]
        with np.errstate(divide="ignore", invalid="ignore"):
            kerr = np.where(oks > 0, np.abs(ks - oks) / oks, np.where(ks == oks, 0.0, np.inf))
        out["modes"][name] = {
            "occupancy_rel_err_max": float(err.max()), "occupancy_rel_err_p99": float(np.percentile(err, 99)),
            "kernel_sum_rel_err_max": float(kerr.max()), "kernel_sum_rel_err_p99": float(np.percentile(kerr, 99)),
            "sign_flips": int(flips), "triangles": int(len(m["triangles"]) // 3),
            "hausdorff_rel": hausdorff(m, om, diag)[0],
        }
        mesh.close()
        occ.close()
    tree.close()
    sc.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--res", type=float, default=0.25)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    ctx = tsr_amd.Context(0)
    r = study(ctx, args.res, args.seed)
    print(json.dumps(r), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(r, fh, indent=1)
    ctx.close()


if __name__ == "__main__":
    main()
