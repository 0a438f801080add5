"""Training and evaluation of the routing network of the certified priors.

No counterpart in the reference.  The priors (Occupancy.cpp:530-568) are the largest
single kernel of the occupancy field.  The certified cascade (TSR_PRIORS_CERTIFIED)
bounds each queried leaf's prior with a depth-capped sphere traversal (stage 0: cap 3
levels above the leaf, stage 1: 2 levels above) and keeps the leaves whose occupancy
class the interval decides, which is all the mesh reads of them; stage 2 computes the
exact prior of the rest.  The DMC mesh equals the plain one by construction.

The network (a 16-64-64-64-1 MLP on per-leaf features) predicts the stage that decides a
leaf, so that leaves a bound would not decide skip that bound (TSR_CULL_LEARNED).  It
never changes a result, only the work: its gain is measured against the cascade without
it and against the plain evaluation.

Training data come from certified fields of seeded synthetic scenes (config 3: unions of
3-8 spheres, boxes and thin sheets, 32 cameras, 320x240): features per queried leaf and
the stage that decided it (target stage / 2).

    python -m tsr_amd.culling --train 0:60 --holdout 1000:1020 --out tsr_amd/culler_weights.npz \\
        --report profiles/culler_report.json

Everything heavy runs on the GPU through the library: scenes, fields, features, stage
labels, MLP training (BF16 MFMA forward/backward + Adam) and the evaluation runs.  Row
subsampling, feature standardisation (folded into the first layer afterwards) and the
comparisons run on the host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import _lib as L
from .api import (MLP_PARAM_COUNT, Context, Culler, DualMarchingCells, Occupancy, Scene, SyntheticScene, Tree,
                  init_mlp_weights, leaf_features, mlp_train)
from .metrics import bbox_diagonal, hausdorff

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "culler_weights.npz")
SCENE_CONFIG = 3
MODES = {"plain": L.TSR_ACCUM_KAHAN_SEGMENTED, "certified": L.TSR_PRIORS_CERTIFIED, "learned": L.TSR_CULL_LEARNED}


def load_culler(ctx, path=DEFAULT_WEIGHTS, threshold=None):
    """Culler with committed weights (and their tuned threshold unless given)."""
    z = np.load(path, allow_pickle=False)
    t = float(z["threshold"]) if threshold is None else threshold
    return Culler(ctx, z["weights"], t)


def scene_rows(ctx, seed, res, rng, max_rows, config=SCENE_CONFIG):
    """Features and stage targets of the queried leaves of one certified field; returns
    (X, y, stage counts, samples)"""
    syn = SyntheticScene(config, seed=seed, resolution_scale=res)
    sc = Scene(ctx, syn.samples, syn.cameras)
    tree = Tree(ctx, sc)
    occ = Occupancy(ctx, tree, sc, flags=L.TSR_PRIORS_CERTIFIED)
    X = leaf_features(ctx, tree, occ)
    lev = occ.prior_levels()
    occ.close()
    tree.close()
    sc.close()
    q = np.flatnonzero(lev != 255)
    idx = q if len(q) <= max_rows else rng.choice(q, max_rows, replace=False)
    counts = np.bincount(lev[q], minlength=3)[:3]
    return X[idx], (lev[idx] / 2.0).astype(np.float32), counts, syn.count


def fold_standardisation(w, mean, std):
    """weights for raw features from weights trained on (x - mean) / std"""
    w = np.array(w, np.float64)
    W1 = w[:64 * 16].reshape(64, 16)
    b1 = w[64 * 16:64 * 16 + 64]
    W1f = W1 / std[None, :]
    b1f = b1 - W1f @ mean
    w[:64 * 16] = W1f.reshape(-1)
    w[64 * 16:64 * 16 + 64] = b1f
    return w.astype(np.float32)


def run_modes(ctx, tree, sc, culler, modes=("plain", "certified", "learned"), reps=2):
    """Occupancy + DMC per mode on one tree, the faster of `reps` runs: (field download, mesh,
    seconds, phase ms, prior stages) per mode"""
    out = {}
    for mode in modes:
        best = None
        for _ in range(reps):   # the first run may pay one-time set-up
            ctx.synchronize()
            t0 = time.perf_counter()
            occ = Occupancy(ctx, tree, sc, flags=MODES[mode], culler=culler if mode == "learned" else None)
            mesh = DualMarchingCells(ctx, tree, occ)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            if best is None or dt < best[2]:
                best = (occ.download(), mesh.download(), dt, [float(x) for x in ctx.phase_times()],
                        occ.prior_stages())
            occ.close()
            mesh.close()
        out[mode] = best
    return out


def compare(plain, other, diagonal):
    """mesh and field of a certified run against the plain one"""
    (pf, pm, pdt, pph, _), (f, m, dt, ph, stages) = plain, other
    same = all(np.array_equal(pm[k], m[k]) for k in ("positions", "triangles", "edge_keys"))
    h, how = hausdorff(pm, m, diagonal)
    po, o = pf["occupancy"], f["occupancy"]
    cls = lambda v: np.where(v < 0, 0, np.where(v <= np.float64(np.float32(1e-6)), 1, 2))
    return {"ms": 1e3 * dt, "priors_ms": ph[3], "phase_ms": ph[1:6], "stages": stages,
            "mesh_identical": bool(same), "hausdorff_rel": h, "hausdorff_method": how,
            "class_changes": int(np.sum(cls(po) != cls(o))),
            "occupancy_max_abs_change": float(np.max(np.abs(po - o))) if len(po) else 0.0,
            "speedup_vs_plain": pdt / dt}


def evaluate(ctx, culler, seed, res, config=SCENE_CONFIG):
    """plain, certified and learned on one scene"""
    syn = SyntheticScene(config, seed=seed, resolution_scale=res)
    sc = Scene(ctx, syn.samples, syn.cameras)
    tree = Tree(ctx, sc)
    diag = bbox_diagonal(syn.samples)
    r = run_modes(ctx, tree, sc, culler)
    out = {"config": config, "seed": seed, "samples": syn.count, "leaves": tree.leaf_count,
           "plain_ms": 1e3 * r["plain"][2], "plain_priors_ms": r["plain"][3][3],
           "certified": compare(r["plain"], r["certified"], diag), "learned": compare(r["plain"], r["learned"], diag)}
    out["learned_vs_certified_priors"] = out["certified"]["priors_ms"] / max(out["learned"]["priors_ms"], 1e-9)
    tree.close()
    sc.close()
    return out


def _seeds(s):
    a, b = s.split(":")
    return list(range(int(a), int(b)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--train", default="0:60", help="training seeds a:b")
    ap.add_argument("--holdout", default="1000:1020", help="held-out seeds a:b")
    ap.add_argument("--res", type=float, default=1.0, help="resolution scale of the 320x240 depth maps")
    ap.add_argument("--rows", type=int, default=60000, help="queried leaves kept per training scene")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--thresholds", default="0.75,1.0,1.25,1.5", help="candidates of the routing threshold t")
    ap.add_argument("--tune", type=int, default=6, help="training scenes re-run to choose the threshold")
    ap.add_argument("--max-hausdorff", type=float, default=1e-3, help="bound, fraction of the bbox diagonal")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--report", default=None, help="JSON report path")
    ap.add_argument("--eval-only", action="store_true",
                    help="no training: evaluate the weights in --out on --holdout seeds of --config")
    ap.add_argument("--config", type=int, default=SCENE_CONFIG, help="scene config of --eval-only")
    args = ap.parse_args(argv)
    ctx = Context(0)
    if args.eval_only:
        culler = load_culler(ctx, args.out)
        rs = []
        for seed in _seeds(args.holdout):
            r = evaluate(ctx, culler, seed, args.res, args.config)
            print(json.dumps(r), flush=True)
            rs.append(r)
        if args.report:
            with open(args.report, "w") as f:
                json.dump(rs, f, indent=1)
        culler.close()
        ctx.close()
        bad = [r["seed"] for r in rs for m in ("certified", "learned") if r[m]["hausdorff_rel"] > args.max_hausdorff]
        return 1 if bad else 0

    rng = np.random.default_rng(0)
    Xs, ys, counts = [], [], np.zeros(3, np.int64)
    t0 = time.perf_counter()
    for k, seed in enumerate(_seeds(args.train)):
        X, y, c, n = scene_rows(ctx, seed, args.res, rng, args.rows)
        Xs.append(X)
        ys.append(y)
        counts += c
        if k % 10 == 0:
            print(f"[culling] scene {seed}: {n} samples, stages {c.tolist()} ({time.perf_counter() - t0:.0f}s)",
                  flush=True)
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    w = np.ones(len(y), np.float32)
    mean, std = X.mean(0).astype(np.float64), X.std(0).astype(np.float64)
    std = np.where(std > 1e-6, std, 1.0)
    Xn = ((X - mean) / std).astype(np.float32)
    print(f"[culling] {len(X)} rows, stage shares {(counts / counts.sum()).round(4).tolist()}; "
          f"training {args.epochs} epochs", flush=True)
    w0 = init_mlp_weights(0)
    wt, losses = mlp_train(ctx, w0, Xn, y, w, epochs=args.epochs, batch=args.batch, learning_rate=args.lr, seed=1)
    print("[culling] loss per epoch " + " ".join(f"{v:.4f}" for v in losses), flush=True)
    weights = fold_standardisation(wt, mean, std)
    assert weights.size == MLP_PARAM_COUNT

    # threshold: the candidate with the least priors time on the tuning scenes
    tune = _seeds(args.train)[:args.tune]
    culler = Culler(ctx, weights, 0.5)
    tuning = []
    for thr in sorted(float(t) for t in args.thresholds.split(",")):
        L.check(L.lib().tsr_culler_set_weights(ctx.h, culler.h, weights.ctypes.data_as(L.f32p), ctypes.c_float(thr)))
        ms = []
        for s in tune:
            syn = SyntheticScene(SCENE_CONFIG, seed=s, resolution_scale=args.res)
            sc = Scene(ctx, syn.samples, syn.cameras)
            tree = Tree(ctx, sc)
            ms.append(run_modes(ctx, tree, sc, culler, modes=("learned",))["learned"][3][3])
            tree.close()
            sc.close()
        tuning.append({"threshold": thr, "mean_priors_ms": float(np.mean(ms))})
        print(f"[culling] threshold {thr}: priors {np.mean(ms):.2f} ms", flush=True)
    chosen = min(tuning, key=lambda t: t["mean_priors_ms"])["threshold"]
    L.check(L.lib().tsr_culler_set_weights(ctx.h, culler.h, weights.ctypes.data_as(L.f32p), ctypes.c_float(chosen)))
    held = []
    for s in _seeds(args.holdout):
        held.append(evaluate(ctx, culler, s, args.res))
        h = held[-1]
        print(f"[culling] held-out {s}: plain {h['plain_ms']:.1f} ms, certified {h['certified']['ms']:.1f} ms, "
              f"learned {h['learned']['ms']:.1f} ms; priors {h['plain_priors_ms']:.2f} / "
              f"{h['certified']['priors_ms']:.2f} / {h['learned']['priors_ms']:.2f} ms; meshes identical "
              f"{h['certified']['mesh_identical'] and h['learned']['mesh_identical']}", flush=True)
    worst = max(r[m]["hausdorff_rel"] for r in held for m in ("certified", "learned"))
    summary = {
        "threshold": chosen,
        "train_seeds": args.train, "holdout_seeds": args.holdout, "resolution_scale": args.res,
        "rows": int(len(X)), "stage_shares": (counts / counts.sum()).tolist(), "epochs": args.epochs,
        "loss_per_epoch": [float(v) for v in losses], "tuning": tuning,
        "holdout_worst_hausdorff_rel": worst,
        "holdout_meshes_identical": all(r[m]["mesh_identical"] for r in held for m in ("certified", "learned")),
        "holdout_mean_speedup_learned_vs_plain": float(np.mean([r["learned"]["speedup_vs_plain"] for r in held])),
        "holdout_mean_speedup_certified_vs_plain": float(np.mean([r["certified"]["speedup_vs_plain"] for r in held])),
        "holdout_mean_priors_ms": {m: float(np.mean([r[m]["priors_ms"] for r in held])) for m in ("certified", "learned")}
        | {"plain": float(np.mean([r["plain_priors_ms"] for r in held]))},
        "holdout": held,
    }
    print(json.dumps({k: v for k, v in summary.items() if k != "holdout"}), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(summary, f, indent=1)
    culler.close()
    ctx.close()
    if worst > args.max_hausdorff:
        print(f"[culling] a held-out mesh is {worst:.2e} of the diagonal from the plain one; no weights written",
              file=sys.stderr, flush=True)
        return 1
    np.savez(args.out, weights=weights, threshold=np.float32(chosen), feature_mean=mean, feature_std=std)
    return 0


if __name__ == "__main__":
    sys.exit(main())
