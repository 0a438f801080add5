"""Training and evaluation of the learned node culler.

The culler has no counterpart in the reference: it predicts, per octree node,
whether the node's subtree can hold part of the surface (a leaf with positive
occupancy or a corner of a dual cell whose corner signs differ).  Nodes it
rejects skip emptiness-cone evaluation and priors, and their leaves become
empty (-EPSILON), so it may only reject far-empty space; every mesh it changes
is measured against the plain field's mesh.

Training data come from plain fields of seeded synthetic scenes (config 3:
unions of 3-8 spheres, boxes and thin sheets, 32 cameras, 320x240):

    python -m tsr_amd.culling --train 0:200 --holdout 1000:1020 --out tsr_amd/culler_weights.npz

Everything heavy runs on the GPU through the library: scenes, plain fields,
features, labels, MLP training (BF16 MFMA forward/backward + Adam) and the
evaluation runs.  Only the candidate-row subsampling, the feature
standardisation (folded into the first layer afterwards) and the mesh distance
run on the host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib as L
from .api import (MLP_PARAM_COUNT, Context, Culler, DualMarchingCells, Occupancy, Scene, SyntheticScene, Tree,
                  init_mlp_weights, mlp_train, node_features, node_labels)

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "culler_weights.npz")
SCENE_CONFIG = 3
MIN_DEPTH = 2          # the device never culls the root or its children


def load_culler(ctx, path=DEFAULT_WEIGHTS, threshold=None):
    """Culler with committed weights (and their tuned threshold unless given)."""
    z = np.load(path, allow_pickle=False)
    t = float(z["threshold"]) if threshold is None else threshold
    return Culler(ctx, z["weights"], t)


def scene_rows(ctx, seed, res, rng, max_rows):
    """Features and labels of the candidate nodes (sampleness flag set, depth >= 2)
    of one plain field; returns (X, y, class counts)"""
    syn = SyntheticScene(SCENE_CONFIG, seed=seed, resolution_scale=res)
    sc = Scene(ctx, syn.samples, syn.cameras)
    tree = Tree(ctx, sc)
    occ = Occupancy(ctx, tree, sc)
    X = node_features(ctx, tree, sc, occ)
    y = node_labels(ctx, tree, occ)
    occ.close()
    tree.close()
    sc.close()
    cand = np.nonzero((X[:, 15] > 0.5) & (X[:, 0] * 19.0 >= MIN_DEPTH - 0.5))[0]
    pos, neg = cand[y[cand] == 1], cand[y[cand] == 0]
    half = max_rows // 2
    ps = pos if len(pos) <= half else rng.choice(pos, half, replace=False)
    ns = neg if len(neg) <= max_rows - len(ps) else rng.choice(neg, max_rows - len(ps), replace=False)
    idx = np.concatenate([ps, ns])
    # row weights restore each class's share of the scene
    w = np.concatenate([np.full(len(ps), len(pos) / max(len(ps), 1)), np.full(len(ns), len(neg) / max(len(ns), 1))])
    return X[idx], y[idx].astype(np.float32), w, (len(pos), len(neg), syn.count)


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


def mesh_points(mesh, per_edge=None):
    """vertices plus a barycentric grid of points on every triangle (3 per edge; vertices
    only for meshes above 2M triangles)"""
    P = mesh["positions"].reshape(-1, 3)
    T = mesh["triangles"].reshape(-1, 3).astype(np.int64)
    if per_edge is None:
        per_edge = 3 if len(T) <= 2_000_000 else 1
    if len(T) == 0 or per_edge <= 1:
        return P
    pts = [P]
    a, b, c = P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]
    for i in range(per_edge + 1):
        for j in range(per_edge + 1 - i):
            u, v = i / per_edge, j / per_edge
            if (i, j) in ((0, 0), (per_edge, 0), (0, per_edge)):
                continue
            pts.append(a * (1 - u - v) + b * u + c * v)
    return np.concatenate(pts)


def hausdorff(mesh_a, mesh_b):
    """symmetric Hausdorff distance between two meshes, approximated on dense
    surface samples (overestimates by at most the sample spacing)"""
    from scipy.spatial import cKDTree
    pa, pb = mesh_points(mesh_a), mesh_points(mesh_b)
    if len(pa) == 0 or len(pb) == 0:
        return 0.0 if len(pa) == len(pb) else float("inf")
    dab = cKDTree(pb).query(pa, workers=-1)[0].max()
    dba = cKDTree(pa).query(pb, workers=-1)[0].max()
    return float(max(dab, dba))


def evaluate(ctx, culler, seed, res, config=SCENE_CONFIG):
    """plain vs learned on one scene: times, work, misses, sign flips, mesh distance"""
    syn = SyntheticScene(config, seed=seed, resolution_scale=res)
    sc = Scene(ctx, syn.samples, syn.cameras)
    tree = Tree(ctx, sc)
    out = {"config": config, "seed": seed, "samples": syn.count, "nodes": tree.node_count,
           "leaves": tree.leaf_count}
    res_ = {}
    for mode in ("plain", "learned"):
        best = None
        for rep in range(2):   # the faster of two runs (the first may pay one-time set-up)
            ctx.synchronize()
            t0 = time.perf_counter()
            if mode == "plain":
                occ = Occupancy(ctx, tree, sc)
            else:
                occ = Occupancy(ctx, tree, sc, flags=L.TSR_CULL_LEARNED, culler=culler)
            mesh = DualMarchingCells(ctx, tree, occ)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            if best is None or dt < best[2]:
                if best is not None:
                    best[0].close()
                best = (occ, mesh.download(), dt, ctx.phase_times(), occ.hit_counts())
            else:
                occ.close()
            mesh.close()
        res_[mode] = best
    (po, pm, pdt, pph, phits), (lo, lm, ldt, lph, lhits) = res_["plain"], res_["learned"]
    lab = node_labels(ctx, tree, po)
    t = tree.download()
    fp, fl = po.download(), lo.download()
    leaves = t["leaf_nodes"]
    culled_leaf = lo.culled_leaves() == 1
    pos_leaf = lab[leaves] == 1
    nodes_c, leaves_c = lo.cull_counts()
    P = pm["positions"].reshape(-1, 3)
    diag = float(np.linalg.norm(P.max(0) - P.min(0))) if len(P) else 1.0
    h = hausdorff(pm, lm)
    out.update({
        "plain_ms": 1e3 * pdt, "learned_ms": 1e3 * ldt,
        "plain_phase_ms": [float(x) for x in pph[1:5]], "learned_phase_ms": [float(x) for x in lph[1:5]],
        "emptiness_hits_plain": int(phits[0]), "emptiness_hits_learned": int(lhits[0]),
        "culled_nodes": nodes_c, "culled_leaves": leaves_c,
        "positive_leaves": int(pos_leaf.sum()),
        "missed_leaves": int(np.sum(pos_leaf & culled_leaf)),
        "sign_flips": int(np.sum((fp["occupancy"] > 0) != (fl["occupancy"] > 0))),
        "triangles_plain": int(len(pm["triangles"]) // 3), "triangles_learned": int(len(lm["triangles"]) // 3),
        "hausdorff": h, "bbox_diagonal": diag, "hausdorff_rel": h / diag,
    })
    po.close()
    lo.close()
    tree.close()
    sc.close()
    return out


def _seeds(s):
    a, b = s.split(":")
    return list(range(int(a), int(b)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--train", default="0:200", help="training seeds a:b")
    ap.add_argument("--holdout", default="1000:1020", help="held-out seeds a:b")
    ap.add_argument("--res", type=float, default=1.0, help="resolution scale of the 320x240 depth maps")
    ap.add_argument("--rows", type=int, default=40000, help="candidate rows kept per training scene")
    ap.add_argument("--pos-weight", type=float, default=8.0, help="extra cost of a missed surface node")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--thresholds", default="0.02,0.05,0.1,0.2,0.3")
    ap.add_argument("--tune", type=int, default=10, help="training scenes re-run to choose the threshold")
    ap.add_argument("--max-hausdorff", type=float, default=1e-3, help="bound, fraction of the bbox diagonal")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--report", default=None, help="JSON report path")
    ap.add_argument("--eval-only", action="store_true",
                    help="no training: evaluate the weights in --out on --holdout seeds of --config")
    ap.add_argument("--config", type=int, default=SCENE_CONFIG, help="scene config of --eval-only")
    args = ap.parse_args(argv)
    if args.eval_only:
        ctx = Context(0)
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
        return 0

    ctx = Context(0)
    rng = np.random.default_rng(0)
    Xs, ys, ws, counts = [], [], [], []
    t0 = time.perf_counter()
    for k, seed in enumerate(_seeds(args.train)):
        X, y, w, c = scene_rows(ctx, seed, args.res, rng, args.rows)
        Xs.append(X)
        ys.append(y)
        ws.append(w)
        counts.append(c)
        if k % 10 == 0:
            print(f"[culling] scene {seed}: {c[2]} samples, {c[0]} surface / {c[1]} other candidate nodes "
                  f"({time.perf_counter() - t0:.0f}s)", flush=True)
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    w = np.concatenate(ws)
    w = np.where(y > 0.5, w * args.pos_weight, w)
    w = (w / w.mean()).astype(np.float32)
    mean, std = X.mean(0).astype(np.float64), X.std(0).astype(np.float64)
    std = np.where(std > 1e-6, std, 1.0)
    Xn = ((X - mean) / std).astype(np.float32)
    print(f"[culling] {len(X)} rows, {y.mean():.3f} positive; training {args.epochs} epochs", flush=True)
    w0 = init_mlp_weights(0)
    wt, losses = mlp_train(ctx, w0, Xn, y, w, epochs=args.epochs, batch=args.batch, learning_rate=args.lr, seed=1)
    print("[culling] loss per epoch " + " ".join(f"{v:.4f}" for v in losses), flush=True)
    weights = fold_standardisation(wt, mean, std)
    assert weights.size == MLP_PARAM_COUNT

    # threshold: the largest candidate whose mesh stays within the bound on the tuning scenes
    tune = _seeds(args.train)[:args.tune]
    chosen, tuning = None, []
    culler = Culler(ctx, weights, 0.5)
    for thr in sorted(float(t) for t in args.thresholds.split(",")):
        L.check(L.lib().tsr_culler_set_weights(ctx.h, culler.h, weights.ctypes.data_as(L.f32p),
                                               ctypes_float(thr)))
        rs = [evaluate(ctx, culler, s, args.res) for s in tune]
        worst = max(r["hausdorff_rel"] for r in rs)
        gain = np.mean([r["plain_ms"] / r["learned_ms"] for r in rs])
        tuning.append({"threshold": thr, "worst_hausdorff_rel": worst, "mean_speedup": gain})
        print(f"[culling] threshold {thr}: worst Hausdorff {worst:.2e} of the diagonal, speedup {gain:.2f}",
              flush=True)
        if worst <= args.max_hausdorff:
            chosen = thr
    if chosen is None:
        print(f"[culling] no threshold keeps the tuning scenes within {args.max_hausdorff:.1e} of the diagonal; "
              "no weights written", file=sys.stderr, flush=True)
        return 1
    L.check(L.lib().tsr_culler_set_weights(ctx.h, culler.h, weights.ctypes.data_as(L.f32p), ctypes_float(chosen)))
    held = [evaluate(ctx, culler, s, args.res) for s in _seeds(args.holdout)]
    summary = {
        "threshold": chosen,
        "train_seeds": args.train, "holdout_seeds": args.holdout, "resolution_scale": args.res,
        "rows": int(len(X)), "positive_fraction": float(y.mean()), "epochs": args.epochs,
        "loss_per_epoch": [float(v) for v in losses], "tuning": tuning,
        "holdout_worst_hausdorff_rel": max(r["hausdorff_rel"] for r in held),
        "holdout_mean_speedup": float(np.mean([r["plain_ms"] / r["learned_ms"] for r in held])),
        "holdout_miss_rate": float(sum(r["missed_leaves"] for r in held) /
                                   max(sum(r["positive_leaves"] for r in held), 1)),
        "holdout_culled_leaf_fraction": float(np.mean([r["culled_leaves"] / r["leaves"] for r in held])),
        "holdout": held,
    }
    np.savez(args.out, weights=weights, threshold=np.float32(chosen), feature_mean=mean, feature_std=std)
    print(json.dumps({k: v for k, v in summary.items() if k != "holdout"}), flush=True)
    if args.report:
        with open(args.report, "w") as f:
            json.dump(summary, f, indent=1)
    culler.close()
    ctx.close()


def ctypes_float(v):
    import ctypes
    return ctypes.c_float(v)


if __name__ == "__main__":
    sys.exit(main())
