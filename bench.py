"""Benchmark: input samples/s from resident SoA samples to occupancy field + DMC mesh.

One step = Tree() + Occupancy(tree) + DualMarchingCells::extractMesh on one MI355X
(HIP), with the samples resident in HBM before timing starts.  The default workload
is the largest single-GPU scene: 30 random primitives including thin sheets in
[-4,4]^3, 120 cameras at 3..20, 1024x768 depth maps, 2 % outliers (36.5M samples
after miss filtering).  `--config 1` runs the thin wedge + sphere scene.
Prints one JSON line.  `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os

import numpy as np
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


WORKLOADS = {
    0: "unit sphere r=1; 12 cameras on a ring of radius 3, 128x96 depth maps, 0.5% depth noise",
    1: "thin wedge (two 1x1 sheets, 0.01 thick, 10 deg) + sphere r=0.4; 64 cameras at 2..8, "
       "640x480 depth maps, 0.3% depth noise, 1% outliers",
    3: "union of 3-8 random spheres/boxes/thin sheets; 32 cameras at 2..8, 320x240 depth maps, "
       "0.3% depth noise, 1% outliers",
    4: "30 random primitives incl. thin sheets in [-4,4]^3; 120 cameras at 3..20, 1024x768 depth maps, "
       "0.3% depth noise, 2% outliers",
}


# workload names of the committed counter profiles (profiles/pmc_traffic_<kernel>_<name>.txt)
WORKLOAD_TAGS = {1: "wedge_sphere_9.2M", 4: "primitives30_36.5M"}


def cpu_info():
    model, cores = "unknown", os.cpu_count() or 1
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return model, cores


# samples of the CPU leg: about 0.57M samples of the same scene (10-30 s of host work)
CPU_RES = {0: 1.0, 1: 0.25, 3: 0.5, 4: 0.125}


def cpu_baseline_and_precision(ctx, config, res, threads):
    """CPU leg.  (1) The oracle in its OpenMP float mode (the reference's Real = float and
    its parallel structure) timed on a bounded sample of the workload: the same scene at a
    reduced resolution.  (2) Precision of the GPU path on that same sample against the
    oracle in FP64: per-leaf field error, sign flips and mesh Hausdorff distance, for the
    default accumulation and for plain FP32 atomics."""
    from oracle.oracle import Oracle
    from tsr_amd import DualMarchingCells, Occupancy, Scene, SyntheticScene, Tree, _lib
    from tsr_amd.metrics import bbox_diagonal, field_error, hausdorff

    syn = SyntheticScene(config, seed=0, resolution_scale=res)
    o = Oracle(syn.samples, syn.cameras)
    t0 = time.perf_counter()
    o.build_tree()
    o.occupancy(threads=threads)
    o.dmc()
    dt = time.perf_counter() - t0
    ph = o.timings()
    # FP64 reference field and mesh on the same tree (OpenMP FP64)
    o.occupancy(threads=threads, fp64=True)
    o.dmc()
    ref_tree, ref_field, ref_mesh = o.tree(), o.field(), o.mesh()
    diag = bbox_diagonal(syn.samples)
    prec = {"samples": syn.count, "resolution_scale": res, "reference": "oracle FP64 (OpenMP)",
            "bbox_diagonal": diag}
    scene = Scene(ctx, syn.samples, syn.cameras)
    tree = Tree(ctx, scene)
    t = tree.download()
    prec["tree_tables_equal"] = bool(all(np.array_equal(t[k], ref_tree[k]) for k in
                                         ("children", "sample_target", "cells")))
    for name, flags in (("fp32_kahan_segmented", _lib.TSR_ACCUM_KAHAN_SEGMENTED),
                        ("fp32_atomic", _lib.TSR_ACCUM_FP32_ATOMIC)):
        occ = Occupancy(ctx, tree, scene, flags=flags | _lib.TSR_PRIORS_ALL_LEAVES)
        mesh = DualMarchingCells(ctx, tree, occ)
        f, m = occ.download(), mesh.download()
        rec = field_error(f["occupancy"], f["kernel_sums"], ref_field) if prec["tree_tables_equal"] else {}
        h, how = hausdorff(m, ref_mesh, diag)
        rec["hausdorff_over_bbox_diagonal"] = h
        rec["hausdorff_method"] = how
        rec["mesh_identical"] = bool(how.startswith("same topology") and h == 0.0)
        prec[name] = rec
        mesh.close()
        occ.close()
    tree.close()
    scene.close()
    return syn.count, dt, ph, prec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--res", type=float, default=1.0)
    ap.add_argument("--cpu-res", type=float, default=None,
                    help="resolution scale of the CPU baseline / precision sample (default: ~0.57M samples)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned-compare", action="store_true",
                    help="skip the learned-cull vs plain comparison after the timed steps")
    ap.add_argument("--learned", action="store_true",
                    help="time the learned form (certified priors routed by the committed network, "
                         "tsr_amd/culler_weights.npz) instead of the plain one")
    ap.add_argument("--calibrate-copy", type=float, default=0,
                    help="also run a plain copy kernel of this many bytes (reference for PMC traffic scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding plan is design-only)", file=sys.stderr)
        sys.exit(2)

    from tsr_amd import Context, DualMarchingCells, Occupancy, Scene, SyntheticScene, Tree, device_memory

    t0 = time.perf_counter()
    syn = SyntheticScene(args.config, seed=0, resolution_scale=args.res)
    gen_s = time.perf_counter() - t0
    n = syn.count
    print(f"[bench] config {args.config} res {args.res}: {n} samples, {len(syn.cameras['x'])} cameras "
          f"(generated in {gen_s:.1f}s)", file=sys.stderr, flush=True)
    ctx = Context(0)
    if args.calibrate_copy > 0:
        ctx.calibration_copy(int(args.calibrate_copy))
    scene = Scene(ctx, syn.samples, syn.cameras)
    culler = None
    if args.learned:
        from tsr_amd import _lib
        from tsr_amd.culling import load_culler
        culler = load_culler(ctx)

    def step():
        tree = Tree(ctx, scene)
        if culler is not None:
            occ = Occupancy(ctx, tree, scene, flags=_lib.TSR_CULL_LEARNED, culler=culler)
        else:
            occ = Occupancy(ctx, tree, scene)
        mesh = DualMarchingCells(ctx, tree, occ)
        info = (tree.node_count, tree.leaf_count, tree.cell_count, mesh.vertex_count, mesh.triangle_count,
                occ.hit_counts())
        mesh.close()
        occ.close()
        tree.close()
        return info

    for w in range(args.warmup):
        info = step()
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    ctx.reset_kernel_stats()
    mem_before, _ = device_memory(reset_peak=True)
    phases = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        info = step()
        phases.append(ctx.phase_times())
        print(f"[bench] step {k} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1000.0 * dt / args.steps
    value = n * args.steps / dt
    nodes, leaves, cells, verts, tris, hits = info
    ph = [float(sum(p[i] for p in phases) / len(phases)) for i in range(6)]

    out = {
        "metric": "input samples/sec to occupancy field + DMC mesh; learned-cull vs plain speedup",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 predicates/ratio)",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS.get(args.config, f"synthetic config {args.config}")
                        + (f" at resolution scale {args.res}" if args.res != 1.0 else "")
                        + ("; learned form: certified priors with network routing" if args.learned else ""),
            "config_index": args.config,
            "resolution_scale": args.res,
            "samples": n,
            "nodes": nodes, "leaves": leaves, "dual_cells": cells,
            "mesh_vertices": verts, "mesh_triangles": tris,
            "hits_emptiness": int(hits[0]), "hits_sampleness": int(hits[1]),
            "phase_ms": {"tree": ph[0], "sampleness": ph[1], "emptiness": ph[2], "priors": ph[3],
                         "occupancy": ph[4], "dmc": ph[5]},
            "device_memory_peak_gib": device_memory()[1] / 2**30,
            "device_memory_before_steps_gib": mem_before / 2**30,
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    # roofline of the dominant instrumented kernel: algorithmic bytes per launch / average
    # launch time (HIP events on the library's stream)
    stats = {k: ctx.kernel_stats(k) for k in Context.KERNELS}
    out["kernels"] = {k: {"ms_per_launch": v[0] / v[1], "algorithmic_bytes_per_launch": v[2] / v[1],
                          "GBps": (v[2] / v[1]) / (v[0] / v[1] * 1e-3) / 1e9}
                      for k, v in stats.items() if v[1] > 0}
    if out["kernels"]:
        dom = max(out["kernels"], key=lambda k: out["kernels"][k]["ms_per_launch"])
        kd = out["kernels"][dom]
        traffic = None
        tag = WORKLOAD_TAGS.get(args.config)
        tfile = os.path.join(ROOT, "profiles", f"pmc_traffic_{dom}_{tag}.txt")
        # committed counter profiles cover the full-size plain workloads
        if tag and os.path.exists(tfile) and args.res == 1.0 and not args.learned:
            try:
                with open(tfile) as f:
                    traffic = json.loads(f.readline())["hbm_bytes_per_launch"]
            except (ValueError, KeyError):
                traffic = None
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": kd["GBps"], "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": kd["GBps"] / HBM_PEAK_GBS, "traffic": traffic,
                           "algorithmic_bytes_per_launch": kd["algorithmic_bytes_per_launch"],
                           "ms_per_launch": kd["ms_per_launch"]}
    # learned vs plain: the occupancy + DMC part of one step per mode on this scene's tree
    # (the tree build is the same for all), after the timed region: plain, the certified
    # priors without the routing network, and with it (TSR_CULL_LEARNED)
    if not args.no_learned_compare:
        from tsr_amd import _lib
        from tsr_amd.culling import load_culler
        from tsr_amd.metrics import bbox_diagonal, hausdorff
        cul = culler if culler is not None else load_culler(ctx)
        tree = Tree(ctx, scene)
        flags = {"plain": _lib.TSR_ACCUM_KAHAN_SEGMENTED, "certified": _lib.TSR_PRIORS_CERTIFIED,
                 "learned": _lib.TSR_CULL_LEARNED}
        times, pri, meshes, stages = {}, {}, {}, {}
        for mode in ("plain", "certified", "learned") * 2:   # second round counts
            ctx.synchronize()
            t1 = time.perf_counter()
            o = Occupancy(ctx, tree, scene, flags=flags[mode], culler=cul if mode == "learned" else None)
            m = DualMarchingCells(ctx, tree, o)
            ctx.synchronize()
            times[mode] = 1000.0 * (time.perf_counter() - t1)
            pri[mode] = float(ctx.phase_times()[3])
            stages[mode] = o.prior_stages()
            meshes[mode] = m.download()
            m.close()
            o.close()
        tree.close()
        diag = bbox_diagonal(syn.samples)
        same = {k: all(np.array_equal(meshes[k][a], meshes["plain"][a]) for a in ("positions", "triangles", "edge_keys"))
                for k in ("certified", "learned")}
        out["learned_vs_plain"] = {
            "plain_ms": times["plain"], "certified_ms": times["certified"], "learned_ms": times["learned"],
            "speedup": times["plain"] / times["learned"],
            "speedup_without_network": times["plain"] / times["certified"],
            "priors_ms": pri, "learned_stages": stages["learned"], "certified_stages": stages["certified"],
            "meshes_identical": same,
            "hausdorff_over_bbox_diagonal": max(hausdorff(meshes[k], meshes["plain"], diag)[0]
                                                for k in ("certified", "learned")),
            "scope": "occupancy field + DMC on the same tree, one run each after a warm-up run; "
                     "routing network trained on config-3 scenes (tsr_amd/culler_weights.npz)",
        }
    if not args.no_cpu_baseline:
        model, cores = cpu_info()
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(cores, 64)
        cres = args.cpu_res if args.cpu_res is not None else CPU_RES.get(args.config, 0.25) * args.res
        cn, cdt, cph, prec = cpu_baseline_and_precision(ctx, args.config, cres, threads)
        print(f"[bench] cpu leg done ({cn} samples, {cdt:.1f}s)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = {
            "value": cn / cdt, "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"same scene at resolution scale {cres:g} ({cn} samples, {cn / n:.4f} of the workload), "
                      f"full path (tree serial as in the reference, OpenMP float occupancy/priors, serial DMC); "
                      f"{model}; value extrapolated to the full workload as the sample's samples/s "
                      f"(linear in samples)",
            "extrapolated": True,
            "phase_s": {"tree": cph[0], "sampleness": cph[1], "emptiness": cph[2], "priors": cph[3],
                        "occupancy": cph[4], "dmc": cph[5]},
        }
        out["precision"] = prec
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
