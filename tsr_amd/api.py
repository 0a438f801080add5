"""Python face of the GPU occupancy + DMC library, shaped like TSR's hot path.

Scene.reconstruct() follows Scene::reconstruct (SR/Scene/Scene.cpp:169-197):
Tree() -> Occupancy(tree) -> DualMarchingCells::extractMesh.  The objects own
device memory through the C ABI (include/tsr_gpu.h); downloads return numpy arrays.
"""
import ctypes

import numpy as np

from . import _lib as L


def _fp(a):
    return a.ctypes.data_as(L.f32p) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(L.f64p) if a is not None else None


def _up(a):
    return a.ctypes.data_as(L.u32p) if a is not None else None


class Context:
    """One HIP device + one stream (tsr_context)."""

    def __init__(self, device=0):
        h = L.vp()
        L.check(L.lib().tsr_context_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            L.lib().tsr_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        L.check(L.lib().tsr_context_synchronize(self.h))

    def phase_times(self):
        """Milliseconds: tree, S pass, E pass, priors, occupancy, DMC (HIP events)."""
        ms = np.zeros(6, np.float32)
        L.check(L.lib().tsr_context_phase_times(self.h, _fp(ms), 6))
        return ms


    KERNELS = {"priors": 0, "emptiness_pass": 1, "emptiness_overflow": 2, "sampleness_pass": 3,
               "sampleness_overflow": 4, "emptiness_gather": 5}

    def kernel_stats(self, name):
        """(ms, launches, algorithmic bytes, units) accumulated for an instrumented kernel."""
        ms = ctypes.c_double(0)
        la, by, un = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        L.check(L.lib().tsr_context_kernel_stats(self.h, self.KERNELS[name], ctypes.byref(ms), ctypes.byref(la),
                                                 ctypes.byref(by), ctypes.byref(un)))
        return ms.value, la.value, by.value, un.value

    def reset_kernel_stats(self):
        L.check(L.lib().tsr_context_reset_kernel_stats(self.h))

    def calibration_copy(self, nbytes):
        L.check(L.lib().tsr_calibration_copy(self.h, int(nbytes)))


def read_samples_file(path):
    """Samples::loadFromFile: SoA arrays (dict) and cameras per sample."""
    n, mc = ctypes.c_uint32(0), ctypes.c_uint32(0)
    L.check(L.lib().tsr_samples_file_info(str(path).encode(), ctypes.byref(n), ctypes.byref(mc)))
    out = {k: np.zeros(n.value, np.float32) for k in ("px", "py", "pz", "nx", "ny", "nz", "scale", "confidence")}
    out["parent_cameras"] = np.zeros(n.value * mc.value, np.uint32)
    L.check(L.lib().tsr_samples_file_read(str(path).encode(), *[_fp(out[k]) for k in
                                                                ("px", "py", "pz", "nx", "ny", "nz", "scale",
                                                                 "confidence")], _up(out["parent_cameras"])))
    return out, mc.value


def device_memory(reset_peak=False):
    """(current, peak) device bytes held by the library in this process (tsr_device_memory)."""
    cur, peak = ctypes.c_uint64(0), ctypes.c_uint64(0)
    L.check(L.lib().tsr_device_memory(ctypes.byref(cur), ctypes.byref(peak), int(bool(reset_peak))))
    return cur.value, peak.value


class SyntheticScene:
    """Seeded analytic scene (config 0..4) generated on the host (tsr_synthetic_*)."""

    def __init__(self, config, seed=0, resolution_scale=1.0):
        h = L.vp()
        L.check(L.lib().tsr_synthetic_create(int(config), int(seed), float(resolution_scale), ctypes.byref(h)))
        self.h = h
        sv, cv = L.SamplesView(), L.CamerasView()
        L.check(L.lib().tsr_synthetic_views(h, ctypes.byref(sv), ctypes.byref(cv)))
        n, k = sv.count, cv.count

        def arr(ptr, count, dt):
            if count == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dt, copy=True)

        self.samples = {
            "px": arr(sv.px, n, np.float32), "py": arr(sv.py, n, np.float32), "pz": arr(sv.pz, n, np.float32),
            "nx": arr(sv.nx, n, np.float32), "ny": arr(sv.ny, n, np.float32), "nz": arr(sv.nz, n, np.float32),
            "scale": arr(sv.scale, n, np.float32), "confidence": arr(sv.confidence, n, np.float32),
            "parent_cameras": arr(sv.parent_cameras, n, np.uint32),
        }
        self.cameras = {"x": arr(cv.x, k, np.float32), "y": arr(cv.y, k, np.float32), "z": arr(cv.z, k, np.float32)}
        L.lib().tsr_synthetic_destroy(h)
        self.h = None

    @property
    def count(self):
        return len(self.samples["px"])


def samples_view(samples, max_cams=1):
    s = {k: np.ascontiguousarray(v) for k, v in samples.items()}
    v = L.SamplesView()
    v.count = len(s["px"])
    v.max_cameras_per_sample = max_cams
    for k in ("px", "py", "pz", "nx", "ny", "nz", "scale", "confidence"):
        setattr(v, k, _fp(s[k]))
    v.parent_cameras = _up(s["parent_cameras"])
    return v, s


def cameras_view(cameras):
    c = {k: np.ascontiguousarray(v, dtype=np.float32) for k, v in cameras.items()}
    v = L.CamerasView()
    v.count = len(c["x"])
    v.x, v.y, v.z = _fp(c["x"]), _fp(c["y"]), _fp(c["z"])
    return v, c


class Scene:
    """Device copy of samples + cameras (the data TSR reads through Scene::getSingleton)."""

    def __init__(self, ctx, samples, cameras, max_cams=1):
        self.ctx = ctx
        sv, keep1 = samples_view(samples, max_cams)
        cv, keep2 = cameras_view(cameras)
        h = L.vp()
        L.check(L.lib().tsr_scene_upload(ctx.h, ctypes.byref(sv), ctypes.byref(cv), ctypes.byref(h)))
        self.h = h
        self.count = sv.count
        self.max_cams = max_cams

    def close(self):
        if self.h:
            L.lib().tsr_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def download(self):
        n = self.count
        out = {k: np.zeros(n, np.float32) for k in ("px", "py", "pz", "nx", "ny", "nz", "scale", "confidence")}
        out["parent_cameras"] = np.zeros(n * self.max_cams, np.uint32)
        L.check(L.lib().tsr_scene_download(self.ctx.h, self.h, *[_fp(out[k]) for k in
                                                                  ("px", "py", "pz", "nx", "ny", "nz", "scale",
                                                                   "confidence")], _up(out["parent_cameras"])))
        return out


class Tree:
    """Tree::Tree() on the GPU (nodes, leaves, dual cells; permutes the scene's samples)."""

    def __init__(self, ctx, scene, params=None):
        self.ctx = ctx
        p = params or L.default_params()
        h = L.vp()
        L.check(L.lib().tsr_tree_build(ctx.h, scene.h, ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        info = L.TreeInfo()
        L.check(L.lib().tsr_tree_get_info(h, ctypes.byref(info)))
        self.node_count, self.leaf_count, self.cell_count = info.node_count, info.leaf_count, info.cell_count
        self.max_depth = info.max_depth
        self.root_min = np.array(info.root_min[:], np.float32)
        self.root_size = np.float32(info.root_size)
        self.sample_count = scene.count
        self.input_count = scene.count   # samples the tree was built from
        self.max_cams = scene.max_cams

    def close(self):
        if self.h:
            L.lib().tsr_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    TABLES = ("children", "parents", "sample_counts", "sample_starts", "node_to_leaf", "leaf_nodes",
              "leaf_scopes", "cells", "sample_target", "node_depth")

    def download(self, keys=None):
        """Host copies of the tables; `keys` limits the download to those tables."""
        N, Lc, C, S = self.node_count, self.leaf_count, self.cell_count, self.input_count
        shape = {"children": (N, np.uint32), "parents": (N, np.uint32), "sample_counts": (N, np.uint32),
                 "sample_starts": (N, np.uint32), "node_to_leaf": (N, np.uint32), "leaf_nodes": (Lc, np.uint32),
                 "leaf_scopes": (4 * Lc, np.float32), "cells": (8 * C, np.uint32), "sample_target": (S, np.uint32),
                 "node_depth": (N, np.uint32)}
        want = self.TABLES if keys is None else tuple(keys)
        o = {k: np.zeros(shape[k][0], shape[k][1]) for k in want}
        g = lambda k, f: f(o[k]) if k in o else None  # noqa: E731
        L.check(L.lib().tsr_tree_download(self.ctx.h, self.h, g("children", _up), g("parents", _up),
                                          g("sample_counts", _up), g("sample_starts", _up), g("node_to_leaf", _up),
                                          g("leaf_nodes", _up), g("leaf_scopes", _fp), g("cells", _up),
                                          g("sample_target", _up), g("node_depth", _up)))
        return o

    def erase_samples(self, inliers):
        """Tree::eraseSamples + Samples::compact: keep the samples with inliers[i] != 0 (node order)"""
        m = np.ascontiguousarray(np.asarray(inliers) != 0, np.uint8)
        L.check(L.lib().tsr_tree_erase_samples(self.ctx.h, self.h, m.ctypes.data_as(L.u8p)))
        self.sample_count = int(m.sum())

    def build_stats(self):
        """sampling refinement statistics: units, refining units, created nodes, device lookups"""
        a = np.zeros(4, np.uint64)
        L.check(L.lib().tsr_tree_build_stats(self.h, a.ctypes.data_as(L.u64p)))
        return {k: int(v) for k, v in zip(("units", "events", "created", "device_lookups"), a)}

    def neighbors(self):
        """Leaves::getNeighbors for all leaves: (offsets[leaf_count + 1], neighbours) CSR."""
        total = ctypes.c_uint64(0)
        L.check(L.lib().tsr_tree_neighbor_count(self.ctx.h, self.h, ctypes.byref(total)))
        offs = np.zeros(self.leaf_count + 1, np.uint32)
        nb = np.zeros(total.value, np.uint32)
        L.check(L.lib().tsr_tree_neighbors_download(self.ctx.h, self.h, _up(offs), _up(nb)))
        return offs, nb

    def save(self, nodes_path=None, leaves_path=None, cells_path=None, samples_path=None):
        """Tree::saveToFiles (.nodes, .leaves, .dualCells) and the samples in node order."""
        enc = lambda p: None if p is None else str(p).encode()  # noqa: E731
        L.check(L.lib().tsr_tree_save(self.ctx.h, self.h, enc(nodes_path), enc(leaves_path), enc(cells_path)))
        if samples_path is not None:
            L.check(L.lib().tsr_tree_samples_save(self.ctx.h, self.h, enc(samples_path)))

    @classmethod
    def load(cls, ctx, scene, nodes_path, leaves_path=None, cells_path=None, params=None):
        """Tree::Tree(const Path&) over a scene holding the tree's samples in node order."""
        t = cls.__new__(cls)
        t.ctx = ctx
        p = params or L.default_params()
        h = L.vp()
        enc = lambda q: None if q is None else str(q).encode()  # noqa: E731
        L.check(L.lib().tsr_tree_load(ctx.h, scene.h, ctypes.byref(p), enc(nodes_path), enc(leaves_path),
                                      enc(cells_path), ctypes.byref(h)))
        t.h = h
        info = L.TreeInfo()
        L.check(L.lib().tsr_tree_get_info(h, ctypes.byref(info)))
        t.node_count, t.leaf_count, t.cell_count = info.node_count, info.leaf_count, info.cell_count
        t.max_depth = info.max_depth
        t.root_min = np.array(info.root_min[:], np.float32)
        t.root_size = np.float32(info.root_size)
        t.sample_count = scene.count
        t.input_count = scene.count
        t.max_cams = scene.max_cams
        return t

    def download_samples(self):
        """The samples in node order (Samples::reorder)."""
        n = self.sample_count
        out = {k: np.zeros(n, np.float32) for k in ("px", "py", "pz", "nx", "ny", "nz", "scale", "confidence")}
        out["parent_cameras"] = np.zeros(n * self.max_cams, np.uint32)
        L.check(L.lib().tsr_tree_samples_download(self.ctx.h, self.h, *[_fp(out[k]) for k in
                                                                        ("px", "py", "pz", "nx", "ny", "nz", "scale",
                                                                         "confidence")], _up(out["parent_cameras"])))
        return out


class Occupancy:
    """Occupancy::Occupancy(tree) on the GPU."""

    def __init__(self, ctx, tree, scene, params=None, flags=L.TSR_ACCUM_KAHAN_SEGMENTED, culler=None):
        self.ctx = ctx
        p = params or L.default_params()
        h = L.vp()
        L.check(L.lib().tsr_occupancy_compute(ctx.h, tree.h, scene.h, ctypes.byref(p), int(flags),
                                              culler.h if culler is not None else None, ctypes.byref(h)))
        self.h = h
        self.leaf_count = tree.leaf_count
        self.node_count = tree.node_count

    def close(self):
        if self.h:
            L.lib().tsr_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def save(self, path, real_bytes=8):
        """Occupancy::saveToFile; real_bytes 4 is the reference's Real = float layout."""
        L.check(L.lib().tsr_field_save(self.ctx.h, self.h, str(path).encode(), int(real_bytes)))

    @classmethod
    def load(cls, ctx, tree, path, params=None):
        """Occupancy(const Path&) for a tree; the occupancy ratio is recomputed."""
        o = cls.__new__(cls)
        o.ctx = ctx
        p = params or L.default_params()
        h = L.vp()
        L.check(L.lib().tsr_field_load(ctx.h, tree.h, ctypes.byref(p), str(path).encode(), ctypes.byref(h)))
        o.h = h
        o.leaf_count = tree.leaf_count
        o.node_count = tree.node_count
        return o

    def add_samples(self, tree, scene, samples):
        """Occupancy::addSamples: the field follows added samples (node-order indices)"""
        idx = np.ascontiguousarray(samples, np.uint32)
        L.check(L.lib().tsr_occupancy_add_samples(self.ctx.h, self.h, tree.h, scene.h, idx.ctypes.data_as(L.u32p),
                                                  len(idx)))

    def erase_samples(self, tree, scene, samples):
        """Occupancy::eraseSamples: the field follows erased samples (node-order indices)"""
        idx = np.ascontiguousarray(samples, np.uint32)
        L.check(L.lib().tsr_occupancy_erase_samples(self.ctx.h, self.h, tree.h, scene.h, idx.ctypes.data_as(L.u32p),
                                                    len(idx)))

    def prior_stages(self):
        """dict of the priors stages: queries, entering stage 0 / 1 / 2 (2 = exact), and
        the queries the routing network sent straight to stage 1 / 2"""
        a = np.zeros(6, np.uint32)
        L.check(L.lib().tsr_field_prior_stages(self.h, a.ctypes.data_as(L.u32p)))
        keys = ("queries", "stage0", "stage1", "exact", "routed_stage1", "routed_exact")
        return {k: int(v) for k, v in zip(keys, a)}

    def prior_levels(self):
        """per leaf: the stage that decided it (0, 1 certified bounds, 2 exact), 255 if not queried"""
        c = np.zeros(self.leaf_count, np.uint8)
        L.check(L.lib().tsr_field_prior_levels(self.ctx.h, self.h, c.ctypes.data_as(L.u8p)))
        return c

    def hit_counts(self):
        h = np.zeros(2, np.uint64)
        L.check(L.lib().tsr_field_hit_counts(self.h, h.ctypes.data_as(L.u64p)))
        return h

    def download(self):
        Lc, N = self.leaf_count, self.node_count
        o = {"kernel_sums": np.zeros(2 * Lc, np.float32), "cone_lengths": np.zeros(2 * Lc, np.float32),
             "priors": np.zeros(Lc, np.float64), "node_states": np.zeros(N, np.uint32),
             "occupancy": np.zeros(Lc, np.float64)}
        m = ctypes.c_double(0.0)
        L.check(L.lib().tsr_field_download(self.ctx.h, self.h, _fp(o["kernel_sums"]), _fp(o["cone_lengths"]),
                                           _dp(o["priors"]), _up(o["node_states"]), _dp(o["occupancy"]),
                                           ctypes.byref(m)))
        o["cone_count"] = m.value
        return o


class DualMarchingCells:
    """DualMarchingCells::extractMesh on the GPU (raw crust mesh)."""

    def __init__(self, ctx, tree, occupancy):
        self.ctx = ctx
        h = L.vp()
        L.check(L.lib().tsr_dmc_extract(ctx.h, tree.h, occupancy.h, ctypes.byref(h)))
        self.h = h
        info = L.MeshInfo()
        L.check(L.lib().tsr_mesh_get_info(h, ctypes.byref(info)))
        self.vertex_count, self.triangle_count = info.vertex_count, info.triangle_count

    def close(self):
        if self.h:
            L.lib().tsr_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def postprocess(self, min_triangle_isle_size=1000):
        """Isle removal, hole filling and vertex normals (the rest of DualMarchingCells' constructor)."""
        L.check(L.lib().tsr_crust_postprocess(self.ctx.h, self.h, int(min_triangle_isle_size)))
        info = L.MeshInfo()
        L.check(L.lib().tsr_mesh_get_info(self.h, ctypes.byref(info)))
        self.vertex_count, self.triangle_count = info.vertex_count, info.triangle_count
        return self

    def normals(self):
        n = np.zeros(3 * self.vertex_count, np.float64)
        L.check(L.lib().tsr_mesh_download_normals(self.ctx.h, self.h, _dp(n)))
        return n

    def download(self):
        V, T = self.vertex_count, self.triangle_count
        o = {"positions": np.zeros(3 * V, np.float64), "scales": np.zeros(V, np.float64),
             "edge_keys": np.zeros(2 * V, np.uint32), "triangles": np.zeros(3 * T, np.uint32)}
        L.check(L.lib().tsr_mesh_download(self.ctx.h, self.h, _dp(o["positions"]), _dp(o["scales"]),
                                          _up(o["edge_keys"]), _up(o["triangles"])))
        return o


class _MeshHandle(DualMarchingCells):
    def __init__(self, ctx, h):   # noqa: super().__init__ is not called: the handle exists
        self.ctx = ctx
        self.h = h
        info = L.MeshInfo()
        L.check(L.lib().tsr_mesh_get_info(h, ctypes.byref(info)))
        self.vertex_count, self.triangle_count = info.vertex_count, info.triangle_count


def dmc_from_values(ctx, tree, occupancy):
    """DualMarchingCells over an explicit per-leaf occupancy (tsr_dmc_extract_values)."""
    o = np.ascontiguousarray(occupancy, np.float64)
    if len(o) != tree.leaf_count:
        raise ValueError(f"expected {tree.leaf_count} values, got {len(o)}")
    h = L.vp()
    L.check(L.lib().tsr_dmc_extract_values(ctx.h, tree.h, _dp(o), ctypes.byref(h)))
    return _MeshHandle(ctx, h)


MLP_PARAM_COUNT = 64 * 16 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1
LEAF_FEATURES = 16


class Culler:
    """Routing network of the certified priors on the device (tsr_culler): weights and the
    threshold t of the first stage floor(2 p + 1 - t)."""

    def __init__(self, ctx, weights, threshold=0.75):
        w = np.ascontiguousarray(weights, np.float32)
        if w.size != MLP_PARAM_COUNT:
            raise ValueError(f"expected {MLP_PARAM_COUNT} weights, got {w.size}")
        h = L.vp()
        L.check(L.lib().tsr_culler_create(ctx.h, _fp(w), ctypes.c_float(threshold), ctypes.byref(h)))
        self.h = h
        self.threshold = threshold

    def close(self):
        if self.h:
            L.lib().tsr_culler_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def leaf_features(ctx, tree, occupancy):
    """LEAF_FEATURES floats per leaf of a tree (from a computed field)"""
    f = np.zeros((tree.leaf_count, LEAF_FEATURES), np.float32)
    L.check(L.lib().tsr_leaf_features(ctx.h, tree.h, occupancy.h, _fp(f)))
    return f


def mlp_forward(ctx, weights, features):
    w = np.ascontiguousarray(weights, np.float32)
    x = np.ascontiguousarray(features, np.float32)
    out = np.zeros(len(x), np.float32)
    L.check(L.lib().tsr_mlp_forward(ctx.h, _fp(w), _fp(x), len(x), _fp(out)))
    return out


def mlp_gradient(ctx, weights, features, labels, row_weights=None):
    """(summed weighted BCE loss, gradient) of one batch"""
    w = np.ascontiguousarray(weights, np.float32)
    x = np.ascontiguousarray(features, np.float32)
    y = np.ascontiguousarray(labels, np.float32)
    rw = np.ascontiguousarray(np.ones(len(x)) if row_weights is None else row_weights, np.float32)
    g = np.zeros(MLP_PARAM_COUNT, np.float32)
    loss = ctypes.c_double(0.0)
    L.check(L.lib().tsr_mlp_gradient(ctx.h, _fp(w), _fp(x), _fp(y), _fp(rw), len(x), _fp(g), ctypes.byref(loss)))
    return loss.value, g


def mlp_train(ctx, weights, features, labels, row_weights=None, epochs=10, batch=4096, learning_rate=1e-3, seed=0):
    """Adam on the device; returns (trained weights, mean loss per epoch)"""
    w = np.array(weights, np.float32).reshape(-1)
    x = np.ascontiguousarray(features, np.float32)
    y = np.ascontiguousarray(labels, np.float32)
    rw = np.ascontiguousarray(np.ones(len(x)) if row_weights is None else row_weights, np.float32)
    losses = np.zeros(epochs, np.float32)
    L.check(L.lib().tsr_mlp_train(ctx.h, _fp(w), _fp(x), _fp(y), _fp(rw), len(x), epochs, batch,
                                  ctypes.c_float(learning_rate), seed, _fp(losses)))
    return w, losses


def init_mlp_weights(seed=0):
    """He-initialised weights in the flat layout W1,b1,W2,b2,W3,b3,W4,b4 ([out][in])."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((16, 64), (64, 64), (64, 64), (64, 1)):
        parts.append(rng.normal(0.0, np.sqrt(2.0 / fan_in), fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def reconstruct(ctx, samples, cameras, params=None, flags=L.TSR_ACCUM_KAHAN_SEGMENTED, culler=None):
    """Scene::reconstruct (Scene.cpp:169-197) up to the raw crust mesh."""
    scene = Scene(ctx, samples, cameras)
    tree = Tree(ctx, scene, params)
    occ = Occupancy(ctx, tree, scene, params, flags, culler)
    mesh = DualMarchingCells(ctx, tree, occ)
    return scene, tree, occ, mesh


class DepthMesh:
    """DepthImage::triangulate + createFlexibleMesh with the view's parent cameras (tsr_depth_mesh).

    view: dict with depths (H x W float32, depth along the pixel ray, <= 0: none), position (3),
    rotation (3 x 3 world-to-view), focal_length (normalised by max(W, H)), principal_point (2,
    normalised), aspect_ratio, camera_index, and optionally views (H x W x C camera indices,
    0xFFFFFFFF: none)."""

    def __init__(self, ctx, view):
        self.ctx = ctx
        d = np.ascontiguousarray(view["depths"], np.float32)
        H, W = d.shape
        v = L.DepthView()
        v.width, v.height = W, H
        v.depths = d.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        v.position[:] = [float(x) for x in np.asarray(view["position"]).reshape(3)]
        v.rotation[:] = [float(x) for x in np.asarray(view["rotation"]).reshape(9)]
        v.focal_length = float(view["focal_length"])
        v.principal_point[:] = [float(x) for x in view.get("principal_point", (0.5, 0.5))]
        v.aspect_ratio = float(view.get("aspect_ratio", 1.0))
        v.camera_index = int(view["camera_index"])
        views = view.get("views")
        if views is not None:
            vw = np.ascontiguousarray(views, np.uint32).reshape(H, W, -1)
            v.views_channels = vw.shape[2]
            v.views = vw.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        else:
            vw = None
            v.views_channels = 0
            v.views = None
        h = L.vp()
        L.check(L.lib().tsr_depth_mesh_create(ctx.h, ctypes.byref(v), ctypes.byref(h)))
        self.h = h
        self.width, self.height = W, H
        a, b, c = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint32(0)
        L.check(L.lib().tsr_depth_mesh_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        self.vertex_count, self.triangle_count, self.cameras_per_sample = a.value, b.value, c.value

    def download(self):
        V, T, C = self.vertex_count, self.triangle_count, self.cameras_per_sample
        o = {"positions": np.zeros(3 * V), "normals": np.zeros(3 * V), "scales": np.zeros(V),
             "triangles": np.zeros(3 * T, np.uint32), "parent_cameras": np.zeros(V * C, np.uint32),
             "pixel_to_vertex": np.zeros(self.width * self.height, np.uint32)}
        L.check(L.lib().tsr_depth_mesh_download(self.ctx.h, self.h, _dp(o["positions"]), _dp(o["normals"]),
                                                _dp(o["scales"]), _up(o["triangles"]), _up(o["parent_cameras"]),
                                                _up(o["pixel_to_vertex"])))
        o["cameras_per_sample"] = C
        return o

    def close(self):
        if self.h:
            L.lib().tsr_depth_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def samples_from_depth_maps(ctx, views):
    """Scene::loadDepthMeshes + Samples::addSamplesViaMeshes (Scene.cpp:484-524, Samples.cpp:224-273):
    one sample per depth-mesh vertex (confidence 1), parent cameras padded with 0xFFFFFFFF to the
    largest cameras-per-sample; returns (samples dict, max cameras per sample)"""
    meshes = []
    for v in views:
        m = DepthMesh(ctx, v)
        meshes.append(m.download())
        m.close()
    mc = max([m["cameras_per_sample"] for m in meshes] + [1])
    keys = ("px", "py", "pz", "nx", "ny", "nz", "scale", "confidence")
    out = {k: [] for k in keys + ("parent_cameras",)}
    for m in meshes:
        P, N = m["positions"].reshape(-1, 3), m["normals"].reshape(-1, 3)
        n = len(P)
        for a, k in enumerate(("px", "py", "pz")):
            out[k].append(P[:, a])
        for a, k in enumerate(("nx", "ny", "nz")):
            out[k].append(N[:, a])
        out["scale"].append(m["scales"])
        out["confidence"].append(np.ones(n))
        cams = np.full((n, mc), 0xFFFFFFFF, np.uint32)
        cams[:, :m["cameras_per_sample"]] = m["parent_cameras"].reshape(n, -1)
        out["parent_cameras"].append(cams.reshape(-1))
    res = {k: np.concatenate(out[k]).astype(np.float32) for k in keys}
    res["parent_cameras"] = np.concatenate(out["parent_cameras"]).astype(np.uint32)
    return res, mc


def write_cameras_file(path, data13, view_ids, real_bytes=4):
    """Cameras::saveToFile: data13 (n x 13: quaternion, position, principal point, distortion 2,
    focal length, aspect ratio) and view IDs"""
    d = np.ascontiguousarray(data13, np.float64).reshape(-1, 13)
    v = np.ascontiguousarray(view_ids, np.uint32)
    L.check(L.lib().tsr_cameras_file_write(str(path).encode(), len(v), _dp(d), _up(v), int(real_bytes)))


def read_cameras_file(path):
    """Cameras::loadFromFile: (n x 13 data, view IDs)"""
    n = ctypes.c_uint32(0)
    L.check(L.lib().tsr_cameras_file_info(str(path).encode(), ctypes.byref(n)))
    d = np.zeros((n.value, 13))
    v = np.zeros(n.value, np.uint32)
    L.check(L.lib().tsr_cameras_file_read(str(path).encode(), _dp(d), _up(v)))
    return d, v
