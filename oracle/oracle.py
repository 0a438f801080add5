"""ctypes wrapper of the CPU oracle (oracle/tsr_oracle.cpp) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker / CPU baseline; the product (tsr_amd) never uses it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        f32p = ctypes.POINTER(ctypes.c_float)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.c_uint32] + [f32p] * 8 + [u32p, ctypes.c_uint32, ctypes.c_uint32] + \
            [f32p] * 3 + [ctypes.POINTER(ctypes.c_double)]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_error.restype = ctypes.c_char_p
        L.oracle_error.argtypes = [ctypes.c_void_p]
        L.oracle_build_tree.argtypes = [ctypes.c_void_p]
        L.oracle_occupancy.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_dmc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_count.restype = ctypes.c_uint64
        L.oracle_count.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.oracle_data.restype = ctypes.c_void_p
        L.oracle_data.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.oracle_mc_case.restype = ctypes.c_char_p
        L.oracle_mc_case.argtypes = [ctypes.c_uint32]
        L.oracle_mc_edges.restype = ctypes.c_uint32
        L.oracle_mc_edges.argtypes = [ctypes.c_uint32]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


class Oracle:
    """Serial FP64 restatement of Tree / Occupancy / DualMarchingCells for one scene."""

    def __init__(self, samples, cameras, params=(2.0, 100.0, 25.0, 1.0, 2.0), max_cams=1):
        L = lib()
        self._keep = {k: np.ascontiguousarray(v) for k, v in samples.items()}
        self._keep.update({"c" + k: np.ascontiguousarray(v, dtype=np.float32) for k, v in cameras.items()})
        s = self._keep
        f = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
        prm = np.array(params, np.float64)
        self._prm = prm
        self.h = L.oracle_create(len(s["px"]), f(s["px"]), f(s["py"]), f(s["pz"]), f(s["nx"]), f(s["ny"]),
                                 f(s["nz"]), f(s["scale"]), f(s["confidence"]),
                                 s["parent_cameras"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), max_cams,
                                 len(s["cx"]), f(s["cx"]), f(s["cy"]), f(s["cz"]),
                                 prm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise OracleError(lib().oracle_error(self.h).decode())

    def build_tree(self):
        self._check(lib().oracle_build_tree(self.h))
        return self

    def occupancy(self, threads=1, brute_force=False, fp64=False):
        """threads > 1: OpenMP float mode (the reference's Real = float), or OpenMP FP64 with fp64=True."""
        mode = (int(threads) << 8) | int(bool(brute_force)) | (2 if fp64 else 0)
        self._check(lib().oracle_occupancy(self.h, mode))
        return self

    def dmc(self, occupancy=None):
        p = None
        if occupancy is not None:
            self._occ_in = np.ascontiguousarray(occupancy, np.float64)
            p = self._occ_in.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self._check(lib().oracle_dmc(self.h, p))
        return self

    def count(self, what):
        return int(lib().oracle_count(self.h, what.encode()))

    def array(self, what, dtype, n):
        ptr = lib().oracle_data(self.h, what.encode())
        if not ptr or n == 0:
            return np.zeros(n, dtype)
        buf = (ctypes.c_char * (np.dtype(dtype).itemsize * n)).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype, count=n).copy()

    def tree(self):
        N, Lc, C, S = self.count("nodes"), self.count("leaves"), self.count("cells"), self.count("samples")
        return {
            "children": self.array("children", np.uint32, N),
            "parents": self.array("parents", np.uint32, N),
            "sample_counts": self.array("sample_counts", np.uint32, N),
            "sample_starts": self.array("sample_starts", np.uint32, N),
            "node_to_leaf": self.array("node_to_leaf", np.uint32, N),
            "leaf_nodes": self.array("leaf_node", np.uint32, Lc),
            "leaf_scopes": self.array("leaf_scope", np.float32, 4 * Lc),
            "leaf_depth": self.array("leaf_depth", np.uint32, Lc),
            "cells": self.array("cells", np.uint32, 8 * C),
            "sample_target": self.array("sample_target", np.uint32, S),
            "sample_depth": self.array("sample_depth", np.uint32, S),
            "root_min": self.array("root_min", np.float32, 3),
            "root_size": self.array("root_size", np.float32, 1)[0],
        }

    def field(self):
        Lc, N = self.count("leaves"), self.count("nodes")
        return {
            "kernel_sums": self.array("kernel_sums", np.float64, 2 * Lc),
            "cone_lengths": self.array("cone_lengths", np.float64, 2 * Lc),
            "priors": self.array("priors", np.float64, Lc),
            "node_states": self.array("node_states", np.uint32, N),
            "occupancy": self.array("occupancy", np.float64, Lc),
            "cone_count": self.array("cone_count", np.float64, 1)[0],
        }

    def mesh(self):
        V, T = self.count("vertices"), self.count("triangles")
        pos = np.stack([self.array("vpos_x", np.float64, V), self.array("vpos_y", np.float64, V),
                        self.array("vpos_z", np.float64, V)], axis=1).reshape(-1)
        keys = np.stack([self.array("vkey0", np.uint32, V), self.array("vkey1", np.uint32, V)], axis=1).reshape(-1)
        return {"positions": pos, "scales": self.array("vscale", np.float64, V), "edge_keys": keys,
                "triangles": self.array("triangles", np.uint32, 3 * T), "conflicts": self.count("conflicts")}

    def leaf_neighbours(self):
        """Leaves::createNeighborsEdges: (offsets[L + 1], neighbours) CSR of the current tree"""
        L = lib()
        L.oracle_leaf_neighbours.restype = ctypes.c_uint64
        L.oracle_leaf_neighbours.argtypes = [ctypes.c_void_p]
        m = int(L.oracle_leaf_neighbours(self.h))
        n = self.count("leaves")
        return self.array("neighbour_offsets", np.uint32, n + 1), self.array("neighbours", np.uint32, m)

    def update(self, removal, samples):
        """Occupancy::addSamples / eraseSamples after an FP64 occupancy(): samples are
        ORIGINAL (input order) indices"""
        L = lib()
        L.oracle_update.restype = ctypes.c_int
        L.oracle_update.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32]
        q = np.ascontiguousarray(samples, np.uint32)
        if L.oracle_update(self.h, int(bool(removal)), q.ctypes.data, len(q)) != 0:
            raise RuntimeError(L.oracle_error(self.h).decode())
        return self

    def prior_bounds(self, leaves, cap_levels, cone_lengths=None):
        """Depth-capped prior intervals (lo, hi) of the given leaves after an FP64
        occupancy(); cone_lengths (2 * leaf count) replaces the field's own"""
        L = lib()
        L.oracle_prior_bounds.restype = ctypes.c_int
        L.oracle_prior_bounds.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        q = np.ascontiguousarray(leaves, np.uint32)
        lo, hi = np.zeros(len(q)), np.zeros(len(q))
        cl = None if cone_lengths is None else np.ascontiguousarray(cone_lengths, np.float32)
        if L.oracle_prior_bounds(self.h, q.ctypes.data, len(q), int(cap_levels),
                                 None if cl is None else cl.ctypes.data, lo.ctypes.data, hi.ctypes.data) != 0:
            raise RuntimeError(L.oracle_error(self.h).decode())
        return lo, hi

    def timings(self):
        return self.array("timings", np.float64, 6)


def mc_case(cfg):
    return lib().oracle_mc_case(cfg).decode()


def mc_edges(cfg):
    return int(lib().oracle_mc_edges(cfg))


def _d3(v):
    a = np.ascontiguousarray(v, np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ray_kernel(z, Z):
    L = lib()
    L.oracle_ray_kernel.restype = ctypes.c_double
    L.oracle_ray_kernel.argtypes = [ctypes.c_double, ctypes.c_double]
    return L.oracle_ray_kernel(z, Z)


def disc_kernel(r2, R):
    L = lib()
    L.oracle_disc_kernel.restype = ctypes.c_double
    L.oracle_disc_kernel.argtypes = [ctypes.c_double, ctypes.c_double]
    return L.oracle_disc_kernel(r2, R)


def cone_kernel(apex, end_normal, end, R, length, x, sampleness, conf=1.0):
    """Returns (scaled kernel, tComplete, tCut, getLength(), degenerate)."""
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.oracle_cone_kernel.restype = ctypes.c_double
    L.oracle_cone_kernel.argtypes = [dp, dp, dp, ctypes.c_double, ctypes.c_double, dp, ctypes.c_int,
                                     ctypes.c_double, dp]
    a, ap = _d3(apex)
    n, np_ = _d3(end_normal)
    e, ep = _d3(end)
    xx, xp = _d3(x)
    out, op = _d3(np.zeros(4))
    k = L.oracle_cone_kernel(ap, np_, ep, R, length, xp, int(sampleness), conf, op)
    return k, out[0], out[1], out[2], bool(out[3])


def cone_separated(apex, end_normal, end, R, length, lo, size):
    L = lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.oracle_cone_separated.restype = ctypes.c_int
    L.oracle_cone_separated.argtypes = [dp, dp, dp, ctypes.c_double, ctypes.c_double, dp, ctypes.c_double]
    a, ap = _d3(apex)
    n, np_ = _d3(end_normal)
    e, ep = _d3(end)
    l, lp = _d3(lo)
    return bool(L.oracle_cone_separated(ap, np_, ep, R, length, lp, size))


def uniform(depth):
    """Oracle handle holding a uniform tree of `depth` levels over [0,1)^3."""
    one = np.ones(1, np.float32)
    samples = {k: one.copy() * 0.5 for k in ("px", "py", "pz")}
    samples.update({"nx": one * 0, "ny": one * 0, "nz": one.copy(), "scale": one.copy() * 0.1,
                    "confidence": one.copy(), "parent_cameras": np.zeros(1, np.uint32)})
    o = Oracle(samples, {"x": one * 3, "y": one * 0, "z": one * 0})
    L = lib()
    L.oracle_uniform_tree.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    if L.oracle_uniform_tree(o.h, depth) != 0:
        raise OracleError(L.oracle_error(o.h).decode())
    return o


def split_tree(codes):
    """Oracle handle holding the tree over [0,1)^3 whose inner nodes are the given location
    codes ((1 << 3d) | Morton code; parents before children), with leaves and dual cells."""
    one = np.ones(1, np.float32)
    samples = {k: one.copy() * 0.5 for k in ("px", "py", "pz")}
    samples.update({"nx": one * 0, "ny": one * 0, "nz": one.copy(), "scale": one.copy() * 0.1,
                    "confidence": one.copy(), "parent_cameras": np.zeros(1, np.uint32)})
    o = Oracle(samples, {"x": one * 3, "y": one * 0, "z": one * 0})
    L = lib()
    L.oracle_split_tree.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32]
    c = np.ascontiguousarray(codes, np.uint64)
    if L.oracle_split_tree(o.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(c)) != 0:
        raise OracleError(L.oracle_error(o.h).decode())
    return o


def leaf_centres(o):
    L = lib()
    L.oracle_leaf_centre.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    n = o.count("leaves")
    out = np.zeros((n, 3))
    buf = np.zeros(3)
    for i in range(n):
        L.oracle_leaf_centre(o.h, i, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        out[i] = buf
    return out


EPS_F = float(np.float32(1e-6))


def occupancy_class(p, e, s, M):
    """Certified-priors judge restatement: what the mesh reads of the occupancy at prior p
    (0: -EPSILON, 1: (0, EPSILON], 2: above EPSILON; Occupancy.cpp:570-587), -1 within 1e-9 max(E, S)/M
    of a class boundary or below 1e-18 (arrays)"""
    p, e, s = (np.asarray(x, np.float64) for x in (p, e, s))
    pE, pS = p * (e / M), (1.0 - p) * (s / M)
    tol = 1e-9 * (np.maximum(e, s) / M)
    v = pS - pE
    c = np.full(v.shape, -1, np.int64)
    c[v < -tol] = 0
    c[(v > np.maximum(tol, 1e-18)) & (v < EPS_F - tol)] = 1
    c[v > EPS_F + tol] = 2
    c[np.isnan(v)] = -1
    return c


def certified(lo, hi, e, s, M):
    """the judge on intervals: the class where both widened ends agree, else -1"""
    lo = np.asarray(lo) * (1.0 - 1e-9)
    hi = np.minimum(1.0, np.asarray(hi) * (1.0 + 1e-9))
    a, b = occupancy_class(lo, e, s, M), occupancy_class(hi, e, s, M)
    ok = (a >= 0) & (a == b) & (lo >= 0) & (hi <= 1) & (lo <= hi)
    return np.where(ok, a, -1)


def erase_samples_in_nodes(counts, starts, inliers):
    """Nodes::eraseSamples (Nodes.cpp:644-671) with the sample offsets Scene::eraseSamples forms
    (Scene.cpp:270-289): offsets[i] = erased samples before i; per node the erased samples of
    its range leave its count, and its start moves back by the erased samples before it.
    Returns (counts, starts) after the erasure."""
    keep = np.asarray(inliers, bool)
    offsets = np.zeros(len(keep) + 1, np.int64)
    offsets[1:] = np.cumsum(~keep)
    counts = np.asarray(counts, np.int64)
    starts = np.asarray(starts, np.int64)
    ends = starts + counts
    deleted = offsets[ends] - offsets[starts]
    return (counts - deleted).astype(np.uint32), (starts - offsets[starts]).astype(np.uint32)
