"""tsr_dmc_extract_noalloc (DMC into caller-provided device buffers, temporaries in the
library's fixed scratch array) against tsr_dmc_extract, through the C ABI: bit-identical
vertices, scales, keys and triangles on two calls in a row, the second on a smaller
scene; capacities that are too small and meshes larger than the scratch maximum fail
with TSR_ERR_INVALID_ARGUMENT and report the mesh size."""
import ctypes

import numpy as np
import pytest

from tests.helpers import scene

pytestmark = pytest.mark.gpu


class DeviceBuffers:
    """Caller-owned device memory (hipMalloc through the HIP runtime)."""

    def __init__(self):
        from tsr_amd import _lib as L
        L.lib()
        # the HIP runtime the library itself loaded (a second copy would be a second runtime)
        path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64.so" in line)
        self.hip = ctypes.CDLL(path)
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(int(nbytes), 8)) == 0
        self.ptrs.append(p)
        return p

    def download(self, p, dtype, n):
        out = np.zeros(n, dtype)
        if n:
            assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device -> host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(ctx, tree, occ, dev, max_v, max_t):
    from tsr_amd import _lib as L
    pos, sc, keys, tris = dev.alloc(24 * max_v), dev.alloc(8 * max_v), dev.alloc(8 * max_v), dev.alloc(12 * max_t)
    nv, nt = ctypes.c_uint32(0), ctypes.c_uint32(0)
    st = L.lib().tsr_dmc_extract_noalloc(ctx.h, tree.h, occ.h, max_v, max_t, pos, sc, keys, tris, ctypes.byref(nv),
                                         ctypes.byref(nt))
    return st, nv.value, nt.value, (pos, sc, keys, tris)


@pytest.mark.parametrize("first,second", [((3, 0.25), (0, 0.5))], ids=["primitives-then-sphere"])
def test_noalloc_matches_allocating_extraction(gpu_ctx, first, second):
    from tsr_amd import DualMarchingCells, Occupancy, Scene, Tree
    dev = DeviceBuffers()
    try:
        sizes = []
        for cfg, res in (first, second):
            samples, cameras = scene(cfg, res)
            print(f"\n[noalloc] config={cfg} res={res} samples={len(samples['px'])}", flush=True)
            sc = Scene(gpu_ctx, samples, cameras)
            tree = Tree(gpu_ctx, sc)
            occ = Occupancy(gpu_ctx, tree, sc)
            ref = DualMarchingCells(gpu_ctx, tree, occ)
            want = ref.download()
            V, T = ref.vertex_count, ref.triangle_count
            st, nv, nt, (pos, scl, keys, tris) = _noalloc(gpu_ctx, tree, occ, dev, V + 16, T + 16)
            assert st == 0, f"status {st}"
            assert (nv, nt) == (V, T)
            assert np.array_equal(dev.download(pos, np.float64, 3 * V), want["positions"])
            assert np.array_equal(dev.download(scl, np.float64, V), want["scales"])
            assert np.array_equal(dev.download(keys, np.uint32, 2 * V), want["edge_keys"])
            assert np.array_equal(dev.download(tris, np.uint32, 3 * T), want["triangles"])
            sizes.append(T)
            # a capacity one short fails and still reports the size
            st, nv, nt, _ = _noalloc(gpu_ctx, tree, occ, dev, V, T - 1)
            assert st == 1 and nt == T
            for x in (ref, occ, tree, sc):
                x.close()
        assert sizes[1] < sizes[0]
    finally:
        dev.free()


def test_noalloc_rejects_meshes_beyond_the_scratch_maximum(gpu_ctx):
    from tsr_amd import Occupancy, Scene, Tree
    from tsr_amd import _lib as L
    samples, cameras = scene(1, 1.0)      # thin wedge + sphere: 4.6M triangles, ~0.5 GB of temporaries
    sc = Scene(gpu_ctx, samples, cameras)
    tree = Tree(gpu_ctx, sc)
    occ = Occupancy(gpu_ctx, tree, sc)
    dev = DeviceBuffers()
    try:
        st, nv, nt, _ = _noalloc(gpu_ctx, tree, occ, dev, 1 << 23, 1 << 23)
        assert st == 1 and b"no-allocation maximum" in L.lib().tsr_last_error()
        assert nt > 4_000_000
    finally:
        dev.free()
        for x in (occ, tree, sc):
            x.close()
