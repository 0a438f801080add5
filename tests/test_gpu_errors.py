"""The reference's failure sites (SURVEY of the path, exception list) forced through the
C ABI; each must return the status that maps the reference's exception:

  no samples / no cameras       Scene.cpp:137-151 (reconstruct returns false)
  non-positive sample scale     Samples.cpp:280-288
  sample outside the root       Nodes.cpp:237-238 (a NaN position is in no node)
  tree deeper than 19 levels    (location codes + depth in 64 bits; the reference's
                                 float sizes underflow instead)
  more than 2^32 - 1 links      Samples.h:395-402
Non-manifold DMC edges and corrupt files are forced in test_gpu_storage.py; node-count
overflow (2^32 nodes), zero-size nodes and unplaced samples cannot be reached with inputs
that fit a device and stay untested.
"""
import ctypes

import numpy as np
import pytest

from tsr_amd import _lib as L

pytestmark = pytest.mark.gpu


def _samples(p, scale, normal=(0.0, 0.0, 1.0)):
    p = np.asarray(p, np.float32).reshape(-1, 3)
    n = len(p)
    return {"px": p[:, 0].copy(), "py": p[:, 1].copy(), "pz": p[:, 2].copy(),
            "nx": np.full(n, normal[0], np.float32), "ny": np.full(n, normal[1], np.float32),
            "nz": np.full(n, normal[2], np.float32), "scale": np.asarray(scale, np.float32).reshape(-1).copy(),
            "confidence": np.ones(n, np.float32), "parent_cameras": np.zeros(n, np.uint32)}


CAM = {"x": np.array([5.0], np.float32), "y": np.zeros(1, np.float32), "z": np.zeros(1, np.float32)}


def _status_of(fn):
    from tsr_amd import TsrError
    with pytest.raises(TsrError) as e:
        fn()
    return e.value.status_name


def test_no_samples_and_no_cameras(gpu_ctx):
    from tsr_amd import Scene
    empty = {k: v[:0] for k, v in _samples([[0, 0, 0]], [0.1]).items()}
    assert _status_of(lambda: Scene(gpu_ctx, empty, CAM)) == "TSR_ERR_NO_SAMPLES"
    nocam = {k: v[:0] for k, v in CAM.items()}
    assert _status_of(lambda: Scene(gpu_ctx, _samples([[0, 0, 0]], [0.1]), nocam)) == "TSR_ERR_NO_SAMPLES"


@pytest.mark.parametrize("bad", [0.0, -1.0])
def test_non_positive_scale(gpu_ctx, bad):
    from tsr_amd import Scene, Tree
    sc = Scene(gpu_ctx, _samples([[0, 0, 0], [1, 0, 0]], [0.1, bad]), CAM)
    assert _status_of(lambda: Tree(gpu_ctx, sc)) == "TSR_ERR_INVALID_SCALE"
    sc.close()


def test_sample_outside_the_root(gpu_ctx):
    from tsr_amd import Scene, Tree
    sc = Scene(gpu_ctx, _samples([[0, 0, 0], [np.nan, 0, 0], [1, 1, 1]], [0.1, 0.1, 0.1]), CAM)
    assert _status_of(lambda: Tree(gpu_ctx, sc)) == "TSR_ERR_SAMPLE_OUTSIDE_ROOT"
    sc.close()


def test_tree_too_deep(gpu_ctx):
    from tsr_amd import Scene, Tree
    # extent 1000, max sampling distance 2e-4: the container would be ~22 levels down
    sc = Scene(gpu_ctx, _samples([[0, 0, 0], [1000, 0, 0]], [1e-4, 1e-4]), CAM)
    assert _status_of(lambda: Tree(gpu_ctx, sc)) == "TSR_ERR_TREE_TOO_DEEP"
    sc.close()


def test_too_many_links_is_refused_before_reading(gpu_ctx):
    # 2^22 samples with 1024 camera slots each: 2^32 links; the arrays are never read
    v = L.SamplesView()
    v.count = 1 << 22
    v.max_cameras_per_sample = 1024
    one = np.zeros(1, np.float32)
    for k in ("px", "py", "pz", "nx", "ny", "nz", "scale", "confidence"):
        setattr(v, k, one.ctypes.data_as(L.f32p))
    cams = np.zeros(1, np.uint32)
    v.parent_cameras = cams.ctypes.data_as(L.u32p)
    cv = L.CamerasView()
    cv.count = 1
    cv.x = cv.y = cv.z = one.ctypes.data_as(L.f32p)
    h = L.vp()
    st = L.lib().tsr_scene_upload(gpu_ctx.h, ctypes.byref(v), ctypes.byref(cv), ctypes.byref(h))
    assert L.STATUS_NAMES[st] == "TSR_ERR_LINK_OVERFLOW"
