"""C ABI library on the host (no GPU compute): it loads, exports every entry
point declared in include/tsr_gpu.h, validates arguments, and the host-side
synthetic scene generator is deterministic and matches its specification."""
import ctypes
import os
import re

import numpy as np
import pytest

import tsr_amd
from tsr_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(tsr_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = declared_entry_points()
    assert len(names) > 25
    for n in sorted(names):
        assert hasattr(lib, n), n
    bound = {s[0] for s in L.SIGNATURES}
    assert names == bound, names ^ bound


def test_params_default_match_app_cfg():
    p = L.default_params()
    assert (p.bandwidth_factor, p.confidence_threshold, p.relative_radius_for_priors,
            p.relative_sample_cone_length, p.max_relative_sampling_distance) == (2.0, 100.0, 25.0, 1.0, 2.0)


def test_argument_validation_without_device():
    lib = L.lib()
    assert lib.tsr_tree_build(None, None, None, None) == 1
    assert lib.tsr_occupancy_compute(None, None, None, None, 0, None, None) == 1
    assert lib.tsr_dmc_extract(None, None, None, None) == 1
    assert b"null" in lib.tsr_last_error()
    assert lib.tsr_mlp_forward(None, None, None, 0, None) == 1
    assert lib.tsr_mlp_gradient(None, None, None, None, None, 0, None, None) == 1
    assert lib.tsr_mlp_train(None, None, None, None, None, 0, 1, 1, 0.0, 0, None) == 1
    assert lib.tsr_culler_create(None, None, 0.5, ctypes.byref(L.vp())) == 1
    assert lib.tsr_leaf_features(None, None, None, None) == 1
    assert lib.tsr_field_prior_levels(None, None, None) == 1
    h = L.vp()
    assert lib.tsr_synthetic_create(7, 0, 1.0, ctypes.byref(h)) == 1     # unknown config
    assert lib.tsr_synthetic_create(0, 0, 0.0, ctypes.byref(h)) == 1     # bad resolution


def test_synthetic_sphere_scene():
    s = tsr_amd.SyntheticScene(0, seed=0, resolution_scale=0.5)
    p = np.stack([s.samples["px"], s.samples["py"], s.samples["pz"]], 1).astype(np.float64)
    n = np.stack([s.samples["nx"], s.samples["ny"], s.samples["nz"]], 1).astype(np.float64)
    cams = np.stack([s.cameras["x"], s.cameras["y"], s.cameras["z"]], 1)
    assert len(cams) == 12 and np.allclose(np.linalg.norm(cams, axis=1), 3.0, atol=1e-6)
    r = np.linalg.norm(p, axis=1)
    # depth noise sigma = 0.5 % of depth (<= 2 units): samples within a few sigma of the sphere
    assert np.all(np.abs(r - 1.0) < 0.06) and abs(np.mean(r) - 1.0) < 2e-3
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
    # normal = ground truth at the noiseless hit, facing the camera
    c = cams[s.samples["parent_cameras"]]
    assert np.all(np.sum(n * (c - p), axis=1) > 0)
    # scale = pixel footprint d' * 2 tan(30 deg) / W
    d = np.linalg.norm(p - c, axis=1)
    assert np.allclose(s.samples["scale"], d * 2 * np.tan(np.pi / 6) / 64, rtol=1e-5)
    assert np.all(s.samples["confidence"] == 1.0)


def test_synthetic_is_seeded_and_has_outliers():
    a = tsr_amd.SyntheticScene(1, seed=0, resolution_scale=0.05)
    b = tsr_amd.SyntheticScene(1, seed=0, resolution_scale=0.05)
    c = tsr_amd.SyntheticScene(1, seed=1, resolution_scale=0.05)
    assert np.array_equal(a.samples["px"], b.samples["px"])
    assert a.count != c.count or not np.array_equal(a.samples["px"], c.samples["px"])
    assert len(a.cameras["x"]) == 64
    cams = np.stack([a.cameras["x"], a.cameras["y"], a.cameras["z"]], 1)
    dist = np.linalg.norm(cams, axis=1)
    assert dist.min() >= 2.0 - 1e-5 and dist.max() <= 8.0 + 1e-5
    # outliers: about 1 % of samples do not face their camera or lie off the primitives
    p = np.stack([a.samples["px"], a.samples["py"], a.samples["pz"]], 1)
    n = np.stack([a.samples["nx"], a.samples["ny"], a.samples["nz"]], 1)
    facing = np.sum(n * (cams[a.samples["parent_cameras"]] - p), axis=1) > 0
    assert 0.001 < 1.0 - facing.mean() < 0.02


def test_samples_file_errors_without_device(tmp_path):
    """Samples::loadFromFile's FileCorruptionException cases map to TSR_ERR_FILE (host only)."""
    lib = L.lib()
    n, mc = ctypes.c_uint32(0), ctypes.c_uint32(0)
    missing = str(tmp_path / "missing.samples").encode()
    assert lib.tsr_samples_file_info(missing, ctypes.byref(n), ctypes.byref(mc)) == 13
    bad = tmp_path / "bad.samples"
    bad.write_bytes(b"XXXX" + bytes(12))
    assert lib.tsr_samples_file_info(str(bad).encode(), ctypes.byref(n), ctypes.byref(mc)) == 13
    assert b"not a file of this kind" in lib.tsr_last_error()
    short = tmp_path / "short.samples"
    short.write_bytes(b"TSRF" + np.array([5, 0, 4, 10, 1], np.uint32).tobytes())   # 10 samples promised
    assert lib.tsr_samples_file_info(str(short).encode(), ctypes.byref(n), ctypes.byref(mc)) == 0
    assert (n.value, mc.value) == (10, 1)
    buf = np.zeros(10, np.float32)
    assert lib.tsr_samples_file_read(str(short).encode(), *([buf.ctypes.data_as(L.f32p)] * 8), None) == 13
    assert b"too short" in lib.tsr_last_error()


def test_cameras_file_round_trip_without_device(tmp_path):
    """Cameras::saveToFile / loadFromFile (Cameras.cpp:116-177): count, then 13 Reals and the view
    ID per camera as CameraData lays them out (CameraData.h:21-36)"""
    from tsr_amd.api import read_cameras_file, write_cameras_file
    rng = np.random.default_rng(0)
    data = rng.normal(size=(3, 13))
    ids = np.array([4, 9, 2], np.uint32)
    p4, p8 = tmp_path / "a.cameras", tmp_path / "b.cameras"
    write_cameras_file(p4, data, ids, real_bytes=4)
    write_cameras_file(p8, data, ids, real_bytes=8)
    raw = p4.read_bytes()
    assert len(raw) == 16 + 4 + 3 * (13 * 4 + 4)
    assert raw[:4] == b"TSRF" and np.frombuffer(raw[4:16], np.uint32).tolist() == [6, 0, 4]
    rec = raw[20:20 + 56]
    assert np.array_equal(np.frombuffer(rec[:52], np.float32), data[0].astype(np.float32))
    assert np.frombuffer(rec[52:56], np.uint32)[0] == 4
    assert len(p8.read_bytes()) == 16 + 4 + 3 * (13 * 8 + 8)
    d4, i4 = read_cameras_file(p4)
    d8, i8 = read_cameras_file(p8)
    assert np.array_equal(d4, data.astype(np.float32).astype(np.float64)) and np.array_equal(i4, ids)
    assert np.array_equal(d8, data) and np.array_equal(i8, ids)
    bad = tmp_path / "c.cameras"
    bad.write_bytes(raw[:-3])
    with pytest.raises(L.TsrError) as e:
        read_cameras_file(bad)
    assert e.value.status_name == "TSR_ERR_FILE"
