"""ctypes binding of libtsr_amd.so (the C ABI in include/tsr_gpu.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C tsr_amd/csrc``).
Loading fails loudly when it is missing: there is no CPU fallback for the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtsr_amd.so")

u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)
u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
u8p = ctypes.POINTER(ctypes.c_uint8)
vp = ctypes.c_void_p


class Params(ctypes.Structure):
    _fields_ = [
        ("bandwidth_factor", ctypes.c_float),
        ("confidence_threshold", ctypes.c_float),
        ("relative_radius_for_priors", ctypes.c_float),
        ("relative_sample_cone_length", ctypes.c_float),
        ("max_relative_sampling_distance", ctypes.c_float),
    ]


class SamplesView(ctypes.Structure):
    _fields_ = [
        ("count", u32),
        ("max_cameras_per_sample", u32),
        ("px", f32p), ("py", f32p), ("pz", f32p),
        ("nx", f32p), ("ny", f32p), ("nz", f32p),
        ("scale", f32p),
        ("confidence", f32p),
        ("parent_cameras", u32p),
    ]


class CamerasView(ctypes.Structure):
    _fields_ = [("count", u32), ("x", f32p), ("y", f32p), ("z", f32p)]


class TreeInfo(ctypes.Structure):
    _fields_ = [
        ("node_count", u32),
        ("leaf_count", u32),
        ("cell_count", u32),
        ("max_depth", u32),
        ("root_min", ctypes.c_float * 3),
        ("root_size", ctypes.c_float),
    ]


class MeshInfo(ctypes.Structure):
    _fields_ = [("vertex_count", u32), ("triangle_count", u32)]


STATUS_NAMES = {
    0: "TSR_OK",
    1: "TSR_ERR_INVALID_ARGUMENT",
    2: "TSR_ERR_NO_SAMPLES",
    3: "TSR_ERR_SAMPLE_OUTSIDE_ROOT",
    4: "TSR_ERR_NODE_OVERFLOW",
    5: "TSR_ERR_TREE_TOO_DEEP",
    6: "TSR_ERR_LINK_OVERFLOW",
    7: "TSR_ERR_NON_MANIFOLD_EDGE",
    8: "TSR_ERR_HIP",
    9: "TSR_ERR_OUT_OF_MEMORY",
    10: "TSR_ERR_ZERO_SIZE_NODE",
    11: "TSR_ERR_SAMPLE_NOT_INSERTED",
    12: "TSR_ERR_INVALID_SCALE",
    13: "TSR_ERR_FILE",
}

TSR_ACCUM_KAHAN_SEGMENTED = 0
TSR_ACCUM_FP32_ATOMIC = 1
TSR_CULL_LEARNED = 0x100
TSR_PRIORS_CERTIFIED = 0x400
TSR_PRIORS_ALL_LEAVES = 0x200
TSR_EMPTINESS_TRAVERSAL = 0x800
MLP_PARAM_COUNT = 64 * 16 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1

# (name, restype, argtypes) of every exported entry point of include/tsr_gpu.h
SIGNATURES = [
    ("tsr_params_default", None, [ctypes.POINTER(Params)]),
    ("tsr_last_error", ctypes.c_char_p, []),
    ("tsr_context_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
    ("tsr_context_destroy", None, [vp]),
    ("tsr_context_synchronize", ctypes.c_int, [vp]),
    ("tsr_context_phase_times", ctypes.c_int, [vp, f32p, ctypes.c_int]),
    ("tsr_context_kernel_stats", ctypes.c_int, [vp, ctypes.c_int, f64p, u64p, u64p, u64p]),
    ("tsr_context_reset_kernel_stats", ctypes.c_int, [vp]),
    ("tsr_calibration_copy", ctypes.c_int, [vp, u64]),
    ("tsr_device_memory", ctypes.c_int, [ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.c_int]),
    ("tsr_scene_upload", ctypes.c_int, [vp, ctypes.POINTER(SamplesView), ctypes.POINTER(CamerasView), ctypes.POINTER(vp)]),
    ("tsr_scene_destroy", None, [vp]),
    ("tsr_scene_download", ctypes.c_int, [vp, vp, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, u32p]),
    ("tsr_tree_build", ctypes.c_int, [vp, vp, ctypes.POINTER(Params), ctypes.POINTER(vp)]),
    ("tsr_tree_get_info", ctypes.c_int, [vp, ctypes.POINTER(TreeInfo)]),
    ("tsr_tree_build_stats", ctypes.c_int, [vp, u64p]),
    ("tsr_tree_download", ctypes.c_int, [vp, vp, u32p, u32p, u32p, u32p, u32p, u32p, f32p, u32p, u32p, u32p]),
    ("tsr_tree_samples_download", ctypes.c_int, [vp, vp, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, u32p]),
    ("tsr_tree_destroy", None, [vp]),
    ("tsr_tree_neighbor_count", ctypes.c_int, [vp, vp, u64p]),
    ("tsr_tree_neighbors_download", ctypes.c_int, [vp, vp, u32p, u32p]),
    ("tsr_tree_save", ctypes.c_int, [vp, vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    ("tsr_tree_load", ctypes.c_int, [vp, vp, ctypes.POINTER(Params), ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.c_char_p, ctypes.POINTER(vp)]),
    ("tsr_tree_samples_save", ctypes.c_int, [vp, vp, ctypes.c_char_p]),
    ("tsr_samples_file_info", ctypes.c_int, [ctypes.c_char_p, u32p, u32p]),
    ("tsr_samples_file_read", ctypes.c_int, [ctypes.c_char_p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p,
                                             u32p]),
    ("tsr_field_save", ctypes.c_int, [vp, vp, ctypes.c_char_p, ctypes.c_int]),
    ("tsr_field_load", ctypes.c_int, [vp, vp, ctypes.POINTER(Params), ctypes.c_char_p, ctypes.POINTER(vp)]),
    ("tsr_occupancy_compute", ctypes.c_int, [vp, vp, vp, ctypes.POINTER(Params), ctypes.c_int, vp, ctypes.POINTER(vp)]),
    ("tsr_field_download", ctypes.c_int, [vp, vp, f32p, f32p, f64p, u32p, f64p, f64p]),
    ("tsr_field_hit_counts", ctypes.c_int, [vp, u64p]),
    ("tsr_occupancy_add_samples", ctypes.c_int, [vp, vp, vp, vp, u32p, ctypes.c_uint32]),
    ("tsr_occupancy_erase_samples", ctypes.c_int, [vp, vp, vp, vp, u32p, ctypes.c_uint32]),
    ("tsr_tree_erase_samples", ctypes.c_int, [vp, vp, u8p]),
    ("tsr_field_prior_stages", ctypes.c_int, [vp, u32p]),
    ("tsr_field_prior_levels", ctypes.c_int, [vp, vp, u8p]),
    ("tsr_field_destroy", None, [vp]),
    ("tsr_dmc_extract", ctypes.c_int, [vp, vp, vp, ctypes.POINTER(vp)]),
    ("tsr_dmc_extract_values", ctypes.c_int, [vp, vp, f64p, ctypes.POINTER(vp)]),
    ("tsr_dmc_extract_noalloc", ctypes.c_int, [vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp, vp,
                                               u32p, u32p]),
    ("tsr_crust_postprocess", ctypes.c_int, [vp, vp, ctypes.c_uint32]),
    ("tsr_mesh_download_normals", ctypes.c_int, [vp, vp, f64p]),
    ("tsr_mesh_get_info", ctypes.c_int, [vp, ctypes.POINTER(MeshInfo)]),
    ("tsr_mesh_download", ctypes.c_int, [vp, vp, f64p, f64p, u32p, u32p]),
    ("tsr_mesh_destroy", None, [vp]),
    ("tsr_culler_create", ctypes.c_int, [vp, f32p, ctypes.c_float, ctypes.POINTER(vp)]),
    ("tsr_culler_set_weights", ctypes.c_int, [vp, vp, f32p, ctypes.c_float]),
    ("tsr_culler_destroy", None, [vp]),
    ("tsr_leaf_features", ctypes.c_int, [vp, vp, vp, f32p]),
    ("tsr_mlp_forward", ctypes.c_int, [vp, f32p, f32p, ctypes.c_uint32, f32p]),
    ("tsr_mlp_gradient", ctypes.c_int, [vp, f32p, f32p, f32p, f32p, ctypes.c_uint32, f32p, f64p]),
    ("tsr_mlp_train", ctypes.c_int, [vp, f32p, f32p, f32p, f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_float, u64, f32p]),
    ("tsr_cameras_file_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, f64p, u32p, ctypes.c_int]),
    ("tsr_cameras_file_info", ctypes.c_int, [ctypes.c_char_p, u32p]),
    ("tsr_cameras_file_read", ctypes.c_int, [ctypes.c_char_p, f64p, u32p]),
    ("tsr_depth_mesh_create", ctypes.c_int, [vp, ctypes.c_void_p, ctypes.POINTER(vp)]),
    ("tsr_depth_mesh_info", ctypes.c_int, [vp, u32p, u32p, u32p]),
    ("tsr_depth_mesh_download", ctypes.c_int, [vp, vp, f64p, f64p, f64p, u32p, u32p, u32p]),
    ("tsr_depth_mesh_destroy", None, [vp]),
    ("tsr_synthetic_create", ctypes.c_int, [ctypes.c_int, u64, ctypes.c_double, ctypes.POINTER(vp)]),
    ("tsr_synthetic_views", ctypes.c_int, [vp, ctypes.POINTER(SamplesView), ctypes.POINTER(CamerasView)]),
    ("tsr_synthetic_destroy", None, [vp]),
]

_lib = None


class TsrError(RuntimeError):
    """Error status returned by the C ABI (message from tsr_last_error)."""

    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.status_name = STATUS_NAMES.get(status, str(status))


def lib():
    """Load libtsr_amd.so once; raises if the library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C tsr_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().tsr_last_error()
        raise TsrError(status, msg.decode() if msg else "")


def default_params():
    p = Params()
    lib().tsr_params_default(ctypes.byref(p))
    return p


class DepthView(ctypes.Structure):
    """tsr_depth_view"""
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("depths", ctypes.POINTER(ctypes.c_float)),
                ("position", ctypes.c_float * 3), ("rotation", ctypes.c_float * 9), ("focal_length", ctypes.c_float),
                ("principal_point", ctypes.c_float * 2), ("aspect_ratio", ctypes.c_float),
                ("camera_index", ctypes.c_uint32), ("views_channels", ctypes.c_uint32),
                ("views", ctypes.POINTER(ctypes.c_uint32))]
