"""MI355X (gfx950) occupancy field + dual marching cells for TSR-style surface reconstruction.

The compute path is libtsr_amd.so (HIP kernels behind the C ABI in include/tsr_gpu.h);
this package is its Python binding.
"""
from ._lib import TsrError, default_params, lib  # noqa: F401
from .api import (Context, Culler, DepthMesh, DualMarchingCells, device_memory, read_samples_file, Occupancy, Scene, SyntheticScene, Tree,  # noqa: F401
                  init_mlp_weights, leaf_features, mlp_forward, mlp_gradient, mlp_train,
                  reconstruct, samples_from_depth_maps)
