"""MI355X-native X-ray projection renderer (HIP kernels behind a C ABI).

Drop-in for the per-pixel line-integral path of
igrega348/xray_projection_render: camera/ray generation, the fixed-step
Beer-Lambert march over the primitive tree and deformation fields, and the
trilinear voxel sampler run as gfx950 kernels in libxrp.so.  XRayRenderer
keeps the reference's Python interface.
"""
from ._lib import (Batch, Context, Scene, Surrogate, XrpError, camera_from_angles,
                   diag_deform_error, diag_l2_bandwidth, diag_mfma_bf16, diag_sincos_error,
                   export_volume, join_surrogate_params,
                   lib, render, render_device, render_device_noalloc, render_volume_projections,
                   sample_count,
                   split_surrogate_params)
from .xray_renderer import XRayRenderer

__all__ = ["Batch", "Context", "Scene", "Surrogate", "XRayRenderer", "XrpError",
           "camera_from_angles", "diag_deform_error", "diag_l2_bandwidth", "diag_mfma_bf16",
           "diag_sincos_error", "export_volume",
           "join_surrogate_params", "lib", "render", "render_device", "render_device_noalloc",
           "render_volume_projections", "sample_count", "split_surrogate_params"]
