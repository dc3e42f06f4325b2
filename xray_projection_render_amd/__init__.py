"""MI355X-native X-ray projection renderer (HIP kernels behind a C ABI).

Drop-in for the per-pixel line-integral path of
igrega348/xray_projection_render: camera/ray generation, the fixed-step
Beer-Lambert march over the primitive tree and deformation fields, and the
trilinear voxel sampler run as gfx950 kernels in libxrp.so.
"""
from ._lib import (Context, Scene, XrpError, camera_from_angles, lib, render,
                   render_device, render_volume_projections, sample_count)

__all__ = ["Context", "Scene", "XrpError", "camera_from_angles", "lib", "render",
           "render_device", "render_volume_projections", "sample_count"]
