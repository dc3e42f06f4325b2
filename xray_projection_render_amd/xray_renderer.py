"""XRayRenderer: the reference's Python binding, served by libxrp.

Same class, method and parameter names as the reference's
xray_projection_render/xray_renderer.py:63-448: ``render(params,
camera_angles)`` forwards a whitelisted JSON parameter set to the shared
library's ``RenderProjections`` (api.go:82-180) and returns its decoded
``RenderResult``.  The shared library is this repository's libxrp.so, whose
``RenderProjections`` renders on the GPU; there is no download fallback.
"""
from __future__ import annotations

import ctypes
import json
import os
from typing import Dict, List, Optional

from . import _lib


class XRayRenderer:
    """Python wrapper for the GPU renderer (xray_renderer.py:63)."""

    def __init__(self, library_path: Optional[str] = None):
        if library_path is None:
            library_path = _lib.LIB_PATH
        if not os.path.exists(library_path):
            raise FileNotFoundError(
                f"Library not found at {library_path}.\n\nBuild it with "
                "`python -m xray_projection_render_amd.build`.")
        self.lib = ctypes.CDLL(library_path)
        self._setup_function_signatures()

    def _setup_function_signatures(self):
        # xray_renderer.py:347-355
        self.lib.RenderProjections.argtypes = [ctypes.c_char_p]
        self.lib.RenderProjections.restype = ctypes.POINTER(ctypes.c_char)
        self.lib.FreeString.argtypes = [ctypes.POINTER(ctypes.c_char)]
        self.lib.FreeString.restype = None

    def render(self, params: Dict,
               camera_angles: Optional[List[Dict[str, float]]] = None) -> Dict:
        """Render X-ray projections (xray_renderer.py:357-448).

        params keys and defaults: input (required), output_dir "images",
        fname_pattern "image_%03d.png", resolution 512, num_images 1,
        out_of_plane False, ds -1.0, R 4.0, fov 40.0, jobs_modulo 1,
        job_num 0, transforms_file "transforms.json", deformation_file "",
        time_label 0.0, transparency False, export_volume False,
        polar_angle 90.0, density_multiplier 1.0, flat_field 0.0,
        integration "hierarchical", log_level "error".  camera_angles is a
        list of {"azimuthal": deg, "polar": deg} and overrides num_images.

        Returns {"success", "error" (on failure), "num_images", "output_dir"}.
        """
        render_params = {
            "input": params.get("input"),
            "output_dir": params.get("output_dir", "images"),
            "fname_pattern": params.get("fname_pattern", "image_%03d.png"),
            "resolution": params.get("resolution", 512),
            "num_images": params.get("num_images", 1),
            "out_of_plane": params.get("out_of_plane", False),
            "ds": params.get("ds", -1.0),
            "R": params.get("R", 4.0),
            "fov": params.get("fov", 40.0),
            "jobs_modulo": params.get("jobs_modulo", 1),
            "job_num": params.get("job_num", 0),
            "transforms_file": params.get("transforms_file", "transforms.json"),
            "deformation_file": params.get("deformation_file", ""),
            "time_label": params.get("time_label", 0.0),
            "transparency": params.get("transparency", False),
            "export_volume": params.get("export_volume", False),
            "polar_angle": params.get("polar_angle", 90.0),
            "density_multiplier": params.get("density_multiplier", 1.0),
            "flat_field": params.get("flat_field", 0.0),
            "integration": params.get("integration", "hierarchical"),
            "log_level": params.get("log_level", "error"),
            "camera_angles": [],
        }
        if camera_angles is not None:
            render_params["camera_angles"] = camera_angles
        elif "camera_angles" in params:
            render_params["camera_angles"] = params["camera_angles"]
        if not render_params["input"]:
            raise ValueError("'input' parameter is required")
        result_ptr = self.lib.RenderProjections(json.dumps(render_params).encode("utf-8"))
        try:
            result_str = ctypes.string_at(result_ptr).decode("utf-8")
        finally:
            self.lib.FreeString(result_ptr)
        return json.loads(result_str)
