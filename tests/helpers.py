"""Shared helpers: build the same scene for libxrp and for the FP64 oracle."""
from __future__ import annotations

import os
import sys
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if os.path.join(ROOT, "oracle") not in sys.path:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")
SCENES = os.path.join(GOLDEN, "scenes")


def oracle_scene(obj: Optional[Dict], deformation: Optional[Dict] = None,
                 voxels: Optional[np.ndarray] = None, numbers: str = "yaml",
                 density_multiplier: float = 1.0, flat_field: float = 0.0,
                 base_dir: str = ".") -> O.Scene:
    if voxels is not None:
        nz, nx, ny = voxels.shape
        node = O.voxel_node(voxels.astype(np.float64).reshape(-1), nx, ny, nz)
    else:
        node = O.build_object(obj, numbers=numbers, base_dir=base_dir)
    d = O.build_deformation(deformation, numbers=numbers) if deformation else None
    return O.Scene(node, d, density_multiplier, flat_field)


def oracle_render(sc: O.Scene, angles: Sequence[Tuple[float, float]], res: int, ds: float,
                  integrator: str = "simple", R: float = 4.0, fov: float = 40.0,
                  rows=None) -> np.ndarray:
    az = [a for a, _ in angles]
    po = [p for _, p in angles]
    return sc.render(az, po, res, ds, R=R, fov=fov, integrator=integrator, rows=rows)


def xrp_scene(obj: Optional[Dict], deformation: Optional[Dict] = None,
              voxels: Optional[np.ndarray] = None, numbers: str = "yaml",
              base_dir: Optional[str] = None):
    import xray_projection_render_amd as X
    if voxels is not None:
        return X.Scene.from_voxels(voxels, deformation, numbers=numbers)
    return X.Scene.from_dict(obj, deformation, numbers=numbers, base_dir=base_dir)


def load_scene_file(name: str):
    """Scene fixtures under tests/golden/scenes (YAML or JSON)."""
    import json

    import yaml
    path = os.path.join(SCENES, name)
    with open(path) as fh:
        text = fh.read()
    if name.endswith(".json"):
        return json.loads(text), "json", text
    return yaml.safe_load(text), "yaml", text
