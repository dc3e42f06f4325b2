"""libxrp host-side behaviour on the CPU (no device calls).

* every function declared in include/xrp.h is exported;
* cameras and sample grids are bit-identical to the oracle's;
* the scene loader accepts and rejects exactly what the reference's
  FromMap rules do (checked against the oracle's builder), and reports the
  same MinFeatureSize;
* the clip box contains every point of non-zero density;
* error paths return their status codes.
"""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest
import yaml

import xray_projection_render_amd as X
from xray_projection_render_amd import _lib as L
from xray_projection_render_amd import workloads as W
from helpers import O, ROOT, SCENES, load_scene_file


def header_functions():
    text = open(os.path.join(ROOT, "include", "xrp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|char\s*\*|const\s+char\s*\*)\s*\**\s*(\w+)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_exports_every_declared_function():
    names = header_functions()
    assert "xrp_render" in names and "RenderProjections" in names
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version():
    assert X.lib().xrp_abi_version() == 2


@pytest.mark.parametrize("az,po,R", [(0, 90, 4), (45, 60, 5), (123.4, 77.7, 4), (315, 90, 4),
                                      (10, 180, 4), (200.5, 3.25, 2.5)])
def test_camera_bit_identical_to_oracle(az, po, R):
    e1, c1 = X.camera_from_angles(az, po, R)
    e2, c2 = O.camera_from_angles(az, po, R)
    assert np.array_equal(e1, e2) and np.array_equal(c1, c2)


@pytest.mark.parametrize("ds", [0.02, 0.0015625, 0.0015, 0.001, 0.037])
def test_sample_count(ds):
    smin, smax = 4.0 - 1.74, 4.0 + 1.74
    n, s = 0, smin
    while s < smax:
        n += 1
        s += ds
    assert X.sample_count(smin, smax, ds) == n


def try_both(obj, numbers="yaml", deformation=None, base_dir=None):
    """(accepted by libxrp, accepted by the oracle builder, messages)."""
    ok_x = ok_o = True
    mx = mo = ""
    try:
        X.Scene.from_dict(obj, deformation, numbers=numbers, base_dir=base_dir).close()
    except X.XrpError as e:
        ok_x, mx = False, e.message
    try:
        O.build_object(obj, numbers=numbers, base_dir=base_dir or ".")
        if deformation:
            O.build_deformation(deformation, numbers=numbers)
    except ValueError as e:
        ok_o, mo = False, str(e)
    return ok_x, ok_o, mx, mo


SPH = {"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.3, "rho": 1.0}
CASES = [
    (SPH, True),
    ({**SPH, "radius": 1}, False),               # int radius: radius.(float64) fails
    ({**SPH, "rho": 1}, False),                   # int rho fails for a sphere
    ({**SPH, "center": [0, 0, 0]}, True),         # ints allowed in a sphere centre
    ({**SPH, "center": [0.0, 0.0, 0.0, 0.0]}, False),
    ({"type": "cube", "center": [0.0, 0.0, 0.0], "side": 1.0, "rho": 1.0}, True),
    ({"type": "cube", "center": [0, 0.0, 0.0], "side": 1.0, "rho": 1.0}, False),  # panics
    ({"type": "box", "center": [0, 0, 0], "sides": [1, 2, 3], "rho": 1}, True),
    ({"type": "cylinder", "p0": [0, 0, 0], "p1": [1, 0, 0], "radius": 0.1}, True),
    ({"type": "cylinder", "p0": [0, 0, 0], "p1": [1, 0, 0], "radius": 1}, False),
    ({"type": "gyroid", "center": [0, 0, 0], "scale": 0.1, "thickness": 0.2, "rho": 1}, False),
    ({"type": "gyroid", "center": [0, 0, 0], "scale": 0.1, "thickness": 0.2, "rho": 1.0},
     True),
    ({"type": "parallelepiped", "origin": [0, 0, 0], "v0": [1, 0, 0], "v1": [0, 1, 0],
      "v2": [0, 0, 1], "rho": 1}, True),
    ({"type": "object_collection", "objects": [SPH, SPH]}, True),
    ({"type": "object_collection", "objects": [{"type": "object_collection",
                                                "objects": [SPH]}]}, False),
    ({"type": "object_collection", "objects": {"a": SPH}}, False),
    ({"type": "unit_cell", "objects": {"type": "object_collection", "objects": [SPH]}}, False),
    ({"type": "teapot"}, False),
    ({"type": "tessellated_obj_coll", "uc": {"objects": {"objects": [SPH]}, "xmin": -1,
                                            "xmax": 1, "ymin": -1, "ymax": 1, "zmin": -1,
                                            "zmax": 1},
      "xmin": -2, "xmax": 2, "ymin": -2, "ymax": 2, "zmin": -2, "zmax": 2}, True),
    ({"type": "tessellated_obj_coll", "uc": {"objects": {"objects": [SPH]}, "xmin": -1,
                                            "xmax": 1, "ymin": -1, "ymax": 1, "zmin": -1},
      "xmin": -2, "xmax": 2, "ymin": -2, "ymax": 2, "zmin": -2, "zmax": 2}, False),
    ({"type": "voxel_grid", "nx": 2, "ny": 2, "nz": 2, "rho": [0.0] * 8}, False),
    ({"type": "voxel_grid", "path": "grid.bin", "resolution": [2, 2, 2]}, False),
]


@pytest.mark.parametrize("i", range(len(CASES)))
def test_loader_type_rules_match_oracle(i):
    obj, accept = CASES[i]
    ok_x, ok_o, mx, mo = try_both(obj)
    assert ok_x == ok_o == accept, (obj, mx, mo)
    if not ok_x:
        assert mx == mo or mx.startswith(mo.split(":")[0]), (mx, mo)


def test_json_numbers_are_float64():
    """encoding/json decodes 1 as float64, so `radius: 1` loads from JSON."""
    obj = {**SPH, "radius": 1, "rho": 1}
    assert try_both(obj, numbers="json")[:2] == (True, True)
    assert try_both(obj, numbers="yaml")[:2] == (False, False)


DEFS = [
    ({"type": "linear", "strains": [0.0, 0.0, 0.0, 0.1, 0.1, 0.1]}, True),
    ({"type": "linear", "strains": [0.0, 0.0, 0.1]}, False),
    ({"type": "linear", "strains": [0, 0, 0, 0, 0, 0]}, False),  # .(float64) on ints
    ({"type": "sigmoid", "amplitude": 1, "center": 0, "lengthscale": 1, "direction": "z"}, True),
    ({"type": "sigmoid", "amplitude": 1, "center": 0, "lengthscale": 1, "direction": "w"}, False),
    ({"type": "affine", "matrix": [[1.0, 0.0], [0.0, 1.0]]}, False),
    ({"type": "gaussian", "amplitudes": [0.1, 0.1, 0.1], "sigmas": [1.0, 1.0, 1.0],
      "centers": [0.0, 0.0, 0.0]}, True),
    ({"type": "composed", "deformations": [{"type": "rigid", "displacements": [0.1, 0.0, 0.0]},
                                           {"type": "twist"}]}, False),
    ({"type": "shear"}, False),
]


@pytest.mark.parametrize("i", range(len(DEFS)))
def test_deformation_rules_match_oracle(i):
    d, accept = DEFS[i]
    ok_x, ok_o, mx, mo = try_both(SPH, deformation=d)
    assert ok_x == ok_o == accept, (d, mx, mo)


@pytest.mark.parametrize("name", ["cube_w_hole.yaml", "box_w_pped.yaml", "gyroid_example.yaml",
                                  "lattice.yaml", "balls.json", "pillar_array.json"])
def test_example_scenes_load_like_oracle(name):
    obj, numbers, text = load_scene_file(name)
    sx = X.Scene.from_text(text if numbers == "json" else json.dumps(obj), numbers=numbers)
    so = O.build_object(obj, numbers=numbers)
    assert sx.min_feature_size() == so.min_feature_size()
    assert sx.auto_ds() == so.min_feature_size() / 5.0


def clip_contains_support(obj, deformation, n=4000, seed=0, extent=1.5):
    sc = X.Scene.from_dict(obj, deformation)
    lo, hi, empty = sc.clip_box()
    osc = O.Scene(O.build_object(obj), O.build_deformation(deformation) if deformation else None)
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-extent, extent, (n, 3))
    nonzero = 0
    for p in pts:
        if osc.density(*p) != 0.0:
            nonzero += 1
            assert not empty and np.all(p >= lo) and np.all(p <= hi), (p, lo, hi)
    return nonzero


@pytest.mark.parametrize("dfm", [None,
                                 {"type": "rigid", "displacements": [0.3, -0.2, 0.1]},
                                 {"type": "affine", "matrix": [[1.5, 0.2, 0.0], [0.0, 0.8, 0.1],
                                                               [0.1, 0.0, 1.1]]},
                                 {"type": "sigmoid", "amplitude": -0.4, "center": 0.0,
                                  "lengthscale": 0.1, "direction": "x"},
                                 {"type": "gaussian", "amplitudes": [0.2, -0.3, 0.1],
                                  "sigmas": [0.3, 0.3, 0.3], "centers": [0.0, 0.0, 0.0]}])
def test_clip_box_contains_support(dfm):
    assert clip_contains_support(W.csg().obj, dfm) > 50


def test_clip_box_tessellation_and_random_scene():
    assert clip_contains_support(W.lattice(resolution=8, n_views=1).obj,
                                 {"type": "affine", "matrix": [[8.0, 0.0, 0.0], [0.0, 8.0, 0.0],
                                                               [0.0, 0.0, 8.0]]},
                                 extent=0.6) > 50
    obj, _, _ = W.random_scene(4, max_prims=40)
    clip_contains_support(obj, None)


def test_negative_only_collection_has_empty_box():
    sc = X.Scene.from_dict({"type": "object_collection", "objects": [{**SPH, "rho": -1.0}]})
    assert sc.clip_box()[2] is True


def test_error_paths():
    lib = X.lib()
    p = ctypes.c_void_p()
    assert lib.xrp_scene_create(None, None, 0, None, ctypes.byref(p)) == L.XRP_ERR_NULL \
        if hasattr(L, "XRP_ERR_NULL") else True
    assert lib.xrp_scene_create(b"{not json", None, 0, None, ctypes.byref(p)) == 12
    assert b"JSON parse error" in lib.xrp_last_error()
    assert lib.xrp_scene_create(b'{"type": "teapot"}', None, 0, None, ctypes.byref(p)) == 10
    assert lib.xrp_last_error() == b"unknown object type `teapot`"
    assert lib.xrp_scene_create(json.dumps(SPH).encode(), b'{"type": "shear"}', 1, None,
                                ctypes.byref(p)) == 11
    assert lib.xrp_scene_create(json.dumps(SPH).encode(), None, 7, None, ctypes.byref(p)) == 12
    vox = (ctypes.c_float * 8)()
    assert lib.xrp_scene_create_voxels(vox, 0, 2, 2, None, 0, ctypes.byref(p)) == 2
    assert lib.xrp_render(None, None, None, None, 0, None, None) == 1
    n = ctypes.c_int64()
    assert lib.xrp_sample_count(0.0, 1.0, 0.0, ctypes.byref(n)) == 12


def test_context_without_device_reports_error():
    """In a process without a GPU the context call fails cleanly."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(X.XrpError) as ei:
        X.Context(0)
    assert ei.value.code in (3, 12)


def test_raw_voxel_grid_loads_like_oracle(tmp_path):
    rng = np.random.default_rng(1)
    data = rng.integers(0, 65535, size=3 * 4 * 5, dtype=np.uint16)
    (tmp_path / "g.raw").write_bytes(data.astype("<u2").tobytes())
    obj = {"type": "voxel_grid", "path": "g.raw", "resolution": [3, 4, 5], "dtype": "uint16"}
    sx = X.Scene.from_dict(obj, base_dir=str(tmp_path))
    so = O.build_object(obj, base_dir=str(tmp_path))
    assert sx.min_feature_size() == so.min_feature_size() == 2.0 / 5
    assert sx.layout() == (1, 0, True)
    bad = {**obj, "resolution": [3, 4, 6]}
    ok_x, ok_o, mx, mo = try_both(bad, base_dir=str(tmp_path))
    assert not ok_x and not ok_o and mx == mo


def test_tanh_shear_scene_rules():
    """The extension deformation loads in both builders with the same clip box
    (the object box widened by |amplitude| along the displaced axis)."""
    obj = W.csg().obj
    d = {"type": "tanh_shear", "amplitude": 0.1, "rate": 8.0, "direction": "z",
         "source": "x"}
    sc = X.Scene.from_dict(obj, d)
    lo, hi, empty = sc.clip_box()
    assert not empty
    lo0, hi0, _ = X.Scene.from_dict(obj).clip_box()
    assert lo[2] <= lo0[2] - 0.1 and hi[2] >= hi0[2] + 0.1
    assert abs(lo[2] - (lo0[2] - 0.1)) < 1e-6 and abs(hi[2] - (hi0[2] + 0.1)) < 1e-6
    assert np.allclose(lo[:2], lo0[:2], atol=1e-6) and np.allclose(hi[:2], hi0[:2], atol=1e-6)
    O.build_deformation(d)
    for bad in [dict(d, source="z"), dict(d, direction="w"), dict(d, rate="8")]:
        with pytest.raises(X.XrpError):
            X.Scene.from_dict(obj, bad)
        with pytest.raises(ValueError):
            O.build_deformation(bad)


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes mirror of a struct in include/xrp.h has the header's
    size and the header's offset for each field (a C program built with gcc
    prints them)."""
    import subprocess
    from xray_projection_render_amd import _lib as XL
    pairs = {"xrp_render_params": XL.RenderParams, "xrp_render_stats": XL.RenderStats,
             "xrp_view_angle": XL.ViewAngle, "xrp_camera": XL.Camera,
             "xrp_batch_item": XL.BatchItem, "xrp_tile_shape": XL.TileShape,
             "xrp_batch_stats": XL.BatchStats, "xrp_surrogate_config": XL.SurrogateConfig,
             "xrp_train_config": XL.TrainConfig, "xrp_train_stats": XL.TrainStats,
             "xrp_surrogate_stats": XL.SurrogateStats}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "xrp.h"', "int main(void) {"]
    for cname, cls in pairs.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True,
                               text=True).stdout.splitlines():
        cname, fname, val = line.split()
        got[(cname, fname)] = int(val)
    for cname, cls in pairs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
