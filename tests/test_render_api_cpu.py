"""RenderProjections' parameter decoding and error results (api.go:82-180),
the parts that finish before any GPU work starts."""
import json
import os

import pytest

import xray_projection_render_amd as X
from helpers import SCENES, load_scene_file

CUBE = os.path.join(SCENES, "cube_w_hole.yaml")


@pytest.fixture(scope="module")
def renderer():
    return X.XRayRenderer()


def raw_call(renderer, text):
    """RenderProjections on a raw parameter string, as a cgo caller sends it."""
    import ctypes
    ptr = renderer.lib.RenderProjections(text.encode())
    try:
        return json.loads(ctypes.string_at(ptr).decode())
    finally:
        renderer.lib.FreeString(ptr)


def test_ds_zero_rejected(renderer):
    r = renderer.render({"input": CUBE, "ds": 0})
    assert r == {"success": False,
                 "error": "ds is 0; use a negative value for automatic step-size selection",
                 "num_images": 0, "output_dir": ""}


def test_density_multiplier_zero_rejected(renderer):
    r = renderer.render({"input": CUBE, "density_multiplier": 0.0})
    assert not r["success"]
    assert r["error"].startswith("density_multiplier is 0;")


def test_missing_input_is_fatal(renderer, tmp_path):
    r = renderer.render({"input": str(tmp_path / "nope.yaml")})
    assert r["error"] == "render failed: fatal: error reading object file"
    r = renderer.render({"input": CUBE, "deformation_file": str(tmp_path / "d.yaml")})
    assert r["error"] == "render failed: fatal: error reading deformation file"


def test_bad_deformation_is_fatal(renderer, tmp_path):
    r = renderer.render({"input": CUBE, "output_dir": str(tmp_path / "images"),
                         "deformation_file": os.path.join(SCENES, "balls.json")})
    assert r["error"].startswith("render failed: fatal: Error loading deformation: ")


def test_python_wrapper_requires_input(renderer):
    with pytest.raises(ValueError):
        renderer.render({})


@pytest.mark.parametrize("text,msg", [
    ('{"resolution": "x", "ds": 1, "density_multiplier": 1}',
     "json: cannot unmarshal string into Go struct field RenderParams.resolution of type int"),
    ('{"resolution": 1.5}',
     "json: cannot unmarshal number 1.5 into Go struct field RenderParams.resolution of type int"),
    ('{"transparency": 1}',
     "json: cannot unmarshal number into Go struct field RenderParams.transparency of type bool"),
    ('{"camera_angles": {}}',
     "json: cannot unmarshal object into Go struct field RenderParams.camera_angles of type "
     "[]main.CameraAngle"),
])
def test_type_errors(renderer, text, msg):
    r = raw_call(renderer, text)
    assert r["success"] is False
    assert r["error"] == "Failed to parse parameters: " + msg


def test_keys_match_case_insensitively(renderer):
    # encoding/json matches "DS" to the field tagged "ds"
    r = raw_call(renderer, '{"DS": 0, "Density_Multiplier": 1}')
    assert r["error"].startswith("ds is 0")
    r = raw_call(renderer, '{"ds": 1, "DENSITY_MULTIPLIER": 0}')
    assert r["error"].startswith("density_multiplier is 0")


def test_result_json_layout(renderer):
    # field order and omitempty of RenderResult (api.go:67-72)
    import ctypes
    ptr = renderer.lib.RenderProjections(b'{"ds": 0}')
    try:
        text = ctypes.string_at(ptr).decode()
    finally:
        renderer.lib.FreeString(ptr)
    assert text == ('{"success":false,"error":"ds is 0; use a negative value for automatic '
                    'step-size selection","num_images":0,"output_dir":""}')


def test_object_json_and_empty_transforms(renderer, tmp_path):
    """With no views the flow still writes transforms.json (frames []) and
    object.json, the ToMap (objects.go:33-985) of the loaded scene, which
    reads back as JSON to the same scene.  No GPU work is needed."""
    name = "lattice.yaml"
    r = renderer.render({"input": os.path.join(SCENES, name), "num_images": 0,
                         "output_dir": str(tmp_path / "images"),
                         "transforms_file": str(tmp_path / "transforms.json")})
    assert r == {"success": True, "num_images": 0, "output_dir": str(tmp_path / "images")}
    assert (tmp_path / "images").is_dir()
    tj = json.loads((tmp_path / "transforms.json").read_text())
    assert tj["frames"] == [] and tj["fl_x"] == 0 and tj["w"] == 512
    text = (tmp_path / "object.json").read_text()
    om = json.loads(text)
    # json.MarshalIndent of map[string]interface{}: sorted keys, two spaces,
    # whole floats without a fraction
    assert text.startswith('{\n  "type": "tessellated_obj_coll",\n  "uc": {\n')
    assert '"rho": 1,' in text and '"radius": 0.025,' in text
    assert sorted(om) == ["type", "uc", "xmax", "xmin", "ymax", "ymin", "zmax", "zmin"]
    assert sorted(om["uc"]) == ["objects", "type", "xmax", "xmin", "ymax", "ymin", "zmax", "zmin"]
    assert om["uc"]["objects"]["type"] == "object_collection"
    src, _, _ = load_scene_file(name)
    kids = om["uc"]["objects"]["objects"]
    assert len(kids) == len(src["uc"]["objects"]["objects"])
    for got, want in zip(kids, src["uc"]["objects"]["objects"]):
        assert sorted(got) == ["p0", "p1", "radius", "rho", "type"]
        assert got["p0"] == want["p0"] and got["p1"] == want["p1"]
        assert got["radius"] == want["radius"]
    a = X.Scene.from_files(os.path.join(SCENES, name))
    b = X.Scene.from_files(str(tmp_path / "object.json"))
    assert a.min_feature_size() == b.min_feature_size()
    assert a.layout() == b.layout()
    assert all((x == y).all() for x, y in zip(a.clip_box()[:2], b.clip_box()[:2]))
