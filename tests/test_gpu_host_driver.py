"""RenderProjections end to end on the GPU (api.go:82-180, main.go:310-636):
PNG frames against the oracle's transmittance quantized by Go's rules,
transforms.json, object.json and volume.raw."""
import json
import os

import numpy as np
import pytest

import xray_projection_render_amd as X
from helpers import O, SCENES, load_scene_file, oracle_render, oracle_scene

pytestmark = pytest.mark.gpu

# Device exp() and the host's libm exp() may differ by one ulp, which moves
# a pixel across a 1/65535 quantization boundary with probability ~1e-12;
# the tests allow that many one-level differences and no larger ones.
MAX_OFF_BY_ONE = 2


@pytest.fixture(scope="module")
def renderer():
    return X.XRayRenderer()


def go_u16(val):
    """uint16(val * 0xffff) of main.go:496 (amd64: int64 conversion, low 16 bits)."""
    x = val * 65535.0
    out = np.zeros(x.shape, dtype=np.int64)
    ok = np.isfinite(x) & (np.abs(x) < 2.0 ** 63)
    out[ok] = np.trunc(x[ok]).astype(np.int64)
    out[~ok] = np.iinfo(np.int64).min
    return out & 0xFFFF


def go_byte(x):
    """byte(x) of a float64 (main.go:620-623)."""
    out = np.zeros(x.shape, dtype=np.int64)
    ok = np.isfinite(x) & (np.abs(x) < 2.0 ** 63)
    out[ok] = np.trunc(x[ok]).astype(np.int64)
    out[~ok] = np.iinfo(np.int64).min
    return (out & 0xFF).astype(np.uint8)


def expected_rgba(val, transparency):
    """image.RGBA of main.go:482-500 as png.Encode stores it: [y][x][c]."""
    res = val.shape[0]
    g = (go_u16(val) >> 8).astype(np.uint8)
    a = np.full_like(g, 255)
    if transparency:
        a[~(val < 1.0)] = 0
        g = np.where(a == 0, 0, g).astype(np.uint8)
    img = np.zeros((res, res, 4), dtype=np.uint8)
    for c in range(3):
        img[:, :, c] = g.T[::-1, :]   # pixel (x=i, y=res-1-j)
    img[:, :, 3] = a.T[::-1, :]
    return img


def read_png(path):
    from PIL import Image
    with Image.open(path) as im:
        mode = im.mode
        arr = np.asarray(im.convert("RGBA"))
    return mode, arr


def assert_pixels(got, want):
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    assert d.max() <= 1, f"max pixel difference {d.max()}"
    assert int((d > 0).any(axis=2).sum()) <= MAX_OFF_BY_ONE


def run(renderer, tmp_path, **params):
    p = {"output_dir": str(tmp_path / "images"),
         "transforms_file": str(tmp_path / "transforms.json")}
    p.update(params)
    r = renderer.render(p)
    assert r["success"], r
    return r


@pytest.mark.parametrize("integration", ["simple", "hierarchical"])
def test_png_frames_match_oracle(renderer, tmp_path, integration):
    name = "cube_w_hole.yaml"
    obj, numbers, _ = load_scene_file(name)
    res, n, ds = 40, 3, 0.01
    r = run(renderer, tmp_path, input=os.path.join(SCENES, name), resolution=res, num_images=n,
            ds=ds, integration=integration)
    assert r["num_images"] == n and r["output_dir"] == str(tmp_path / "images")
    angles = O.generate_camera_angles(n)
    ref = oracle_render(oracle_scene(obj, numbers=numbers), angles, res, ds,
                        integrator=integration)
    for k in range(n):
        mode, img = read_png(tmp_path / "images" / f"image_{k:03d}.png")
        assert mode == "RGB"  # every alpha opaque: png.Encode writes 8-bit RGB
        assert_pixels(img, expected_rgba(ref[k], False))


def test_transparency_flat_field_multiplier(renderer, tmp_path):
    name = "box_w_pped.yaml"
    obj, numbers, _ = load_scene_file(name)
    res, ds, ff, dm = 36, 0.01, 0.0, 2.5
    angles = [(30.0, 70.0), (200.0, 120.0)]
    run(renderer, tmp_path, input=os.path.join(SCENES, name), resolution=res, ds=ds,
        integration="simple", transparency=True, flat_field=ff, density_multiplier=dm,
        camera_angles=[{"azimuthal": a, "polar": p} for a, p in angles], fname_pattern="f%d.png")
    ref = oracle_render(oracle_scene(obj, numbers=numbers, density_multiplier=dm, flat_field=ff),
                        angles, res, ds)
    assert (ref >= 1.0).any() and (ref < 1.0).any()
    for k in range(len(angles)):
        mode, img = read_png(tmp_path / "images" / f"f{k}.png")
        assert mode == "RGBA"
        assert_pixels(img, expected_rgba(ref[k], True))


def test_flat_field_darkens_every_pixel(renderer, tmp_path):
    name = "balls.json"
    obj, numbers, _ = load_scene_file(name)
    res, ds, ff = 32, 0.02, 0.4
    run(renderer, tmp_path, input=os.path.join(SCENES, name), resolution=res, ds=ds,
        integration="simple", flat_field=ff, num_images=1)
    ref = oracle_render(oracle_scene(obj, numbers=numbers, flat_field=ff),
                        O.generate_camera_angles(1), res, ds)
    assert ref.max() <= np.exp(-ff) + 1e-15
    _, img = read_png(tmp_path / "images" / "image_000.png")
    assert_pixels(img, expected_rgba(ref[0], False))
    tj = json.loads((tmp_path / "transforms.json").read_text())
    assert tj["flat_field"] == np.exp(-ff)


def test_transforms_json(renderer, tmp_path):
    name = "gyroid_example.yaml"
    res, fov, R, t = 24, 35.0, 3.5, 0.25
    angles = [(10.0, 90.0), (100.0, 45.0), (250.0, 135.0)]
    run(renderer, tmp_path, input=os.path.join(SCENES, name), resolution=res, ds=0.05,
        fov=fov, R=R, time_label=t, integration="simple",
        camera_angles=[{"azimuthal": a, "polar": p} for a, p in angles])
    text = (tmp_path / "transforms.json").read_text()
    tj = json.loads(text)
    # struct field order of TransformParams / OneFrameParams (main.go:288-306)
    assert list(tj) == ["flat_field", "camera_angle_x", "fl_x", "fl_y", "w", "h", "cx", "cy",
                        "frames"]
    assert tj["flat_field"] == 1.0 and '"flat_field": 1,' in text
    assert tj["camera_angle_x"] == fov * np.pi / 180.0
    fl = 1 / np.tan(fov / 2 * np.pi / 180) * float(res) / 2.0
    assert tj["fl_x"] == fl and tj["fl_y"] == fl
    assert tj["w"] == res and tj["h"] == res and tj["cx"] == res / 2 and tj["cy"] == res / 2
    assert len(tj["frames"]) == len(angles)
    for k, (fr, (a, p)) in enumerate(zip(tj["frames"], angles)):
        assert list(fr) == ["file_path", "time", "transform_matrix"]
        assert fr["file_path"] == f"images/image_{k:03d}.png"
        assert fr["time"] == t
        _, cam = O.camera_from_angles(a, p, R)
        assert np.array_equal(np.array(fr["transform_matrix"]), cam)


def test_volume_raw_matches_oracle(renderer, tmp_path):
    name = "cube_w_hole.yaml"
    obj, numbers, _ = load_scene_file(name)
    res, dm = 28, 1.7
    run(renderer, tmp_path, input=os.path.join(SCENES, name), resolution=res, ds=0.05,
        num_images=1, integration="simple", export_volume=True, density_multiplier=dm)
    got = np.fromfile(tmp_path / "volume.raw", dtype=np.uint8)
    assert got.size == res ** 3
    vol = oracle_scene(obj, numbers=numbers, density_multiplier=dm).export_volume(res)
    want = go_byte(vol / vol.max() * 255)
    assert np.array_equal(got, want)
