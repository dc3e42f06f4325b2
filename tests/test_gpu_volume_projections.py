"""xrp_render_volume_projections, the replacement of RenderVolumeProjectionsCUDA
(cuda_backend.h:101-112), against the FP64 oracle, and the ABI rules that
need a device: output struct sizes, argument limits, the volume cache.

The scenes are the two of the reference's CPU-vs-GPU harness
(cuda_test.go:41-214): a voxelized sphere of radius 0.3 on a 32^3 grid and a
non-cubic 24x32x16 box.  The reference's own GPU kernel is allowed 0.10 /
0.12 max-abs against its CPU path (texture voxel-centre convention); this
library follows the CPU renderer's (N-1) convention and is held to the
plain-form bar of 1e-5.
"""
import ctypes

import numpy as np
import pytest

import xray_projection_render_amd as X
from xray_projection_render_amd import _lib as XL
from helpers import O, oracle_render, oracle_scene, xrp_scene

pytestmark = pytest.mark.gpu

TOL = 1e-5


def sphere_volume(n=32):
    """cuda_test.go:41-72: x = i/N*2-1 (i/N, not i/(N-1)), rho 1 inside r 0.3."""
    c = np.arange(n) / n * 2.0 - 1.0
    z, x, y = np.meshgrid(c, c, c, indexing="ij")
    return ((x * x + y * y + z * z) < 0.09).astype(np.float32)


def box_volume(nx=24, ny=32, nz=16):
    """cuda_test.go:131-146, volume[k][i][j] with i along x, j along y."""
    X_ = np.arange(nx) / nx * 2.0 - 1.0
    Y_ = np.arange(ny) / ny * 2.0 - 1.0
    Z_ = np.arange(nz) / nz * 2.0 - 1.0
    z, x, y = np.meshgrid(Z_, X_, Y_, indexing="ij")
    return ((np.abs(x) < 0.4) & (np.abs(y) < 0.2) & (np.abs(z) < 0.3)).astype(np.float32)


ANGLES = [(0.0, 90.0), (45.0, 60.0), (200.0, 120.0), (300.0, 30.0)]


def _cameras(angles, R=4.0, fov=40.0):
    out = []
    for az, po in angles:
        eye, cam = O.camera_from_angles(az, po, R)
        out.append((eye, cam, fov, R))
    return out


def _oracle(vol, angles, res, ds, R=4.0, fov=40.0, ff=0.0):
    sc = oracle_scene(None, voxels=vol, flat_field=ff)
    return oracle_render(sc, angles, res, ds, R=R, fov=fov)


@pytest.mark.parametrize("name", ["sphere32", "box24x32x16"])
def test_reference_harness_scenes(ctx, name):
    vol = sphere_volume() if name == "sphere32" else box_volume()
    nz, nx, ny = vol.shape
    ds = 2.0 / min(nx, ny, nz) / 5.0  # cuda_test.go:65, 149-159
    res = 32
    got = X.render_volume_projections(ctx, vol, _cameras(ANGLES), res, ds)
    ref = _oracle(vol, ANGLES, res, ds)
    err = float(np.max(np.abs(got.astype(np.float64) - ref)))
    print(f"{name}: max_abs={err:.3e} min T={ref.min():.3f}", flush=True)
    assert err <= TOL
    assert ref.min() < 0.7  # the object attenuates


def test_groups_of_fov_and_R_and_flat_field(ctx):
    """Cameras with different (fov, R) are rendered in separate launches."""
    vol = box_volume()
    ds, res = 2.0 / 16 / 5.0, 40
    cams = _cameras(ANGLES[:2]) + _cameras(ANGLES[2:], R=5.0, fov=30.0) + \
        _cameras(ANGLES[:1], R=4.0, fov=40.0)
    got = X.render_volume_projections(ctx, vol, cams, res, ds, flat_field=0.3)
    ref = np.concatenate([_oracle(vol, ANGLES[:2], res, ds, ff=0.3),
                          _oracle(vol, ANGLES[2:], res, ds, R=5.0, fov=30.0, ff=0.3),
                          _oracle(vol, ANGLES[:1], res, ds, ff=0.3)])
    assert float(np.max(np.abs(got.astype(np.float64) - ref))) <= TOL


def test_volume_cache_follows_contents(ctx):
    """The context keeps the last volume's device copies; a call with the
    same buffer but new values must not reuse them."""
    vol = sphere_volume()
    ds, res = 2.0 / 32 / 5.0, 24
    cams = _cameras(ANGLES[:2])
    a = X.render_volume_projections(ctx, vol, cams, res, ds)
    b = X.render_volume_projections(ctx, vol, cams, res, ds)
    assert np.array_equal(a, b)
    vol *= 0.5
    c = X.render_volume_projections(ctx, vol, cams, res, ds)
    ref = _oracle(vol, ANGLES[:2], res, ds)
    assert float(np.max(np.abs(c.astype(np.float64) - ref))) <= TOL
    assert not np.array_equal(a, c)


def test_volume_argument_errors(ctx):
    vol = sphere_volume(8)
    cams = _cameras(ANGLES[:1])
    with pytest.raises(X.XrpError) as e:
        X.render_volume_projections(ctx, vol, cams, 40000, 0.05)
    assert e.value.code == 2
    with pytest.raises(X.XrpError) as e:
        X.render_volume_projections(ctx, vol, cams, 16, 1e-9)
    assert e.value.code == 12
    bad = [(cams[0][0], np.full((4, 4), np.nan), 40.0, 4.0)]
    with pytest.raises(X.XrpError) as e:
        X.render_volume_projections(ctx, vol, bad, 16, 0.05)
    assert e.value.code == 12
    with pytest.raises(X.XrpError) as e:
        X.render_volume_projections(ctx, vol, cams, 16, -0.05)
    assert e.value.code == 12


def _raw_render(ctx, scene, st):
    p = XL.make_params(16, 0.02)
    out = np.empty((1, 16, 16), dtype=np.float32)
    return XL.lib().xrp_render(ctx.ptr, scene.ptr, ctypes.byref(p), XL._views([(30.0, 80.0)]),
                               1, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st))


def test_stats_struct_size_is_honoured(ctx):
    """The library writes no byte past the caller's struct_size and rejects
    a struct_size that was never set."""
    sc = xrp_scene({"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.3, "rho": 1.0})
    full = XL.RenderStats()
    assert _raw_render(ctx, sc, full) == 0 and full.rays == 256 and full.samples > 0
    part = XL.RenderStats()
    ctypes.memset(ctypes.byref(part), 0xAB, ctypes.sizeof(part))
    part.struct_size = 16  # header + rays only
    assert _raw_render(ctx, sc, part) == 0
    assert part.rays == 256
    raw = bytes(ctypes.string_at(ctypes.addressof(part), ctypes.sizeof(part)))
    assert raw[16:] == b"\xab" * (len(raw) - 16)
    unset = XL.RenderStats()
    unset.struct_size = 0
    assert _raw_render(ctx, sc, unset) == 12


def test_batch_create_destroy_releases_device_memory(ctx):
    """Every device buffer of a batch is released with it."""
    import torch
    sc = xrp_scene({"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.3, "rho": 1.0})
    items = [(sc, [(10.0, 80.0)], 32, 0.02, {"integrator": "simple"})]
    X.Batch(ctx, items).close()
    free0, _ = torch.cuda.mem_get_info(0)
    for _ in range(200):
        b = X.Batch(ctx, items)
        b.run()
        b.close()
    free1, _ = torch.cuda.mem_get_info(0)
    print(f"free before {free0 / 2**20:.1f} MiB after {free1 / 2**20:.1f} MiB", flush=True)
    assert free0 - free1 < 2 * 2**20


def test_export_volume_above_2_to_32_voxels(ctx):
    """res = 1632 has more than 2^32 voxels (res^3 = 4.35e9): the export's
    grid is two-dimensional (slices along y), so it launches and every slice
    is written."""
    sc = xrp_scene({"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.5, "rho": 1.0})
    res = 1632
    vol = X.export_volume(ctx, sc, res)
    c = np.arange(res) / res * 2.0 - 1.0
    for k in (0, res // 3, res // 2, res - 1):
        x = c[:, None]
        y = c[None, :]
        want = np.where(x * x + y * y + c[k] * c[k] < 0.25, 255, 0).astype(np.uint8)
        assert np.array_equal(vol[k], want), k
    del vol
