"""xrp_render_device_noalloc: the same images as xrp_render_device, from a
fixed per-GPU scratch region instead of library allocations.

Checked bit for bit against xrp_render_device (itself held to the oracle by
the parity tests) on two calls in a row, the second smaller than the first,
for an analytic scene under both integrators and a voxel scene; plus the
scratch size limit and the resident-grid rule for voxel scenes."""
import numpy as np
import pytest

import xray_projection_render_amd as X
from xray_projection_render_amd import workloads as W
from helpers import xrp_scene

pytestmark = pytest.mark.gpu


def _pair(ctx, sc, angles, res, ds, **kw):
    import torch
    a = torch.full((len(angles), res, res), -1.0, dtype=torch.float32, device="cuda:0")
    b = torch.full((len(angles), res, res), -2.0, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    sa = X.render_device(ctx, sc, angles, res, ds, a.data_ptr(), **kw)
    sb = X.render_device_noalloc(ctx, sc, angles, res, ds, b.data_ptr(), **kw)
    return a.cpu().numpy(), b.cpu().numpy(), sa, sb


@pytest.mark.parametrize("integrator", ["simple", "hierarchical"])
def test_noalloc_equals_render_device_csg(ctx, integrator):
    w = W.csg(resolution=96, n_views=6)
    sc = xrp_scene(w.obj)
    for angles, res in ((w.angles, 96), (w.angles[:2], 40)):
        print(f"csg {integrator} views={len(angles)} res={res}", flush=True)
        a, b, sa, sb = _pair(ctx, sc, angles, res, w.ds, integrator=integrator)
        assert np.array_equal(a, b)
        assert a.min() < 0.6 and a.max() == 1.0
        assert sa["samples"] == sb["samples"] and sa["rays"] == sb["rays"]
    sc.close()


def test_noalloc_equals_render_device_lattice(ctx):
    w = W.lattice(resolution=64, n_views=3)
    sc = xrp_scene(w.obj, w.deformation)
    for angles, res in ((w.angles, 64), (w.angles[:1], 32)):
        print(f"lattice views={len(angles)} res={res}", flush=True)
        a, b, _, _ = _pair(ctx, sc, angles, res, 0.004)
        assert np.array_equal(a, b)
    sc.close()


def test_noalloc_voxel_scene_needs_resident_grid(ctx):
    import torch
    grid = W.blob_grid(32)
    vw = W.voxel(resolution=48, n_views=4, n=32, grid=grid)
    sc = xrp_scene(None, vw.deformation, grid)
    out = torch.empty((4, 48, 48), dtype=torch.float32, device="cuda:0")
    with pytest.raises(X.XrpError) as e:
        X.render_device_noalloc(ctx, sc, vw.angles, 48, vw.ds, out.data_ptr())
    assert e.value.code == 13
    for angles, res in ((vw.angles, 48), (vw.angles[:2], 24)):
        print(f"voxel views={len(angles)} res={res}", flush=True)
        a, b, _, _ = _pair(ctx, sc, angles, res, vw.ds)
        assert np.array_equal(a, b)
    sc.close()


def test_noalloc_rejects_requests_beyond_the_scratch_region(ctx):
    import torch
    sc = xrp_scene({"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.3, "rho": 1.0})
    angles = [(float(k), 90.0) for k in range(14000)]  # 14000 x 160 B of views > 2 MiB
    out = torch.empty((len(angles), 2, 2), dtype=torch.float32, device="cuda:0")
    with pytest.raises(X.XrpError) as e:
        X.render_device_noalloc(ctx, sc, angles, 2, 0.05, out.data_ptr())
    assert e.value.code == 2
    # a sample table above the region: ds = 1e-5 gives 348,000 samples
    small = torch.empty((1, 2, 2), dtype=torch.float32, device="cuda:0")
    with pytest.raises(X.XrpError) as e:
        X.render_device_noalloc(ctx, sc, angles[:1], 2, 1e-5, small.data_ptr())
    assert e.value.code == 2
    sc.close()
