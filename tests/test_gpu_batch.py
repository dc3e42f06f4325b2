"""Batched persistent launch (xrp_batch_*): many scenes of uneven size and
views in one graph-captured launch per (precision, integrator) group.

The batch runs the same per-ray march as the per-scene kernel with the same
plan (precision, guard bands, sample table), so its images must equal the
per-scene renders bit for bit, for every tile shape and on every graph
replay, and stay within the parity bar of the FP64 oracle."""
import numpy as np
import pytest

import xray_projection_render_amd as X
from xray_projection_render_amd import workloads as W
from helpers import oracle_render, oracle_scene, xrp_scene

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _items():
    """(obj, deformation, voxels, angles, res, ds, kwargs) of uneven scenes."""
    out = []
    for k, seed in enumerate([0, 3, 5, 8]):
        obj, _, _ = W.random_scene(seed, max_prims=25)
        ds = max(xrp_scene(obj).auto_ds(), 0.006)
        angles = [(20.0 + 70 * k, 50.0 + 20 * k), (200.0 - 30 * k, 120.0)][: 1 + k % 2]
        out.append((obj, None, None, angles, [33, 48, 40, 24][k], ds, {}))
    lat = W.lattice(resolution=40, n_views=2)
    out.append((lat.obj, lat.deformation, None, lat.angles, 40, 0.004, {}))
    csg = W.csg(resolution=36, n_views=3)
    out.append((csg.obj, None, None, csg.angles, 36, csg.ds,
                {"integrator": "hierarchical", "flat_field": 0.2, "density_multiplier": 1.5}))
    grid = W.blob_grid(24)
    vw = W.voxel(resolution=32, n_views=2, n=24, grid=grid)
    out.append((None, vw.deformation, grid, vw.angles, 32, vw.ds, {}))
    return out


@pytest.fixture(scope="module")
def batch_setup(ctx):
    specs = _items()
    scenes = [xrp_scene(obj, d, v) for obj, d, v, *_ in specs]
    items = [(sc, angles, res, ds, dict({"integrator": "simple"}, **kw))
             for sc, (_, _, _, angles, res, ds, kw) in zip(scenes, specs)]
    b = X.Batch(ctx, items)
    yield specs, scenes, items, b
    b.close()


def test_batch_equals_per_scene_renders(ctx, batch_setup):
    specs, scenes, items, b = batch_setup
    st = b.run()
    assert st["launches"] == 3  # guarded simple, guarded hierarchical, FP64 (voxel) simple
    assert st["rays"] == sum(len(a) * r * r for _, a, r, _, _ in items)
    total = 0
    for k, (sc, angles, res, ds, kw) in enumerate(items):
        got = b.read(k)
        if specs[k][2] is None:
            ref, rst = X.render(ctx, sc, angles, res, ds, **kw)
            assert np.array_equal(got, ref), k
            total += rst["samples"]
        osc = oracle_scene(specs[k][0], specs[k][1], specs[k][2],
                           flat_field=kw.get("flat_field", 0.0),
                           density_multiplier=kw.get("density_multiplier", 1.0))
        oref = oracle_render(osc, angles, res, ds, integrator=kw["integrator"])
        err = float(np.max(np.abs(got.astype(np.float64) - oref)))
        print(f"item {k}: max_abs={err:.3e}", flush=True)
        assert err <= TOL
    assert st["samples"] >= total > 0


def test_batch_shapes_and_replays(ctx, batch_setup):
    """Every tile shape tiles the same pixels; each replay of the captured
    graph (with eager work on the stream between replays) gives the same
    images."""
    specs, scenes, items, b = batch_setup
    b.set_shape(1, 1, 2)
    b.run()
    base = [b.read(k) for k in range(len(items))]
    sc0, a0, r0, ds0, kw0 = items[0]
    for shape in [(2, 2, 2), (2, 1, 4), (1, 4, 1), (4, 1, 3), (1, 2, 8), (8, 1, 2),
                  (4, 4, 1), (16, 1, 2)]:
        b.set_shape(*shape)
        for rep in range(3):
            st = b.run()
            assert st["tiles"] > 0
            for k in range(len(items)):
                assert np.array_equal(b.read(k), base[k]), (shape, rep, k)
            X.render(ctx, sc0, a0, r0, ds0, **kw0)  # eager work between replays


def test_batch_tune_and_latency(ctx, batch_setup):
    specs, scenes, items, b = batch_setup
    shape = b.tune(repeats=1)
    assert 1 <= shape[0] * shape[1] <= 16 and shape[2] >= 1
    st = b.run()
    lat = b.latencies()
    assert np.all(lat > 0)
    assert lat.max() <= st["kernel_ms"] * 1.5 + 1.0
    for k in range(len(items)):
        if specs[k][2] is None:
            sc, angles, res, ds, kw = items[k]
            ref, _ = X.render(ctx, sc, angles, res, ds, **kw)
            assert np.array_equal(b.read(k), ref)
    # a second batch with the same signature starts from the tuned shape
    b2 = X.Batch(ctx, items)
    try:
        b2.run()
        for k in range(len(items)):
            assert np.array_equal(b2.read(k), b.read(k))
    finally:
        b2.close()


def test_batch_rejects_bad_input(ctx):
    sc = xrp_scene(W.csg().obj)
    with pytest.raises(X.XrpError):
        X.Batch(ctx, [(sc, [(0.0, 90.0)], 16, 0.02, {"output": "gray8_alpha8"})])
    with pytest.raises(X.XrpError):
        X.Batch(ctx, [(sc, [(0.0, 90.0)], 0, 0.02)])
    b = X.Batch(ctx, [(sc, [(0.0, 90.0)], 16, 0.02)])
    try:
        with pytest.raises(X.XrpError):
            b.set_shape(6, 3, 1)
        with pytest.raises(X.XrpError):
            b.output_ptr(5)  # the library's own range check
        with pytest.raises(IndexError):
            b.read(5)
    finally:
        b.close()
