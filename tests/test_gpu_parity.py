"""GPU parity: libxrp against the FP64 oracle on the same inputs.

Bar (stated in the benchmark definition): max-abs transmittance error
<= 1e-5 for the plain form.  The guarded FP32 march is built to reproduce
the oracle's per-sample densities exactly, so the observed error is the
float32 rounding of the stored transmittance (<= 6e-8) and the audit
counter (FP32-clear samples whose FP64 value differs) must be zero.
"""
import math
import os

import numpy as np
import pytest

import xray_projection_render_amd as X
from xray_projection_render_amd import workloads as W
from helpers import load_scene_file, oracle_render, oracle_scene, xrp_scene, SCENES

pytestmark = pytest.mark.gpu

# the child processes below import the package from the repository root
_REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

TOL = 1e-5


def _compare(ctx, obj, angles, res, ds, deformation=None, voxels=None, numbers="yaml",
             integrator="simple", precision="auto", audit=True, ff=0.0, dm=1.0,
             base_dir=None, R=4.0, fov=40.0):
    sc = xrp_scene(obj, deformation, voxels, numbers=numbers, base_dir=base_dir)
    img, st = X.render(ctx, sc, angles, res, ds, integrator=integrator, precision=precision,
                       audit=audit, flat_field=ff, density_multiplier=dm, R=R, fov=fov)
    osc = oracle_scene(obj, deformation, voxels, numbers=numbers, flat_field=ff,
                       density_multiplier=dm, base_dir=base_dir or ".")
    ref = oracle_render(osc, angles, res, ds, integrator=integrator, R=R, fov=fov)
    err = float(np.max(np.abs(img.astype(np.float64) - ref)))
    print(f"max_abs={err:.3e} stats={st}", flush=True)
    return err, st, img, ref


@pytest.mark.parametrize("precision", ["guarded", "fp64"])
def test_csg_full_config(ctx, precision):
    w = W.csg()
    err, st, img, ref = _compare(ctx, w.obj, w.angles, w.resolution, w.ds, precision=precision)
    assert err <= TOL
    assert st["audit_mismatch"] == 0
    # the scene is non-trivial: the cube attenuates, the hole shows through
    assert ref.min() < 0.5 and ref.max() == 1.0


def test_csg_hierarchical(ctx):
    w = W.csg(resolution=128)
    for precision in ("guarded", "fp64"):
        err, st, _, _ = _compare(ctx, w.obj, w.angles, w.resolution, w.ds,
                                 integrator="hierarchical", precision=precision)
        assert err <= TOL
        assert st["audit_mismatch"] == 0
        assert st["fine_samples"] > 0


@pytest.mark.parametrize("name,deformation", [
    ("cube_w_hole.yaml", None),
    ("box_w_pped.yaml", None),
    ("balls.json", None),
    ("gyroid_example.yaml", None),
    ("pillar_array.json", None),
    ("lattice.yaml", None),
    ("cube_w_hole.yaml", "deformation_linear.yaml"),
    ("box_w_pped.yaml", "deformation_sigmoid.yaml"),
])
def test_example_scenes(ctx, name, deformation):
    obj, numbers, _ = load_scene_file(name)
    dfm = load_scene_file(deformation)[0] if deformation else None
    sc = xrp_scene(obj, numbers=numbers)
    ds = sc.auto_ds()
    res = 64
    angles = [(0.0, 90.0), (33.0, 70.0), (90.0, 90.0), (200.0, 120.0)]
    for integrator in ("simple", "hierarchical"):
        err, st, _, _ = _compare(ctx, obj, angles, res, ds, deformation=dfm, numbers=numbers,
                                 integrator=integrator)
        assert err <= TOL, (name, integrator)
        assert st["audit_mismatch"] == 0


@pytest.mark.parametrize("dfm", [
    {"type": "gaussian", "amplitudes": [0.1, -0.05, 0.08], "sigmas": [0.3, 0.2, 0.4],
     "centers": [0.1, 0.0, -0.1]},
    {"type": "rigid", "displacements": [0.1, -0.2, 0.05]},
    {"type": "affine", "matrix": [[1.1, 0.1, 0.0], [0.0, 0.9, 0.05], [0.02, 0.0, 1.2]]},
    {"type": "linear", "strains": [0.05, -0.02, 0.03, 0.01, 0.02, -0.01]},
    {"type": "sigmoid", "amplitude": 0.2, "center": 0.0, "lengthscale": 0.1, "direction": "x"},
    {"type": "composed", "deformations": [
        {"type": "rigid", "displacements": [0.05, 0.0, 0.0]},
        {"type": "sigmoid", "amplitude": -0.1, "center": 0.1, "lengthscale": 0.2,
         "direction": "y"}]},
    {"type": "tanh_shear", "amplitude": 0.1, "rate": 8.0, "direction": "z", "source": "x"},
])
def test_deformations(ctx, dfm):
    w = W.csg(resolution=96, n_views=3)
    err, st, _, _ = _compare(ctx, w.obj, w.angles, w.resolution, w.ds, deformation=dfm)
    assert err <= TOL
    assert st["audit_mismatch"] == 0


def test_lattice_config_reduced(ctx):
    w = W.lattice(resolution=96, n_views=3)
    err, st, _, _ = _compare(ctx, w.obj, w.angles, w.resolution, w.ds,
                             deformation=w.deformation)
    assert err <= TOL
    assert st["audit_mismatch"] == 0


@pytest.mark.parametrize("name", ["csg", "lattice", "gyroid", "random"])
def test_safe_radius_skipping(ctx, name):
    """The guarded simple march settles samples inside an evaluated
    sample's safe radius without evaluating them.  With audit each settled
    sample is also evaluated in FP64 and compared (audit_mismatch counts
    any difference); without audit the image is the same."""
    dfm = None
    if name == "csg":
        w = W.csg(resolution=96, n_views=3)
    elif name == "lattice":
        w = W.lattice(resolution=64, n_views=2)
        dfm = w.deformation
    elif name == "gyroid":
        w = W.gyroid_lattice(resolution=48, n_views=2, ds=0.004, shear=True)
        dfm = w.deformation
    else:
        obj, _, _ = W.random_scene(4, max_prims=80)
        w = W.csg(resolution=48, n_views=2)
        w.obj, w.ds = obj, max(xrp_scene(obj).auto_ds(), 0.004)
    err, st, img, ref = _compare(ctx, w.obj, w.angles, w.resolution, w.ds, deformation=dfm)
    assert err <= TOL
    assert st["audit_mismatch"] == 0
    assert st["precision_used"] == 1  # XRP_PRECISION_GUARDED
    assert 0 < st["skipped_samples"] < st["samples"]
    print(f"{name}: skipped {st['skipped_samples'] / st['samples']:.1%} of samples", flush=True)
    sc = xrp_scene(w.obj, dfm)
    img2, st2 = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator="simple",
                         audit=False)
    assert np.array_equal(img2, img)
    if st2["two_pass"] and name == "random":
        # the second pass groups rays into other waves, whose culled child
        # lists (and so safe radii) differ: same densities, other skip counts
        assert abs(st2["skipped_samples"] - st["skipped_samples"]) <= 1e-3 * st["samples"]
    else:
        assert st2["skipped_samples"] == st["skipped_samples"]


_NO_SKIP_SCRIPT = r"""
import sys, numpy as np
import xray_projection_render_amd as X
from xray_projection_render_amd import workloads as W
ctx = X.Context(0)
out = {}
for name in ("csg", "lattice", "gyroid"):
    if name == "csg":
        w, dfm = W.csg(resolution=64, n_views=2), None
    elif name == "lattice":
        w = W.lattice(resolution=64, n_views=2); dfm = w.deformation
    else:
        w = W.gyroid_lattice(resolution=48, n_views=2, ds=0.004, shear=True); dfm = w.deformation
    sc = X.Scene.from_dict(w.obj, dfm)
    for integ in ("simple", "hierarchical"):
        img, st = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator=integ)
        assert st["skipped_samples"] == 0
        out[name + "_" + integ] = img
    sc.close()
ctx.close()
np.savez(sys.argv[1], **out)
"""


def test_safe_radius_settling_changes_no_bits(ctx, tmp_path):
    """The same renders with settling switched off (XRP_SKIP=0, read once per
    process, hence a child process) give the same bits."""
    import subprocess
    import sys as _sys
    path = tmp_path / "noskip.npz"
    env = dict(os.environ, XRP_SKIP="0")
    subprocess.run([_sys.executable, "-c", _NO_SKIP_SCRIPT, str(path)], env=env, check=True,
                   timeout=300, cwd=_REPO)
    ref = np.load(path)
    for name in ("csg", "lattice", "gyroid"):
        if name == "csg":
            w, dfm = W.csg(resolution=64, n_views=2), None
        elif name == "lattice":
            w = W.lattice(resolution=64, n_views=2)
            dfm = w.deformation
        else:
            w = W.gyroid_lattice(resolution=48, n_views=2, ds=0.004, shear=True)
            dfm = w.deformation
        sc = xrp_scene(w.obj, dfm)
        for integ in ("simple", "hierarchical"):
            img, st = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator=integ)
            assert st["skipped_samples"] > 0, (name, integ)
            assert np.array_equal(img, ref[name + "_" + integ]), (name, integ)


_CULL_DEFORMATION = {"type": "affine",
                     "matrix": [[1.2, 0.15, -0.1], [-0.2, 0.9, 0.25], [0.1, -0.15, 1.1]]}

_NO_CULL_SCRIPT = r"""
import sys, numpy as np
import xray_projection_render_amd as X
from xray_projection_render_amd import workloads as W
ctx = X.Context(0)
dfm = %r
out = {}
for seed, n, d in ((4, 80, None), (1, 200, None), (6, 120, dfm)):
    obj, _, _ = W.random_scene(seed, max_prims=n)
    sc = X.Scene.from_dict(obj, d)
    ds = max(sc.auto_ds(), 0.004)
    for integ in ("simple", "hierarchical"):
        img, st = X.render(ctx, sc, [(17.0, 60.0), (250.0, 130.0)], 40, ds, integrator=integ)
        out["%%d_%%s" %% (seed, integ)] = img
    sc.close()
ctx.close()
np.savez(sys.argv[1], **out)
"""


def test_child_culling_changes_no_bits(ctx, tmp_path):
    """Per-wave culling of the root collection's children (bounding spheres
    against each ray's line in object space, also under an affine map)
    leaves out only children that are zero along the wave's rays: audit
    finds no FP32-clear sample that differs from FP64, and renders with
    culling off (XRP_CULL=0, read once per process, hence a child process)
    give the same bits."""
    import subprocess
    import sys as _sys
    path = tmp_path / "nocull.npz"
    env = dict(os.environ, XRP_CULL="0")
    subprocess.run([_sys.executable, "-c", _NO_CULL_SCRIPT % (_CULL_DEFORMATION,), str(path)],
                   env=env, check=True, timeout=300, cwd=_REPO)
    ref = np.load(path)
    for seed, n, d in ((4, 80, None), (1, 200, None), (6, 120, _CULL_DEFORMATION)):
        obj, _, _ = W.random_scene(seed, max_prims=n)
        sc = xrp_scene(obj, d)
        ds = max(sc.auto_ds(), 0.004)
        for integ in ("simple", "hierarchical"):
            angles = [(17.0, 60.0), (250.0, 130.0)]
            img, st = X.render(ctx, sc, angles, 40, ds, integrator=integ, audit=True)
            assert st["audit_mismatch"] == 0, (seed, integ)
            assert np.array_equal(img, ref["%d_%s" % (seed, integ)]), (seed, integ)
        err, _, _, _ = _compare(ctx, obj, [(17.0, 60.0)], 40, ds, deformation=d)
        assert err <= TOL


@pytest.mark.parametrize("name", ["csg", "lattice"])
def test_safe_radius_skipping_hierarchical(ctx, name):
    """Coarse samples of the hierarchical march inside a safe radius are
    settled without evaluation (no toggle can happen there); audit checks
    each one against FP64."""
    if name == "csg":
        w, dfm = W.csg(resolution=96, n_views=3), None
    else:
        w = W.lattice(resolution=64, n_views=2)
        dfm = w.deformation
    err, st, img, _ = _compare(ctx, w.obj, w.angles, w.resolution, w.ds, deformation=dfm,
                               integrator="hierarchical")
    assert err <= TOL
    assert st["audit_mismatch"] == 0
    assert 0 < st["skipped_samples"] < st["samples"]
    assert st["fine_samples"] > 0
    print(f"{name}: skipped {st['skipped_samples'] / st['samples']:.1%} of coarse samples",
          flush=True)
    sc = xrp_scene(w.obj, dfm)
    img2, _ = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator="hierarchical",
                       audit=False)
    assert np.array_equal(img2, img)


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 5, 8])
def test_batch_scenes_reduced(ctx, seed):
    obj, _, _ = W.random_scene(seed, max_prims=60)
    sc = xrp_scene(obj)
    ds = max(sc.auto_ds(), 0.004)
    angles = [(17.0 + 40 * seed, 60.0), (250.0, 130.0)]
    err, st, _, _ = _compare(ctx, obj, angles, 48, ds)
    assert err <= TOL
    assert st["audit_mismatch"] == 0


@pytest.mark.parametrize("precision", ["fp64", "guarded"])
def test_voxel_reduced(ctx, precision):
    grid = W.blob_grid(48)
    w = W.voxel(resolution=64, n_views=3, n=48, grid=grid)
    err, st, _, _ = _compare(ctx, None, w.angles, w.resolution, w.ds,
                             deformation=w.deformation, voxels=grid, precision=precision,
                             audit=False)
    assert err <= TOL


@pytest.fixture(scope="module")
def blob256():
    return W.blob_grid(256)


VOXEL_LAYOUTS = ["lds", "brick4", "linear", "cached", "linear_cached", "pair", "quad",
                 "quad_slide", "quad_step"]


@pytest.mark.parametrize("layout", VOXEL_LAYOUTS)
def test_voxel_full_config(ctx, blob256, layout, monkeypatch):
    """The 256^3 blob grid at 512^2 through the dedicated voxel kernel,
    8 of the benchmark's 64 views, against the FP64 oracle."""
    monkeypatch.setenv("XRP_VOXEL_LAYOUT", layout)
    w = W.voxel(grid=blob256)
    angles = w.angles[:8]
    err, st, img, ref = _compare(ctx, None, angles, w.resolution, w.ds,
                                 deformation=w.deformation, voxels=blob256, audit=False)
    assert err <= TOL
    assert ref.min() < 0.6  # the blobs absorb


def test_voxel_layouts_bit_identical(ctx, blob256, monkeypatch):
    """Brick order and the per-lane corner cache change where corners are
    read from, not the arithmetic: every gather layout gives the same bits."""
    w = W.voxel(grid=blob256, n_views=6)
    sc = xrp_scene(None, w.deformation, voxels=blob256)
    imgs = {}
    for layout in ["brick4", "linear", "cached", "linear_cached", "pair", "quad", "quad_slide"]:
        monkeypatch.setenv("XRP_VOXEL_LAYOUT", layout)
        imgs[layout], _ = X.render(ctx, sc, w.angles, w.resolution, w.ds,
                                   integrator="simple")
    for layout, img in imgs.items():
        assert np.array_equal(img, imgs["brick4"]), layout


def test_voxel_uneven_grid_all_layouts(ctx, monkeypatch):
    """A 29 x 37 x 22 (z, x, y) grid -- no axis a multiple of the 4^3 brick,
    every axis a different length -- under a general affine map, so rays
    cross cells along x, y and z and reach the clamped upper faces: every
    gather layout matches the oracle and gives the same bits."""
    rng = np.random.default_rng(7)
    nz, nx, ny = 29, 37, 22
    z, x, y = np.meshgrid(np.linspace(-1, 1, nz), np.linspace(-1, 1, nx),
                          np.linspace(-1, 1, ny), indexing="ij")
    grid = (0.5 + 0.4 * np.sin(3.1 * x + 1.7 * y) * np.cos(2.3 * z - 0.9 * x)
            + 0.05 * rng.random((nz, nx, ny))).astype(np.float32)
    deformation = {"type": "affine",
                   "matrix": [[1.7, 0.3, -0.2], [-0.25, 1.9, 0.35], [0.15, -0.3, 1.6]]}
    angles = [(17.0, 63.0), (100.0, 90.0), (250.0, 131.0), (333.0, 45.0)]
    imgs = {}
    for layout in ["brick4", "linear", "cached", "linear_cached", "pair", "quad", "quad_slide",
                   "lds", "quad_step"]:
        monkeypatch.setenv("XRP_VOXEL_LAYOUT", layout)
        err, st, img, ref = _compare(ctx, None, angles, 48, 0.0123, deformation=deformation,
                                     voxels=grid, audit=False)
        assert err <= TOL, layout
        assert ref.min() < 0.9  # the grid absorbs
        imgs[layout] = img
    for layout, img in imgs.items():
        # lds and quad_step round their positions differently (same error size)
        if layout not in ("lds", "quad_step"):
            assert np.array_equal(img, imgs["brick4"]), layout


def test_voxel_quad_step_multi_cell_steps(ctx, monkeypatch):
    """Steps longer than a cell (2.5 to 7 cells per sample on a 64^3 grid):
    the stepping march moves the cell by the weights' floors, several cells
    at once, and still matches the oracle and the gathering layouts."""
    rng = np.random.default_rng(11)
    n = 64
    z, x, y = np.meshgrid(*(np.linspace(-1, 1, n),) * 3, indexing="ij")
    grid = (0.5 + 0.45 * np.sin(4.1 * x - 2.2 * y + 1.3 * z)
            + 0.05 * rng.random((n, n, n))).astype(np.float32)
    angles = [(31.0, 80.0), (200.0, 110.0)]
    for ds in (0.08, 0.2):
        out = {}
        for layout in ("quad_step", "quad_slide"):
            monkeypatch.setenv("XRP_VOXEL_LAYOUT", layout)
            err, _, img, _ = _compare(ctx, None, angles, 40, ds,
                                      deformation=W.VOXEL_DEFORMATION, voxels=grid, audit=False)
            assert err <= TOL, (layout, ds)
            out[layout] = img
        assert float(np.abs(out["quad_step"] - out["quad_slide"]).max()) <= TOL


@pytest.mark.parametrize("layout", ["brick4", "cached", "pair", "quad", "quad_slide",
                                    "quad_step"])
def test_voxel_boundary_ties(ctx, layout, monkeypatch):
    """Axis-aligned views put samples on the grid faces x = +-1 exactly;
    the per-ray FP64 range must count them as the reference does."""
    monkeypatch.setenv("XRP_VOXEL_LAYOUT", layout)
    grid = np.ones((9, 9, 9), dtype=np.float32)
    for ds in (0.025, 0.0125, 0.01):
        err, st, img, ref = _compare(ctx, None, [(0.0, 90.0), (90.0, 90.0), (180.0, 90.0)],
                                     33, ds, deformation=W.VOXEL_DEFORMATION, voxels=grid,
                                     audit=False)
        assert err <= TOL


def test_polar_pole_views_are_blank(ctx):
    """polar 0 gives eye on +z, so LookAt's side vector normalizes a zero
    vector and the camera is NaN: every density test is false in the
    reference, so pixels equal exp(-flat_field).  polar 180 is not
    degenerate in FP64 (sin(pi) != 0) and renders normally."""
    w = W.csg(resolution=32)
    err, st, img, ref = _compare(ctx, w.obj, [(0.0, 0.0), (45.0, 0.0)], 32, w.ds, ff=0.5)
    assert err <= TOL
    assert np.allclose(ref, math.exp(-0.5), rtol=0, atol=1e-15)
    assert st["samples"] == 0  # NaN rays are skipped, not marched
    err, st, img, ref = _compare(ctx, w.obj, [(10.0, 180.0)], 32, w.ds, ff=0.5)
    assert err <= TOL
    assert ref.min() < 0.5


def test_flat_field_and_multiplier(ctx):
    w = W.csg(resolution=64, n_views=2)
    err, st, _, _ = _compare(ctx, w.obj, w.angles, 64, w.ds, ff=0.3, dm=2.5)
    assert err <= TOL


def test_gyroid_lattice_tanh_shear_reduced(ctx):
    """The surrogate workload's exact scene (6^3 gyroid cells under
    z += 0.1 tanh(8 x)) at reduced size, both integrators."""
    w = W.gyroid_lattice(resolution=48, n_views=2, ds=0.004, shear=True)
    for integ in ("simple", "hierarchical"):
        err, st, _, ref = _compare(ctx, w.obj, w.angles, w.resolution, w.ds,
                                   deformation=w.deformation, integrator=integ)
        assert err <= TOL
        assert st["audit_mismatch"] == 0
        assert ref.min() < 0.95


@pytest.mark.parametrize("integrator", ["simple", "hierarchical"])
def test_tile_orientation_changes_no_bits(ctx, monkeypatch, integrator):
    """Rays of a wave as 2 x 32 pixels (XRP_TILE=2x32, read per render)
    instead of 32 x 2: the same per-ray march, so the same bits."""
    cases = [W.csg(resolution=70, n_views=2), W.lattice(resolution=48, n_views=1),
             W.gyroid_lattice(resolution=40, n_views=1, ds=0.002, shear=True)]
    for w in cases:
        sc = xrp_scene(w.obj, w.deformation)
        ds = w.ds if integrator == "simple" else 10 * w.ds
        monkeypatch.setenv("XRP_TILE", "32x2")
        a, _ = X.render(ctx, sc, w.angles, w.resolution, ds, integrator=integrator)
        monkeypatch.setenv("XRP_TILE", "2x32")
        b, _ = X.render(ctx, sc, w.angles, w.resolution, ds, integrator=integrator)
        assert np.array_equal(a, b), w.name
        sc.close()


def test_block_size_changes_no_bits(ctx, monkeypatch):
    """Samples per node visit of the guarded simple march (XRP_BLOCK, read
    per render): 1, 2 and 4 evaluate the same samples in the same order,
    so the images are the same bits; the render reports the block it ran
    (1 for gyroids settled by their second-order reach, 2 for small trees)."""
    cases = [(W.csg(resolution=70, n_views=2), 2), (W.lattice(resolution=48, n_views=1), 2),
             (W.gyroid_lattice(resolution=40, n_views=1, ds=0.002, shear=True), 1)]
    for w, default in cases:
        sc = xrp_scene(w.obj, w.deformation)
        monkeypatch.delenv("XRP_BLOCK", raising=False)
        ref, st = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator="simple")
        assert st["block"] == default, w.name
        for b in (1, 2, 4):
            monkeypatch.setenv("XRP_BLOCK", str(b))
            img, st = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator="simple")
            assert st["block"] == b
            assert np.array_equal(img, ref), (w.name, b)
        _, st = X.render(ctx, sc, w.angles, w.resolution, 10 * w.ds,
                         integrator="hierarchical")
        assert st["block"] == 0
        sc.close()
