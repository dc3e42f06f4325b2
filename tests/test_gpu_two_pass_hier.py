"""The two-pass guarded hierarchical march (march.hip march_hier_try_kernel,
march_hier_check_kernel, march_hier_redo_kernel; main.go:159-199).

The first pass marches every ray without the FP64 evaluator.  Samples
inside a guard band, coarse and fine, keep their FP32 values and are
listed; the pixel is stored.  The check pass evaluates the listed samples
in FP64, and a ray with any FP32 value that was not exact is marched again
by the one-pass march, as is a ray with more listed samples than it keeps.
Every pixel must have the one-pass march's bits, coarse and fine sample
counts must be the same, and the images stay within the oracle's 1e-5 bar.
XRP_TWO_PASS=1 takes two passes for the small jobs here; XRP_TWO_PASS=0
takes one (read per render)."""
import numpy as np
import pytest

import xray_projection_render_amd as X
from xray_projection_render_amd import workloads as W
from helpers import load_scene_file, oracle_render, oracle_scene, xrp_scene

pytestmark = pytest.mark.gpu

TOL = 1e-5
ANGLES = [(0.0, 90.0), (33.0, 70.0), (90.0, 90.0), (200.0, 120.0)]


def _cases():
    out = []
    for name, dfm in [("cube_w_hole.yaml", None), ("box_w_pped.yaml", None),
                      ("balls.json", None), ("gyroid_example.yaml", None),
                      ("pillar_array.json", None), ("lattice.yaml", None),
                      ("cube_w_hole.yaml", "deformation_linear.yaml"),
                      ("box_w_pped.yaml", "deformation_sigmoid.yaml")]:
        out.append(pytest.param("file", (name, dfm), id=f"{name}+{dfm}"))
    for name in ("csg", "lattice", "gyroid"):
        out.append(pytest.param("workload", name, id=name))
    for seed in (0, 1, 2, 4, 5):
        out.append(pytest.param("random", seed, id=f"random{seed}"))
    return out


def _setup(kind, arg):
    """(object, deformation, numbers, angles, resolution, ds)."""
    if kind == "file":
        name, dname = arg
        obj, numbers, _ = load_scene_file(name)
        dfm = load_scene_file(dname)[0] if dname else None
        sc = xrp_scene(obj, dfm, numbers=numbers)
        ds = sc.auto_ds()
        sc.close()
        return obj, dfm, numbers, ANGLES, 64, ds
    if kind == "workload":
        w = {"csg": lambda: W.csg(resolution=96, n_views=3),
             "lattice": lambda: W.lattice(resolution=64, n_views=2),
             "gyroid": lambda: W.gyroid_lattice(resolution=64, n_views=2, ds=0.002,
                                                shear=True)}[arg]()
        return w.obj, w.deformation, "yaml", w.angles, w.resolution, w.ds
    obj, _, _ = W.random_scene(arg, max_prims=20)
    sc = xrp_scene(obj)
    ds = sc.auto_ds()
    sc.close()
    return obj, None, "yaml", ANGLES, 64, ds


def _pair(ctx, monkeypatch, sc, angles, res, ds, **env):
    monkeypatch.setenv("XRP_TWO_PASS", "1")
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    a, sa = X.render(ctx, sc, angles, res, ds, integrator="hierarchical", precision="guarded")
    for k in env:
        monkeypatch.delenv(k)
    monkeypatch.setenv("XRP_TWO_PASS", "0")
    b, sb = X.render(ctx, sc, angles, res, ds, integrator="hierarchical", precision="guarded")
    return a, sa, b, sb


def _same(a, b):
    n_diff = int(np.count_nonzero(a != b))
    assert n_diff == 0, n_diff
    return n_diff


@pytest.mark.parametrize("kind,arg", _cases())
def test_hier_two_pass_matches_one_pass_and_oracle(ctx, monkeypatch, kind, arg):
    obj, dfm, numbers, angles, res, ds = _setup(kind, arg)
    sc = xrp_scene(obj, dfm, numbers=numbers)
    a, sa, b, sb = _pair(ctx, monkeypatch, sc, angles, res, ds)
    sc.close()
    n_diff = _same(a, b)
    print(f"{kind} {arg}: block {sa['block']} two_pass {sa['two_pass']} checked rays "
          f"{sa['deferred_rays']} redo_rays {sa['redo_rays']} of {sa['rays']}, reevaluated "
          f"{sa['reevaluated']} / {sb['reevaluated']}, fine {sa['fine_samples']}", flush=True)
    # two passes at 1 or 2 samples per visit; large trees (24 nodes and up,
    # 4 per visit) keep one, and then report no block
    assert sa["two_pass"] == (1 if sa["block"] in (1, 2) else 0) and sb["two_pass"] == 0
    assert sb["redo_rays"] == 0 and sb["deferred_rays"] == 0
    assert sa["samples"] == sb["samples"] and sa["fine_samples"] == sb["fine_samples"]
    # every ray that met a guard-band sample was checked or marched again
    moved = sa["redo_rays"] + sa["deferred_rays"]
    assert not sa["two_pass"] or (sa["reevaluated"] > 0) == (moved > 0)
    if kind != "random":  # (culling per wave: the redo waves hold other rays)
        assert abs(sa["reevaluated"] - sb["reevaluated"]) <= 0.02 * sb["reevaluated"] + 8
    osc = oracle_scene(obj, dfm, numbers=numbers)
    ref = oracle_render(osc, angles, res, ds, integrator="hierarchical")
    assert float(np.max(np.abs(a.astype(np.float64) - ref))) <= TOL


_AFFINE = {"type": "affine",
           "matrix": [[1.2, 0.15, -0.1], [-0.2, 0.9, 0.25], [0.1, -0.15, 1.1]]}


@pytest.mark.parametrize("seed", [0, 2, 4, 6, 14])
def test_hier_two_pass_culled_collections(ctx, monkeypatch, seed):
    """Root collections of 8-20 children, culled per wave, under an affine
    map: the first pass guesses with the wave's cull list, the check pass
    computes the guesses again without one (the padded cull spheres keep
    culled children clear of every guard band), and the images keep the
    one-pass march's bits."""
    obj, _, _ = W.random_scene(seed, max_prims=20)
    assert 8 <= len(obj["objects"]) <= 20
    sc = xrp_scene(obj, _AFFINE)
    ds = sc.auto_ds()
    angles = [(20.0, 80.0), (110.0, 95.0), (250.0, 60.0)]
    a, sa, b, sb = _pair(ctx, monkeypatch, sc, angles, 80, ds)
    sc.close()
    print(f"seed {seed}: checked {sa['deferred_rays']} redo {sa['redo_rays']} of "
          f"{sa['rays']}", flush=True)
    assert sa["two_pass"] == 1
    _same(a, b)
    assert sa["samples"] == sb["samples"] and sa["fine_samples"] == sb["fine_samples"]
    osc = oracle_scene(obj, _AFFINE)
    ref = oracle_render(osc, angles, 80, ds, integrator="hierarchical")
    assert float(np.max(np.abs(a.astype(np.float64) - ref))) <= TOL


@pytest.mark.parametrize("most", ["0", "1"])
def test_hier_two_pass_marches_again_what_does_not_fit(ctx, monkeypatch, most):
    """A ray's slots cut to hold at most 0 or 1 listed samples
    (XRP_DEFER_MAX): rays with more stop in the first pass and are marched
    again."""
    w = W.gyroid_lattice(resolution=64, n_views=2, ds=0.002, shear=True)
    sc = xrp_scene(w.obj, w.deformation)
    a, sa, b, sb = _pair(ctx, monkeypatch, sc, w.angles, w.resolution, w.ds,
                         XRP_DEFER_MAX=most)
    sc.close()
    print(f"at most {most}: deferred_rays {sa['deferred_rays']} redo_rays {sa['redo_rays']}",
          flush=True)
    assert sa["two_pass"] == 1 and sa["redo_rays"] > 0
    assert (sa["deferred_rays"] == 0) == (most == "0")
    _same(a, b)
    assert sa["samples"] == sb["samples"] and sa["fine_samples"] == sb["fine_samples"]


@pytest.mark.parametrize("defer_cap,redo_cap,expect_two_pass", [("3", "100000", 1),
                                                                ("3", "2", 0)])
def test_hier_two_pass_lists_full(ctx, monkeypatch, defer_cap, redo_cap, expect_two_pass):
    """Lists shortened (XRP_DEFER_CAP, XRP_REDO_CAP): rays finding the
    deferral list full are marched again; a full redo list renders the job,
    and the scene's later renders, in one pass."""
    w = W.gyroid_lattice(resolution=64, n_views=2, ds=0.002, shear=True)
    sc = xrp_scene(w.obj, w.deformation)
    a, sa, b, sb = _pair(ctx, monkeypatch, sc, w.angles, w.resolution, w.ds,
                         XRP_DEFER_CAP=defer_cap, XRP_REDO_CAP=redo_cap)
    monkeypatch.setenv("XRP_TWO_PASS", "1")
    _, again = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator="hierarchical")
    sc.close()
    print(f"caps {defer_cap}/{redo_cap}: two_pass {sa['two_pass']} deferred_rays "
          f"{sa['deferred_rays']} redo_rays {sa['redo_rays']}", flush=True)
    assert sa["two_pass"] == expect_two_pass and again["two_pass"] == expect_two_pass
    if expect_two_pass:
        assert sa["deferred_rays"] <= 3 and sa["redo_rays"] > 0
    _same(a, b)
    assert sa["samples"] == sb["samples"] and sa["fine_samples"] == sb["fine_samples"]


def test_hier_two_pass_split_launches(ctx, monkeypatch):
    """A large job in launches of a few views (XRP_LAUNCH_VIEWS)."""
    w = W.csg(resolution=2048, n_views=5)
    sc = xrp_scene(w.obj)
    a, sa, b, sb = _pair(ctx, monkeypatch, sc, w.angles, 2048, w.ds, XRP_LAUNCH_VIEWS="2")
    sc.close()
    print(f"deferred_rays {sa['deferred_rays']} redo_rays {sa['redo_rays']}", flush=True)
    assert sa["two_pass"] == 1
    _same(a, b)
    assert sa["samples"] == sb["samples"] and sa["fine_samples"] == sb["fine_samples"]


@pytest.mark.parametrize("kw", [dict(output="gray8_alpha8", transparency=True),
                                dict(flat_field=0.3, density_multiplier=1.7, R=3.5, fov=50.0)],
                         ids=["gray8_alpha8", "flat_field_dm"])
def test_hier_two_pass_outputs_and_parameters(ctx, monkeypatch, kw):
    """Byte output with transparency, and a flat field, density multiplier,
    camera distance and field of view other than the defaults: the same
    bytes or bits as one pass, and (float output) the oracle's image."""
    w = W.gyroid_lattice(resolution=72, n_views=3, ds=0.002, shear=True)
    sc = xrp_scene(w.obj, w.deformation)
    monkeypatch.setenv("XRP_TWO_PASS", "1")
    a, sa = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator="hierarchical", **kw)
    monkeypatch.setenv("XRP_TWO_PASS", "0")
    b, sb = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator="hierarchical", **kw)
    sc.close()
    print(f"{kw}: checked {sa['deferred_rays']} redo {sa['redo_rays']}", flush=True)
    assert sa["two_pass"] == 1 and sb["two_pass"] == 0
    assert np.array_equal(a, b)
    if "flat_field" in kw:
        osc = oracle_scene(w.obj, w.deformation, density_multiplier=kw["density_multiplier"],
                           flat_field=kw["flat_field"])
        ref = oracle_render(osc, w.angles, w.resolution, w.ds, integrator="hierarchical",
                            R=kw["R"], fov=kw["fov"])
        assert float(np.max(np.abs(a.astype(np.float64) - ref))) <= TOL

