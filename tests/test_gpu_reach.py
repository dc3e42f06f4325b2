"""The gyroid's second-order reach (eval_block.hpp gyroid_reach): samples
settled along the ray from the value and slope of g at an evaluated sample
and a bound on its curvature, under every deformation kind.

Settling changes no term of the sum: with audit every settled sample is also
evaluated in FP64 and compared (audit_mismatch must stay 0), the renders
must equal the ones made with the reach switched off (XRP_REACH=0, read once
per process, hence a child process) bit for bit, and stay at the parity bar
of the FP64 oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

import xray_projection_render_amd as X
from xray_projection_render_amd import workloads as W
from helpers import oracle_render, oracle_scene, xrp_scene

pytestmark = pytest.mark.gpu

TOL = 1e-5

DEFORMATIONS = {
    "tanh": W.TANH_SHEAR,
    "sigmoid": {"type": "sigmoid", "amplitude": 0.05, "center": 0.1, "lengthscale": 0.08,
                "direction": "y"},
    "gaussian": {"type": "gaussian", "amplitudes": [0.03, -0.02, 0.04],
                 "sigmas": [0.2, 0.3, 0.25], "centers": [0.1, 0.0, -0.1]},
    "affine": {"type": "affine", "matrix": [[1.05, 0.02, 0.0], [0.0, 0.97, 0.03],
                                            [0.01, 0.0, 1.02]]},
    "composed": {"type": "composed", "deformations": [
        {"type": "rigid", "displacements": [0.01, -0.02, 0.0]}, W.TANH_SHEAR]},
}

CASES = [(name, integ) for name in DEFORMATIONS for integ in ("simple", "hierarchical")]


def _case(name):
    w = W.gyroid_lattice(resolution=80, n_views=2, ds=0.001, shear=False)
    w.angles = [(20.0, 90.0), (133.0, 70.0)]
    return w, DEFORMATIONS[name]


_SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, %r)
sys.path.insert(0, %r)
import torch  # noqa: F401  (the runtime the tests use)
import xray_projection_render_amd as X
from test_gpu_reach import _case, CASES
ctx = X.Context(0)
out = {}
for name, integ in CASES:
    w, d = _case(name)
    sc = X.Scene.from_dict(w.obj, d)
    ds = w.ds if integ == "simple" else 10 * w.ds
    img, st = X.render(ctx, sc, w.angles, w.resolution, ds, integrator=integ)
    out[name + "_" + integ] = img
    out[name + "_" + integ + "_skipped"] = np.array([st["skipped_samples"]])
    sc.close()
ctx.close()
np.savez(sys.argv[1], **out)
"""


def _child(tmp_path_factory, **switches):
    """The CASES renders made in a child process with the given switches
    (each read once per process by the library)."""
    here = os.path.dirname(os.path.abspath(__file__))
    path = tmp_path_factory.mktemp("reach") / "renders.npz"
    env = dict(os.environ, **switches)
    subprocess.run([sys.executable, "-c", _SCRIPT % (os.path.dirname(here), here), str(path)],
                   env=env, check=True, timeout=300)
    return np.load(path)


@pytest.fixture(scope="module")
def no_reach(tmp_path_factory):
    return _child(tmp_path_factory, XRP_REACH="0")


@pytest.mark.parametrize("switches", [{"XRP_SEAMLESS": "0"}, {"XRP_QUEUE": "1"},
                                      {"XRP_QUEUE": "0"}])
def test_seamless_cells_and_queue_change_no_bits(ctx, tmp_path_factory, switches):
    """Seamless tessellations (no seam guard, the radius crossing cell faces)
    and the per-wave queue of FP64 re-evaluations change which samples are
    evaluated, not the sum: the images equal those made without them."""
    other = _child(tmp_path_factory, **switches)
    for name, integ in CASES:
        w, d = _case(name)
        sc = xrp_scene(w.obj, d)
        ds = w.ds if integ == "simple" else 10 * w.ds
        img, st = X.render(ctx, sc, w.angles, w.resolution, ds, integrator=integ)
        assert np.array_equal(img, other[name + "_" + integ]), (switches, name, integ)
        sc.close()


@pytest.mark.parametrize("name,integ", CASES)
def test_reach_settles_exactly(ctx, no_reach, name, integ):
    w, d = _case(name)
    sc = xrp_scene(w.obj, d)
    ds = w.ds if integ == "simple" else 10 * w.ds
    img, st = X.render(ctx, sc, w.angles, w.resolution, ds, integrator=integ, audit=True)
    print(f"{name} {integ}: skipped {st['skipped_samples']} of {st['samples']} "
          f"(without reach {int(no_reach[name + '_' + integ + '_skipped'][0])}), "
          f"reevaluated {st['reevaluated']}", flush=True)
    assert st["audit_mismatch"] == 0
    img2, st2 = X.render(ctx, sc, w.angles, w.resolution, ds, integrator=integ)
    assert np.array_equal(img, img2)
    assert np.array_equal(img2, no_reach[name + "_" + integ])
    ref = oracle_render(oracle_scene(w.obj, d, numbers="json"), w.angles, w.resolution, ds,
                        integrator=integ)
    assert float(np.max(np.abs(img2.astype(np.float64) - ref))) <= TOL
    sc.close()


def test_reach_settles_more_on_the_sheared_lattice(ctx, no_reach):
    """On the metric's scene the reach must pay: markedly more samples
    settled than with the isotropic radius alone."""
    key = "tanh_simple"
    w, d = _case("tanh")
    sc = xrp_scene(w.obj, d)
    _, st = X.render(ctx, sc, w.angles, w.resolution, w.ds, integrator="simple")
    before = int(no_reach[key + "_skipped"][0])
    print(f"settled {st['skipped_samples']} vs {before} of {st['samples']}", flush=True)
    assert st["skipped_samples"] > 1.05 * before
    sc.close()


def _seamed_cases():
    """Tessellations whose cell content is not continuous across the cell
    faces: gyroids of 1.5 and 0.7 periods per cell, and a sphere set off the
    cell's centre so that the faces cut it."""
    out = []
    for periods in (1.5, 0.7):
        w = W.gyroid_lattice(resolution=64, n_views=2, ds=0.002, shear=False)
        g = w.obj["uc"]["objects"]["objects"][0]
        g["scale"] = g["scale"] / periods
        out.append((f"gyroid{periods}", w.obj))
    cell = 0.25
    out.append(("sphere_cut", {
        "type": "tessellated_obj_coll",
        "uc": {"objects": {"type": "object_collection", "objects": [
            {"type": "sphere", "center": [0.06, -0.04, 0.02], "radius": 0.1, "rho": 1.0}]},
               "xmin": -cell / 2, "xmax": cell / 2, "ymin": -cell / 2, "ymax": cell / 2,
               "zmin": -cell / 2, "zmax": cell / 2},
        "xmin": -0.5, "xmax": 0.5, "ymin": -0.5, "ymax": 0.5, "zmin": -0.5, "zmax": 0.5}))
    return out


@pytest.mark.parametrize("integ", ["simple", "hierarchical"])
@pytest.mark.parametrize("dname", ["none", "tanh"])
def test_seamed_tessellations_settle_exactly(ctx, integ, dname):
    """Where the cell content jumps at the faces, a safe radius may cross a
    face only when the neighbouring cell's density at the sample is the
    sample's own: with audit every settled sample is compared in FP64."""
    angles = [(20.0, 90.0), (133.0, 70.0)]
    d = None if dname == "none" else W.TANH_SHEAR
    for name, obj in _seamed_cases():
        sc = xrp_scene(obj, d)
        ds = 0.002 if integ == "simple" else 0.02
        img, st = X.render(ctx, sc, angles, 64, ds, integrator=integ, audit=True)
        print(f"{name} {dname} {integ}: skipped {st['skipped_samples']} of {st['samples']}, "
              f"mismatches {st['audit_mismatch']}", flush=True)
        assert st["audit_mismatch"] == 0, name
        ref = oracle_render(oracle_scene(obj, d, numbers="json"), angles, 64, ds,
                            integrator=integ)
        assert float(np.max(np.abs(img.astype(np.float64) - ref))) <= TOL, name
        sc.close()
