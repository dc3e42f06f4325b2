"""GPU parity over seeded random scenes: collections of up to 200 mixed
primitives (spheres, boxes, parallelepipeds, cylinders, gyroids) under
affine, strain and composed rigid + strain maps, random views, both
integrators.  Every render runs with
audit (each FP32-decided sample also evaluated in FP64; any difference is
counted) and is compared with the FP64 oracle at the 1e-5 bar.  These
scenes take the per-wave child culling, the queued FP64 re-evaluation of
large collections and the hardware sine/cosine of the gyroid evaluator."""
import numpy as np
import pytest

import xray_projection_render_amd as X
from xray_projection_render_amd import workloads as W
from helpers import oracle_render, oracle_scene, xrp_scene

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _affine(seed: int):
    rng = np.random.default_rng(1000 + seed)
    m = np.eye(3) * rng.uniform(0.8, 1.3) + rng.uniform(-0.2, 0.2, (3, 3))
    return {"type": "affine", "matrix": m.tolist()}


def _deformation(seed: int):
    """Seeds cycle through no map, an affine map, a symmetric strain and a
    composed rigid + strain chain: every chain the culling composes into
    one affine map."""
    k = seed % 4
    if k == 1:
        return _affine(seed)
    if k == 2:
        return {"type": "linear", "strains": [0.08, -0.05, 0.1, 0.04, -0.06, 0.03]}
    if k == 3:
        return {"type": "composed", "deformations": [
            {"type": "rigid", "displacements": [0.05, -0.1, 0.02]},
            {"type": "linear", "strains": [-0.1, 0.06, 0.0, 0.05, 0.02, -0.04]}]}
    return None


@pytest.mark.parametrize("seed", [11, 12, 13, 14, 15, 16])
def test_random_scene_sweep(ctx, seed):
    obj, _, _ = W.random_scene(seed, max_prims=200)
    dfm = _deformation(seed)
    sc = xrp_scene(obj, dfm)
    ds = max(sc.auto_ds(), 0.004)
    rng = np.random.default_rng(seed)
    angles = [(float(rng.uniform(0, 360)), float(rng.uniform(20, 160))) for _ in range(2)]
    osc = oracle_scene(obj, dfm)
    for integ in ("simple", "hierarchical"):
        img, st = X.render(ctx, sc, angles, 24, ds, integrator=integ, audit=True)
        ref = oracle_render(osc, angles, 24, ds, integrator=integ)
        err = float(np.max(np.abs(img.astype(np.float64) - ref)))
        print(f"seed {seed} {integ}: {len(obj['objects'])} children, max_abs {err:.2e}, "
              f"audit {st['audit_mismatch']}", flush=True)
        assert st["audit_mismatch"] == 0, (seed, integ)
        assert err <= TOL, (seed, integ)
    sc.close()


_NONLINEAR = [
    None,
    {"type": "tanh_shear", "amplitude": 0.1, "rate": 8.0, "direction": "z", "source": "x"},
    {"type": "sigmoid", "amplitude": 0.05, "center": 0.1, "lengthscale": 0.08,
     "direction": "y"},
    {"type": "gaussian", "amplitudes": [0.03, -0.02, 0.04], "sigmas": [0.2, 0.3, 0.25],
     "centers": [0.1, 0.0, -0.1]},
    {"type": "composed", "deformations": [
        {"type": "rigid", "displacements": [0.01, -0.02, 0.0]},
        {"type": "tanh_shear", "amplitude": 0.05, "rate": 6.0, "direction": "x",
         "source": "y"}]},
]


def _random_tessellation(seed: int):
    """A tessellation of a random unit cell: 1-5 primitives (spheres, boxes,
    cylinders, gyroids of random period) placed so that the cell faces may
    cut them, in cells of random size and offset, under one of the
    nonlinear deformations."""
    rng = np.random.default_rng(500 + seed)
    cell = float(rng.uniform(0.12, 0.3))
    off = rng.uniform(-0.3, 0.3, 3) * cell
    lo = -cell / 2 + off
    objs = []
    for _ in range(int(rng.integers(1, 6))):
        kind = ["sphere", "box", "cylinder", "gyroid"][int(rng.integers(0, 4))]
        c = (lo + rng.uniform(0.0, 1.0, 3) * cell).tolist()
        size = float(rng.uniform(0.15, 0.5)) * cell
        rho = float(rng.choice([1.0, 0.5, -0.5]))
        if kind == "sphere":
            objs.append({"type": "sphere", "center": c, "radius": size, "rho": rho})
        elif kind == "box":
            objs.append({"type": "box", "center": c,
                         "sides": [size, float(rng.uniform(0.2, 0.8)) * cell, size], "rho": rho})
        elif kind == "cylinder":
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            objs.append({"type": "cylinder", "p0": (np.array(c) - size * u).tolist(),
                         "p1": (np.array(c) + size * u).tolist(), "radius": 0.3 * size,
                         "rho": rho})
        else:
            periods = float(rng.choice([1.0, 1.5, 0.7, 2.0]))
            objs.append({"type": "gyroid", "center": c,
                         "scale": cell / (2.0 * np.pi * periods),
                         "thickness": float(rng.uniform(0.05, 0.2)), "rho": rho})
    obj = {"type": "tessellated_obj_coll",
           "uc": {"objects": {"type": "object_collection", "objects": objs},
                  "xmin": float(lo[0]), "xmax": float(lo[0] + cell),
                  "ymin": float(lo[1]), "ymax": float(lo[1] + cell),
                  "zmin": float(lo[2]), "zmax": float(lo[2] + cell)},
           "xmin": -0.45, "xmax": 0.45, "ymin": -0.45, "ymax": 0.4, "zmin": -0.4, "zmax": 0.45}
    return obj, _NONLINEAR[seed % len(_NONLINEAR)]


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 6, 7, 8, 9, 10])  # 5: negative content only
def test_random_tessellation_sweep(ctx, monkeypatch, seed):
    """Random tessellations under nonlinear deformations: every settled or
    FP32-decided sample audited in FP64 (both integrators), against the
    FP64 oracle at 1e-5; and the two-pass marches without audit, whose
    images must have the audited one-pass march's bits (for the hierarchical
    one the first pass guesses guard-band samples near the cells' faces)."""
    obj, dfm = _random_tessellation(seed)
    sc = xrp_scene(obj, dfm)
    ds = 0.003
    rng = np.random.default_rng(seed)
    angles = [(float(rng.uniform(0, 360)), float(rng.uniform(20, 160))) for _ in range(2)]
    osc = oracle_scene(obj, dfm)
    for integ in ("simple", "hierarchical"):
        d = ds if integ == "simple" else 10 * ds
        img, st = X.render(ctx, sc, angles, 40, d, integrator=integ, audit=True)
        ref = oracle_render(osc, angles, 40, d, integrator=integ)
        err = float(np.max(np.abs(img.astype(np.float64) - ref)))
        print(f"seed {seed} {integ}: {len(obj['uc']['objects']['objects'])} children, "
              f"max_abs {err:.2e}, audit {st['audit_mismatch']}, skipped "
              f"{st['skipped_samples']} of {st['samples']}", flush=True)
        assert st["audit_mismatch"] == 0, (seed, integ)
        assert err <= TOL, (seed, integ)
        monkeypatch.setenv("XRP_TWO_PASS", "1")
        img2, st2 = X.render(ctx, sc, angles, 40, d, integrator=integ)
        monkeypatch.delenv("XRP_TWO_PASS")
        print(f"  two passes: checked {st2['deferred_rays']} redo {st2['redo_rays']}",
              flush=True)
        assert st2["two_pass"] == 1
        assert np.array_equal(img2, img), (seed, integ)
    sc.close()
