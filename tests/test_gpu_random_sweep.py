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
