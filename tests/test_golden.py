"""Golden images (tests/golden/images.npz, written by make_golden.py from
the FP64 oracle): the oracle must reproduce them bit for bit, and libxrp
must match them within the parity bar."""
import json
import os

import numpy as np
import pytest

import xray_projection_render_amd as X
from helpers import GOLDEN, SCENES, oracle_render, oracle_scene

TOL = 1e-5


def _load():
    z = np.load(os.path.join(GOLDEN, "images.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return z, meta


CASES = sorted(_load()[1])


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(name):
    z, meta = _load()
    m = meta[name]
    vox = z[name + "__voxels"] if m["voxels"] else None
    sc = oracle_scene(m["object"], m["deformation"], vox, numbers=m["numbers"],
                      flat_field=m["flat_field"], density_multiplier=m["density_multiplier"],
                      base_dir=SCENES)
    img = oracle_render(sc, [tuple(a) for a in m["angles"]], m["res"], m["ds"],
                        integrator=m["integrator"])
    assert np.array_equal(img, z[name])


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_golden(ctx, name):
    z, meta = _load()
    m = meta[name]
    if m["voxels"]:
        sc = X.Scene.from_voxels(z[name + "__voxels"], m["deformation"], numbers=m["numbers"])
    else:
        sc = X.Scene.from_dict(m["object"], m["deformation"], numbers=m["numbers"],
                               base_dir=SCENES)
    img, st = X.render(ctx, sc, [tuple(a) for a in m["angles"]], m["res"], m["ds"],
                       integrator=m["integrator"], flat_field=m["flat_field"],
                       density_multiplier=m["density_multiplier"])
    err = float(np.max(np.abs(img.astype(np.float64) - z[name])))
    print(f"{name}: max_abs={err:.3e}", flush=True)
    assert err <= TOL
