"""The surrogate's FP64 restatement (oracle/surrogate_ref.py) pinned by
central finite differences and by the encoding's defining properties."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import surrogate_ref as S  # noqa: E402


def _params(cfg, rng, scale=0.5):
    T = 1 << cfg.log2_table
    fin = cfg.levels * cfg.features
    return {
        "tables": rng.uniform(-scale, scale, (cfg.levels, T, cfg.features)),
        "W1": rng.normal(0, 1 / math.sqrt(fin), (fin, cfg.hidden)),
        "b1": rng.normal(0, 0.1, cfg.hidden),
        "W2": rng.normal(0, 1 / math.sqrt(cfg.hidden), (cfg.hidden, cfg.hidden)),
        "b2": rng.normal(0, 0.1, cfg.hidden),
        "w3": rng.normal(0, 1 / math.sqrt(cfg.hidden), cfg.hidden),
        "b3": np.array(0.1),
    }


def test_resolutions_and_indexing():
    cfg = S.Config()
    res = cfg.resolutions()
    assert res[0] == 16 and res[-1] == 2048 and len(res) == 16
    assert all(a <= b for a, b in zip(res, res[1:]))
    # dense levels index every corner once
    c = np.stack(np.meshgrid(*[np.arange(17)] * 3, indexing="ij"), -1).reshape(-1, 3)
    idx = S.corner_index(c, 16, 1 << 19)
    assert len(np.unique(idx)) == 17 ** 3 and idx.max() < 17 ** 3
    # hashed levels: the spatial hash of the published method
    h = S.corner_index(np.array([[3, 5, 7]]), 2048, 1 << 19)[0]
    assert h == ((3 * 1) ^ (5 * 2654435761 % 2 ** 32) ^ (7 * 805459861 % 2 ** 32)) % (1 << 19)


def test_encoding_interpolates_affine_tables():
    """A table holding an affine function of the corner coordinates (dense
    level) is reproduced exactly by trilinear interpolation."""
    cfg = S.Config(levels=1, n_min=8, n_max=8, log2_table=12)
    T = 1 << cfg.log2_table
    tables = np.zeros((1, T, 2))
    c = np.stack(np.meshgrid(*[np.arange(9)] * 3, indexing="ij"), -1).reshape(-1, 3)
    idx = S.corner_index(c, 8, T)
    tables[0, idx, 0] = 1.0 + 2 * c[:, 0] - 3 * c[:, 1] + 0.5 * c[:, 2]
    tables[0, idx, 1] = c[:, 2]
    u = np.random.default_rng(1).uniform(0, 1, (50, 3))
    f, _ = S.encode(u, tables, cfg)
    g = u * 8
    assert np.allclose(f[:, 0], 1.0 + 2 * g[:, 0] - 3 * g[:, 1] + 0.5 * g[:, 2], atol=1e-12)
    assert np.allclose(f[:, 1], g[:, 2], atol=1e-12)


@pytest.mark.parametrize("cfg", [
    S.Config(levels=3, log2_table=6, n_min=2, n_max=8, hidden=64),
    S.Config(levels=16, log2_table=8, n_min=4, n_max=64, hidden=64),
])
def test_gradients_match_finite_differences(cfg):
    rng = np.random.default_rng(7)
    params = _params(cfg, rng)
    u = rng.uniform(0.02, 0.98, (24, 3))
    t = (rng.uniform(0, 1, 24) > 0.7).astype(np.float64)
    loss, g = S.loss_and_grads(u, t, params, cfg)
    eps = 1e-6
    checks = [("W1", (1, 3)), ("W1", (5, 60)), ("b1", (7,)), ("W2", (10, 20)),
              ("W2", (63, 0)), ("b2", (33,)), ("w3", (12,)), ("b3", ())]
    # table entries actually touched by the samples
    _, cache = S.encode(u, params["tables"], cfg)
    for l in (0, cfg.levels - 1):
        i = int(cache[l][0][0, 0])
        checks.append(("tables", (l, i, 1)))
    for name, ix in checks:
        p = params[name]
        old = p[ix].copy() if p.shape else p.copy()
        p[ix] = old + eps
        lp, _ = S.loss_and_grads(u, t, params, cfg)
        p[ix] = old - eps
        lm, _ = S.loss_and_grads(u, t, params, cfg)
        p[ix] = old
        fd = (lp - lm) / (2 * eps)
        an = g[name][ix]
        assert abs(fd - an) <= 1e-7 + 1e-5 * abs(fd), (name, ix, fd, an)
