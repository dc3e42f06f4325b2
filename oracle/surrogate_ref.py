"""FP64 restatement of the learned density surrogate (test infrastructure).

The surrogate has no counterpart in the reference renderer; this module is
the CPU statement of its mathematics that the GPU kernels are checked
against (forward values and gradients), and it is itself checked by central
finite differences (tests/test_surrogate_cpu.py).

Model
-----
* Input: an object-space point p (after the exact deformation), mapped to
  u = (p - lo) / (hi - lo) in [0, 1]^3; with ``wrap`` (tessellated scenes)
  p is first wrapped into the unit cell the way objects.go:568-582 wraps it,
  and lo/hi are the cell bounds.
* Multiresolution hash encoding (Mueller et al. 2022, "Instant neural
  graphics primitives"): L levels, F = 2 features per level, table size
  T = 2**log2_table per level.  Level l has resolution
  N_l = floor(N_min * b**l), b = exp((ln N_max - ln N_min) / (L - 1)).
  g = u * N_l, corner c = floor(g) + {0,1}^3, weights trilinear in
  g - floor(g) (floor(g) capped at N_l - 1).  Corner index: dense c_x + (N_l+1)(c_y + (N_l+1) c_z) when
  (N_l+1)^3 <= T, else (c_x * 1 ^ c_y * 2654435761 ^ c_z * 805459861)
  mod 2**32 mod T.
* MLP: x (L*F = 32) -> h1 = relu(W1^T x + b1) (64) -> h2 = relu(W2^T h1 +
  b2) (64) -> z = w3 . h2 + b3 -> density = sigmoid(z).
* Loss: mean squared error against the exact density.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

PRIMES = (1, 2654435761, 805459861)


@dataclass
class Config:
    levels: int = 16
    features: int = 2
    log2_table: int = 19
    n_min: int = 16
    n_max: int = 2048
    hidden: int = 64

    def resolutions(self):
        if self.levels == 1:
            return [self.n_min]
        b = math.exp((math.log(self.n_max) - math.log(self.n_min)) / (self.levels - 1))
        return [int(math.floor(self.n_min * b ** l)) for l in range(self.levels)]


def corner_index(c: np.ndarray, n: int, table: int) -> np.ndarray:
    """c: (..., 3) int64 corner coordinates of a level with resolution n."""
    if (n + 1) ** 3 <= table:
        return (c[..., 0] + (n + 1) * (c[..., 1] + (n + 1) * c[..., 2])).astype(np.int64)
    h = (c[..., 0].astype(np.uint64) * np.uint64(PRIMES[0])) & np.uint64(0xFFFFFFFF)
    h ^= (c[..., 1].astype(np.uint64) * np.uint64(PRIMES[1])) & np.uint64(0xFFFFFFFF)
    h ^= (c[..., 2].astype(np.uint64) * np.uint64(PRIMES[2])) & np.uint64(0xFFFFFFFF)
    return (h % np.uint64(table)).astype(np.int64)


def encode(u: np.ndarray, tables: np.ndarray, cfg: Config):
    """u: (n, 3) in [0,1]; tables: (L, T, F).  Returns features (n, L*F) and
    the (index, weight) pairs per level for the backward pass."""
    n = u.shape[0]
    T = 1 << cfg.log2_table
    feats = np.zeros((n, cfg.levels * cfg.features))
    cache = []
    for l, N in enumerate(cfg.resolutions()):
        g = u * N
        base = np.minimum(np.floor(g), N - 1)  # u = 1 stays in the last cell
        w = g - base
        base = base.astype(np.int64)
        idxs, wts = [], []
        acc = np.zeros((n, cfg.features))
        for corner in range(8):
            o = np.array([(corner >> 2) & 1, (corner >> 1) & 1, corner & 1])
            c = base + o
            wc = np.prod(np.where(o == 1, w, 1.0 - w), axis=1)
            idx = corner_index(c, N, T)
            acc += wc[:, None] * tables[l, idx, :]
            idxs.append(idx)
            wts.append(wc)
        feats[:, l * cfg.features:(l + 1) * cfg.features] = acc
        cache.append((np.stack(idxs, 1), np.stack(wts, 1)))
    return feats, cache


def forward(x: np.ndarray, params: dict):
    h1p = x @ params["W1"] + params["b1"]
    h1 = np.maximum(h1p, 0.0)
    h2p = h1 @ params["W2"] + params["b2"]
    h2 = np.maximum(h2p, 0.0)
    z = h2 @ params["w3"] + params["b3"]
    y = 1.0 / (1.0 + np.exp(-z))
    return y, (x, h1p, h1, h2p, h2, z)


def loss_and_grads(u: np.ndarray, target: np.ndarray, params: dict, cfg: Config):
    """MSE loss and its gradients with respect to every parameter."""
    x, cache = encode(u, params["tables"], cfg)
    y, (x, h1p, h1, h2p, h2, z) = forward(x, params)
    n = u.shape[0]
    loss = float(np.mean((y - target) ** 2))
    dy = 2.0 * (y - target) / n
    dz = dy * y * (1.0 - y)
    g = {"b3": np.array(dz.sum()), "w3": h2.T @ dz}
    dh2 = np.outer(dz, params["w3"]) * (h2p > 0)
    g["W2"] = h1.T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ params["W2"].T) * (h1p > 0)
    g["W1"] = x.T @ dh1
    g["b1"] = dh1.sum(0)
    dx = dh1 @ params["W1"].T
    gt = np.zeros_like(params["tables"])
    for l, (idx, wts) in enumerate(cache):
        for f in range(cfg.features):
            contrib = wts * dx[:, l * cfg.features + f][:, None]
            np.add.at(gt[l, :, f], idx.reshape(-1), contrib.reshape(-1))
    g["tables"] = gt
    return loss, g


def predict(u: np.ndarray, params: dict, cfg: Config) -> np.ndarray:
    x, _ = encode(u, params["tables"], cfg)
    return forward(x, params)[0]
