"""Constants behind the guarded march's safe radius (eval_block.hpp,
compute_guards in scene.cpp), checked against the mathematics they bound.

The gyroid bound uses |grad g| <= 1.77 for g = sin x cos y + sin y cos z +
sin z cos x: |grad g|^2 peaks at 3 (at the origin), and between the points
of a 721^3 grid of the period it can grow by at most 0.128."""
import math

import numpy as np


def grad_sq(x, y, z):
    gx = np.cos(x) * np.cos(y) - np.sin(x) * np.sin(z)
    gy = np.cos(y) * np.cos(z) - np.sin(y) * np.sin(x)
    gz = np.cos(z) * np.cos(x) - np.sin(z) * np.sin(y)
    return gx * gx + gy * gy + gz * gz


def test_gyroid_gradient_bound_random_points():
    rng = np.random.default_rng(7)
    p = rng.uniform(0.0, 2.0 * math.pi, (3, 2_000_000))
    assert np.sqrt(grad_sq(*p).max()) <= 1.77


def test_gyroid_gradient_peak_and_slack():
    # the peak value the grid argument starts from
    assert abs(grad_sq(0.0, 0.0, 0.0) - 3.0) < 1e-15
    # Hessian entries: diagonal |-sin x cos y - cos x sin z| <= sqrt 2,
    # off-diagonal <= 1, so |H|_F <= sqrt 12; |grad g| <= sqrt 6 (each
    # partial <= sqrt 2); grid spacing 2 pi / 720, half-diagonal sqrt(3)/2
    h = 2.0 * math.pi / 720.0
    growth = 2.0 * math.sqrt(12.0) * math.sqrt(6.0) * h * math.sqrt(3.0) / 2.0
    assert math.sqrt(3.0 + growth) < 1.77


def test_gyroid_gradient_grid_maximum():
    # the grid maximum the bound relies on, on a coarser grid (every 4th
    # point of the 721^3 grid, so the same peak is on it)
    t = np.linspace(0.0, 2.0 * math.pi, 181)
    y, z = np.meshgrid(t, t, indexing="ij")
    m = max(float(grad_sq(x, y, z).max()) for x in t)
    assert abs(m - 3.0) < 1e-12
