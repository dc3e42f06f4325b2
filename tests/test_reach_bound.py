"""Soundness of the gyroid's second-order reach (eval_block.hpp gyroid_reach,
constants from scene.cpp compute_guards), checked in FP64 on the CPU.

Along a ray through the sheared gyroid lattice, g(s) = G(D(o + s d) /
scale) with D the deformation chain.  The march settles the samples within
the reach of an evaluated one: the first root of c + (sigma g' - e1) s -
M s^2 / 2, M = sqrt(12) (L / scale)^2 + 1.77 K / scale.  Here the reach is
recomputed in FP64 (without the FP32 margins) at random points and
directions and g is sampled densely along it: it must never reach the
level |g| = th, and it must go much further than the isotropic radius
(|g| - th) scale / (1.77 L) on average, or it would not be worth its cost.
"""
import math

import numpy as np

HESS = 2.25  # bound on the spectral norm of G's Hessian (pinned below)


def gyroid(q):
    x, y, z = q
    return math.sin(x) * math.cos(y) + math.sin(y) * math.cos(z) + math.sin(z) * math.cos(x)


def grad(q):
    x, y, z = q
    return np.array([math.cos(x) * math.cos(y) - math.sin(z) * math.sin(x),
                     math.cos(y) * math.cos(z) - math.sin(x) * math.sin(y),
                     math.cos(z) * math.cos(x) - math.sin(y) * math.sin(z)])


def quad_reach(c, b, M):
    return 2 * c / (math.sqrt(b * b + 2 * M * c) - b) if c > 0 else 0.0


def reach(g0, g1, th, M):
    if abs(g0) >= th:
        return quad_reach(abs(g0) - th, math.copysign(1.0, g0) * g1, M)
    return min(quad_reach(th - g0, -g1, M), quad_reach(g0 + th, g1, M))


def test_reach_never_crosses_the_level_under_tanh_shear():
    A, k = 0.1, 8.0                       # z += A tanh(k x)
    scale, th = (1.0 / 6.0) / (2 * math.pi), 0.1
    a = A * k
    L = (a + math.sqrt(a * a + 4)) / 2     # spectral norm of the shear I + a e_z e_x^T
    K = A * k * k * 0.7699                # max |tanh''| = 4 / (3 sqrt 3)
    inv = 1.0 / scale
    M = HESS * (L * inv) ** 2 + 1.77 * K * inv

    def D(p):
        return np.array([p[0], p[1], p[2] + A * math.tanh(k * p[0])])

    def J(p, v):
        return np.array([v[0], v[1], v[2] + A * k * (1 - math.tanh(k * p[0]) ** 2) * v[0]])

    rng = np.random.default_rng(11)
    gain, n = [], 0
    while n < 3000:
        p = rng.uniform(-0.5, 0.5, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        q = D(p) * inv
        g0 = gyroid(q)
        if abs(abs(g0) - th) < 1e-3:
            continue
        g1 = float(grad(q) @ J(p, d)) * inv
        r = reach(g0, g1, th, M)
        iso = (abs(abs(g0) - th)) * scale / (1.77 * L)
        assert r >= 0.0
        side = abs(g0) < th
        for t in np.linspace(0.0, r, 200)[1:]:
            gt = gyroid(D(p + t * d) * inv)
            assert (abs(gt) < th) == side, (p, d, r, t, g0, gt)
        gain.append(r / iso)
        n += 1
    gain = np.array(gain)
    print(f"reach / isotropic radius: median {np.median(gain):.2f}, mean {gain.mean():.2f}")
    assert np.median(gain) > 1.0 and gain.mean() > 1.5


def test_curvature_constant_bounds_sampled_second_derivatives():
    """|g''| along random lines never exceeds M (finite differences)."""
    A, k = 0.1, 8.0
    scale = (1.0 / 6.0) / (2 * math.pi)
    inv = 1.0 / scale
    a = A * k
    L = (a + math.sqrt(a * a + 4)) / 2
    M = HESS * (L * inv) ** 2 + 1.77 * A * k * k * 0.7699 * inv
    rng = np.random.default_rng(5)
    h = 1e-4
    worst = 0.0
    for _ in range(20000):
        p = rng.uniform(-0.5, 0.5, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        f = [gyroid(np.array([x[0], x[1], x[2] + A * math.tanh(k * x[0])]) * inv)
             for x in (p - h * d, p, p + h * d)]
        worst = max(worst, abs(f[0] - 2 * f[1] + f[2]) / (h * h))
    print(f"max sampled |g''| {worst:.0f} vs M {M:.0f}")
    assert worst <= M


def test_shear_lipschitz_is_the_spectral_norm():
    """|J v| <= (a + sqrt(a^2 + 4)) / 2 |v| for J = I + b e_z e_x^T, |b| <= a."""
    a = 0.8
    bound = (a + math.sqrt(a * a + 4)) / 2
    worst = 0.0
    for b in np.linspace(-a, a, 41):
        J = np.eye(3)
        J[2, 0] = b
        worst = max(worst, np.linalg.norm(J, 2))
    assert worst <= bound * (1 + 1e-12)
    assert abs(worst - bound) < 1e-9


def test_gyroid_hessian_spectral_bound():
    """Largest |eigenvalue| of G's Hessian on a 241^3 grid of the period,
    plus the most it can grow between grid points (entries' gradients at
    most 2 per component: Frobenius rate sqrt(27 * 4)), stays below 2.25."""
    n = 241
    t = np.linspace(0.0, 2 * math.pi, n, endpoint=False)
    worst = 0.0
    for z in t:
        X, Y = np.meshgrid(t, t, indexing="ij")
        sx, cx, sy, cy = np.sin(X), np.cos(X), np.sin(Y), np.cos(Y)
        sz, cz = math.sin(z), math.cos(z)
        hxx = -sx * cy - sz * cx
        hyy = -sx * cy - sy * cz
        hzz = -sy * cz - sz * cx
        hxy = -cx * sy
        hxz = -cz * sx
        hyz = -cy * sz
        H = np.stack([np.stack([hxx, hxy, hxz], -1), np.stack([hxy, hyy, hyz], -1),
                      np.stack([hxz, hyz, hzz], -1)], -2).reshape(-1, 3, 3)
        worst = max(worst, float(np.abs(np.linalg.eigvalsh(H)).max()))
    slack = math.sqrt(27 * 4.0) * (2 * math.pi / n) * math.sqrt(3) / 2
    print(f"grid max {worst:.5f} + slack {slack:.4f}")
    assert worst + slack <= HESS
