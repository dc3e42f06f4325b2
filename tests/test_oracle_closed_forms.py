"""Pins of the FP64 oracle by mathematics rather than by the reference's tests.

* cameras against the closed-form gluLookAt basis;
* the simple integrator against the interval-count identity: for convex
  primitives T = ds * sum_k rho * [t_in < s_k < t_out] with the analytic
  chord (t_in, t_out), using the same accumulated s_k grid;
* trilinear interpolation reproduces affine fields exactly;
* tessellation periodicity, gyroid and deformation closed forms.
A dropped term, a wrong sign or a swapped index in the oracle fails one of
these.
"""
import math

import numpy as np
import pytest

from helpers import O


def lookat_closed_form(az, po, R):
    th, ph = math.radians(az), math.radians(po)
    eye = np.array([R * math.cos(th) * math.sin(ph), R * math.sin(th) * math.sin(ph),
                    R * math.cos(ph)])
    f = -eye / np.linalg.norm(eye)
    s = np.cross(f, [0, 0, 1.0])
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    cam = np.eye(4)
    cam[:3, 0], cam[:3, 1], cam[:3, 2], cam[:3, 3] = s, u, -f, eye
    return eye, cam


@pytest.mark.parametrize("az", [0, 45, 90, 135, 180, 225, 270, 315])
def test_camera_equator_closed_form(az):
    eye, cam = O.camera_from_angles(az, 90.0, 4.0)
    th = math.radians(az)
    # at polar 90: s = (-sin, cos, 0), u = (0, 0, 1), -f = eye / R
    np.testing.assert_allclose(cam[:3, 0], [-math.sin(th), math.cos(th), 0], atol=1e-12)
    np.testing.assert_allclose(cam[:3, 1], [0, 0, 1], atol=1e-12)
    np.testing.assert_allclose(cam[:3, 2], eye / 4.0, atol=1e-12)
    np.testing.assert_allclose(cam[3], [0, 0, 0, 1], atol=1e-12)


@pytest.mark.parametrize("az,po,R", [(45, 60, 5.0), (123.4, 17.5, 3.0), (300, 150, 4.0)])
def test_camera_general_closed_form(az, po, R):
    eye, cam = O.camera_from_angles(az, po, R)
    e2, c2 = lookat_closed_form(az, po, R)
    np.testing.assert_allclose(eye, e2, atol=1e-12)
    np.testing.assert_allclose(cam, c2, atol=1e-12)


def s_grid(smin, smax, ds):
    s, x = [], smin
    while x < smax:
        s.append(x)
        x += ds
    return np.array(s)


def chord_sphere(o, d, c, r):
    oc = o - c
    b = np.dot(oc, d)
    disc = b * b - (np.dot(oc, oc) - r * r)
    if disc <= 0:
        return None
    q = math.sqrt(disc)
    return -b - q, -b + q


def chord_box(o, d, c, h):
    t0, t1 = -math.inf, math.inf
    for a in range(3):
        lo, hi = c[a] - h[a], c[a] + h[a]
        if d[a] == 0:
            if not (lo < o[a] < hi):
                return None
            continue
        u, v = (lo - o[a]) / d[a], (hi - o[a]) / d[a]
        t0, t1 = max(t0, min(u, v)), min(t1, max(u, v))
    return (t0, t1) if t0 < t1 else None


def chord_cylinder(o, d, p0, p1, r):
    v = p1 - p0
    L2 = np.dot(v, v)
    w = o - p0
    # radial: |w + t d - ((w + t d).v / L2) v|^2 = r^2
    dp = d - np.dot(d, v) / L2 * v
    wp = w - np.dot(w, v) / L2 * v
    a, b, c = np.dot(dp, dp), 2 * np.dot(wp, dp), np.dot(wp, wp) - r * r
    disc = b * b - 4 * a * c
    if a == 0 or disc <= 0:
        return None
    t0, t1 = (-b - math.sqrt(disc)) / (2 * a), (-b + math.sqrt(disc)) / (2 * a)
    # caps: 0 <= (w + t d).v / L2 <= 1
    dv, wv = np.dot(d, v) / L2, np.dot(w, v) / L2
    if dv != 0:
        u, z = (0 - wv) / dv, (1 - wv) / dv
        t0, t1 = max(t0, min(u, z)), min(t1, max(u, z))
    elif not (0 <= wv <= 1):
        return None
    return (t0, t1) if t0 < t1 else None


def random_rays(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        th, ph = rng.uniform(0, 2 * math.pi), rng.uniform(0.2, math.pi - 0.2)
        o = 4.0 * np.array([math.cos(th) * math.sin(ph), math.sin(th) * math.sin(ph),
                            math.cos(ph)])
        target = rng.uniform(-0.35, 0.35, 3)
        d = target - o
        out.append((o, d / np.linalg.norm(d)))
    return out


def count_inside(s, chord):
    if chord is None:
        return 0
    return int(np.sum((s > chord[0]) & (s < chord[1])))


@pytest.mark.parametrize("prim", ["sphere", "box", "cylinder"])
def test_interval_count_identity(prim):
    ds, smin, smax = 0.013, 4.0 - 1.74, 4.0 + 1.74
    s = s_grid(smin, smax, ds)
    if prim == "sphere":
        obj = {"type": "sphere", "center": [0.05, -0.02, 0.1], "radius": 0.33, "rho": 0.7}
        chord = lambda o, d: chord_sphere(o, d, np.array(obj["center"]), obj["radius"])
    elif prim == "box":
        obj = {"type": "box", "center": [0.05, -0.02, 0.1], "sides": [0.5, 0.3, 0.7],
               "rho": 0.7}
        chord = lambda o, d: chord_box(o, d, np.array(obj["center"]),
                                       np.array(obj["sides"]) / 2)
    else:
        obj = {"type": "cylinder", "p0": [-0.3, 0.1, -0.2], "p1": [0.25, -0.1, 0.3],
               "radius": 0.12, "rho": 0.7}
        chord = lambda o, d: chord_cylinder(o, d, np.array(obj["p0"]), np.array(obj["p1"]),
                                            obj["radius"])
    sc = O.Scene(O.build_object(obj))
    hits = 0
    for o, d in random_rays(300, 7):
        n = count_inside(s, chord(o, d))
        hits += n > 0
        T = -math.log(sc.integrate_simple(o, d, ds, smin, smax))
        assert abs(T - 0.7 * ds * n) <= 1e-12, (o, d, n, T)
    assert hits > 50


def test_interval_count_csg_collection():
    """cube minus sphere minus cylinder: per sample clamp(sum of member rho)."""
    ds, smin, smax = 0.02, 2.26, 5.74
    s = s_grid(smin, smax, ds)
    obj = {"type": "object_collection", "objects": [
        {"type": "cube", "center": [0.0, 0.0, 0.0], "side": 1.0, "rho": 1.0},
        {"type": "sphere", "center": [0.0, 0.0, 0.0], "radius": 0.3, "rho": -1.0},
        {"type": "cylinder", "p0": [-0.6, 0.0, 0.0], "p1": [0.6, 0.0, 0.0], "radius": 0.15,
         "rho": -1.0}]}
    sc = O.Scene(O.build_object(obj))
    for o, d in random_rays(300, 11):
        cb = chord_box(o, d, np.zeros(3), np.full(3, 0.5))
        sp = chord_sphere(o, d, np.zeros(3), 0.3)
        cy = chord_cylinder(o, d, np.array([-0.6, 0, 0]), np.array([0.6, 0, 0]), 0.15)
        tot = np.zeros_like(s)
        for ch, rho in ((cb, 1.0), (sp, -1.0), (cy, -1.0)):
            if ch is not None:
                tot += rho * ((s > ch[0]) & (s < ch[1]))
        want = ds * np.sum(np.clip(tot, 0.0, 1.0))
        T = -math.log(sc.integrate_simple(o, d, ds, smin, smax))
        assert abs(T - want) <= 1e-12


def test_trilinear_reproduces_affine_fields():
    rng = np.random.default_rng(3)
    nx, ny, nz = 5, 7, 4
    a, b, c, e = rng.uniform(-1, 1, 4)
    # node (k, i, j) at world ((i/(nx-1))*2-1, (j/(ny-1))*2-1, (k/(nz-1))*2-1)
    X = np.arange(nx) / (nx - 1) * 2 - 1
    Y = np.arange(ny) / (ny - 1) * 2 - 1
    Z = np.arange(nz) / (nz - 1) * 2 - 1
    rho = (a * X[None, :, None] + b * Y[None, None, :] + c * Z[:, None, None] + e)
    g = O.voxel_node(rho.reshape(-1), nx, ny, nz)
    for p in rng.uniform(-1, 1, (200, 3)):
        assert abs(g.density(*p) - (a * p[0] + b * p[1] + c * p[2] + e)) <= 1e-12


def test_tessellation_periodicity():
    lat = O.build_object({
        "type": "tessellated_obj_coll",
        "uc": {"objects": {"type": "object_collection", "objects": [
            {"type": "sphere", "center": [0.1, 0.0, 0.05], "radius": 0.2, "rho": 1.0},
            {"type": "cylinder", "p0": [-0.25, -0.25, -0.25], "p1": [0.25, 0.25, 0.25],
             "radius": 0.05, "rho": 0.5}]},
            "xmin": -0.25, "xmax": 0.25, "ymin": -0.25, "ymax": 0.25, "zmin": -0.25,
            "zmax": 0.25},
        "xmin": -1.0, "xmax": 1.0, "ymin": -1.0, "ymax": 1.0, "zmin": -1.0, "zmax": 1.0})
    rng = np.random.default_rng(5)
    for p in rng.uniform(-0.7, 0.45, (400, 3)):
        for ax in range(3):
            q = p.copy()
            q[ax] += 0.5
            assert lat.density(*p) == lat.density(*q)
    assert lat.density(1.2, 0, 0) == 0.0


def test_gyroid_closed_form():
    g = O.build_object({"type": "gyroid", "center": [0.0, 0.0, 0.0], "scale": 0.1,
                        "thickness": 0.3, "rho": 2.0})
    assert g.density(0, 0, 0) == 2.0  # G(0) = 0
    for x in np.linspace(-0.5, 0.5, 101):
        want = 2.0 if abs(math.sin(x / 0.1)) < 0.3 else 0.0  # G(x,0,0) = sin x
        assert g.density(x, 0, 0) == want


def test_deformation_closed_forms():
    sg = O.build_deformation({"type": "sigmoid", "amplitude": 0.4, "center": 0.2,
                              "lengthscale": 0.1, "direction": "y"})
    assert sg.apply(1.0, 0.2, 3.0) == (1.0, 0.2 + 0.2, 3.0)
    x, y, z = sg.apply(0.0, 0.2 + 0.1 * math.log(3.0), 0.0)  # sigma = 3/4
    assert abs(y - (0.2 + 0.1 * math.log(3.0) + 0.3)) <= 1e-12
    st = [0.01, 0.02, 0.03, 0.04, 0.05, 0.06]
    ln = O.build_deformation({"type": "linear", "strains": st})
    M = np.eye(3) + np.array([[st[0], st[5], st[4]], [st[5], st[1], st[3]],
                              [st[4], st[3], st[2]]])
    p = np.array([0.3, -0.7, 1.1])
    np.testing.assert_allclose(ln.apply(*p), M @ p, atol=1e-15)
    rg = O.build_deformation({"type": "rigid", "displacements": [0.1, -0.2, 0.3]})
    np.testing.assert_allclose(rg.apply(*p), p + [0.1, -0.2, 0.3], atol=1e-15)
    comp = O.build_deformation({"type": "composed", "deformations": [
        {"type": "rigid", "displacements": [0.1, -0.2, 0.3]},
        {"type": "linear", "strains": st}]})
    np.testing.assert_allclose(comp.apply(*p), M @ (p + [0.1, -0.2, 0.3]), atol=1e-15)


def test_unit_parallelepiped_is_open_unit_box():
    pp = O.build_object({"type": "parallelepiped", "origin": [0.0, 0.0, 0.0],
                         "v0": [1.0, 0.0, 0.0], "v1": [0.0, 1.0, 0.0], "v2": [0.0, 0.0, 1.0],
                         "rho": 1.0})
    box = O.build_object({"type": "box", "center": [0.5, 0.5, 0.5], "sides": [1.0, 1.0, 1.0],
                          "rho": 1.0})
    rng = np.random.default_rng(9)
    pts = np.concatenate([rng.uniform(-0.2, 1.2, (500, 3)),
                          np.array([[0, 0.5, 0.5], [1, 0.5, 0.5], [0.5, 0.5, 1.0]])])
    for p in pts:
        assert pp.density(*p) == box.density(*p)
    # sheared: point = origin + 0.3 v0 + 0.6 v1 + 0.2 v2 is inside
    sh = O.build_object({"type": "parallelepiped", "origin": [0.1, 0.0, -0.1],
                         "v0": [0.7, 0.0, 0.0], "v1": [0.1, 0.6, 0.0], "v2": [0.1, 0.1, 0.5],
                         "rho": -1.0})
    o, v0, v1, v2 = map(np.array, ([0.1, 0, -0.1], [0.7, 0, 0], [0.1, 0.6, 0], [0.1, 0.1, 0.5]))
    assert sh.density(*(o + 0.3 * v0 + 0.6 * v1 + 0.2 * v2)) == -1.0
    assert sh.density(*(o + 1.1 * v0 + 0.6 * v1 + 0.2 * v2)) == 0.0


def test_tanh_shear_closed_form():
    """Extension deformation z += A tanh(k x): closed form at a few points,
    identity in the other axes, odd in x."""
    d = O.build_deformation({"type": "tanh_shear", "amplitude": 0.1, "rate": 8.0,
                             "direction": "z", "source": "x"})
    for p in [(0.0, 0.2, 0.3), (0.1, -0.4, 0.0), (-0.1, 0.0, -0.2), (0.5, 0.5, 0.5)]:
        q = d.apply(*p)
        assert q[0] == p[0] and q[1] == p[1]
        assert abs(q[2] - (p[2] + 0.1 * math.tanh(8.0 * p[0]))) <= 1e-16
    a = d.apply(0.07, 0.0, 0.0)
    b = d.apply(-0.07, 0.0, 0.0)
    assert a[2] == -b[2]
    with pytest.raises(ValueError):
        O.build_deformation({"type": "tanh_shear", "amplitude": 0.1, "rate": 8.0,
                             "direction": "z", "source": "z"})
