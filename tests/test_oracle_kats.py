"""The FP64 oracle against the reference's own known-answer tests.

Each test restates one reference test (file:line) with the reference's
inputs and tolerances; together they pin the oracle's integrators, objects
and deformations before it is used to check the HIP kernels.
"""
import math

import numpy as np
import pytest

from helpers import O

E1 = math.exp(-1.0)


def scene(obj, ff=0.0, dm=1.0, deform=None):
    return O.Scene(obj, deform, dm, ff)


def sphere(c=(0.0, 0.0, 0.0), r=0.5, rho=1.0):
    return O.build_object({"type": "sphere", "center": list(c), "radius": r, "rho": rho})


# ---- main_test.go:326-522 (integrator physics) ----------------------------
def test_simple_sphere_center_ray():  # main_test.go:326-339
    got = scene(sphere()).integrate_simple([0, 0, -5], [0, 0, 1], 0.001, 4.5, 5.5)
    assert abs(got - E1) <= 2 * 0.001


def test_simple_slab_attenuation():  # main_test.go:343-361
    slab = O.build_object({"type": "box", "center": [0, 0, 0], "sides": [10, 10, 1],
                           "rho": 2.0})
    got = scene(slab).integrate_simple([0, 0, -5], [0, 0, 1], 0.001, 0, 10)
    assert abs(got - math.exp(-2.0)) <= 2 * 0.001 * 2.0


def test_hierarchical_sphere_center_ray():  # main_test.go:365-378
    got = scene(sphere()).integrate_hierarchical([0, 0, -5], [0, 0, 1], 0.05, 4.5, 5.5)
    assert abs(got - E1) <= 0.05


def test_simple_vs_hierarchical_agreement():  # main_test.go:382-394
    sc = scene(sphere())
    s = sc.integrate_simple([0, 0, -5], [0, 0, 1], 0.001, 4.5, 5.5)
    h = sc.integrate_hierarchical([0, 0, -5], [0, 0, 1], 0.05, 4.5, 5.5)
    assert abs(s - h) <= 0.01


def test_flat_field_applied_once():  # main_test.go:399-420
    sc = scene(sphere((0, 0, 10), 0.01, 1.0), ff=1.0)
    for fn in (sc.integrate_simple, sc.integrate_hierarchical):
        assert abs(fn([0, 0, -5], [0, 0, 1], 0.01, 0, 1) - E1) <= 1e-9


def test_density_multiplier():  # main_test.go:424-437
    got = scene(sphere(), dm=2.0).integrate_simple([0, 0, -5], [0, 0, 1], 0.001, 4.5, 5.5)
    assert abs(got - math.exp(-2.0)) <= 2 * 0.001 * 2.0


@pytest.mark.parametrize("ds,tol", [(0.01, 5e-2), (0.001, 5e-3), (0.0001, 5e-4)])
def test_simple_ds_convergence(ds, tol):  # main_test.go:443-465
    got = scene(sphere()).integrate_simple([0, 0, -5], [0, 0, 1], ds, 4.5, 5.5)
    assert abs(got - E1) < tol


def test_hierarchical_boundary_accuracy():  # main_test.go:476-497
    sc = scene(sphere())
    ref = sc.integrate_simple([0, 0, -5], [0, 0, 1], 0.0001, 4.5, 5.5)
    for DS in (0.1, 0.05, 0.02):
        got = sc.integrate_hierarchical([0, 0, -5], [0, 0, 1], DS, 4.5, 5.5)
        assert abs(got - ref) <= DS


def test_hierarchical_off_center():  # main_test.go:501-522
    sc = scene(sphere())
    d = 0.3
    want = math.exp(-2 * math.sqrt(0.25 - d * d))
    ref = sc.integrate_simple([d, 0, -5], [0, 0, 1], 0.0001, 4.5, 5.5)
    got = sc.integrate_hierarchical([d, 0, -5], [0, 0, 1], 0.05, 4.5, 5.5)
    assert abs(ref - want) <= 1e-3
    assert abs(got - ref) <= 0.05


# ---- main_test.go:90-283 (cameras and angle lists) -------------------------
@pytest.mark.parametrize("az,po,R,eye", [(0, 90, 4.0, (4, 0, 0)), (90, 90, 4.0, (0, 4, 0)),
                                          (0, 0, 4.0, (0, 0, 4))])
def test_camera_eye(az, po, R, eye):  # main_test.go:90-195
    e, cam = O.camera_from_angles(az, po, R)
    assert np.linalg.norm(e - np.array(eye)) <= 1e-6
    if po != 0:
        assert np.max(np.abs(cam)) > 1e-10


def test_camera_distance():  # main_test.go:153-165
    e, _ = O.camera_from_angles(45, 60, 5.0)
    assert abs(np.linalg.norm(e) - 5.0) <= 1e-6


def test_generate_camera_angles():  # main_test.go:197-283
    assert O.generate_camera_angles(4) == [(90.0, 90.0), (180.0, 90.0), (270.0, 90.0),
                                           (360.0, 90.0)]
    assert [a for a, _ in O.generate_camera_angles(8, 1, 2)] == [135.0, 225.0, 315.0, 405.0]
    assert all(p == 45.0 for _, p in O.generate_camera_angles(2, polar_angle=45.0))


# ---- objects/objects_test.go ----------------------------------------------
def test_sphere_truth_table():  # objects_test.go:14-42
    s = sphere((0, 0, 0), 0.5, 2.0)
    assert [s.density(0, 0, z) for z in (0, 0.4, 0.5, 0.6)] == [2.0, 2.0, 0.0, 0.0]
    s2 = sphere((1, 2, 3), 0.3, 5.0)
    assert s2.density(1, 2, 3) == 5.0 and s2.density(1, 2, 3.31) == 0.0


def test_box_truth_table():  # objects_test.go:44-67
    b = O.build_object({"type": "box", "center": [0, 0, 0], "sides": [2, 1, 0.5], "rho": 1.0})
    cases = [((0, 0, 0), 1), ((0.9, 0, 0), 1), ((1.0, 0, 0), 0), ((0, 0.4, 0), 1),
             ((0, 0.5, 0), 0), ((0, 0, 0.24), 1), ((0, 0, 0.25), 0)]
    for p, want in cases:
        assert b.density(*p) == want, p


def test_cylinder_truth_table():  # objects_test.go:69-93
    c = O.build_object({"type": "cylinder", "p0": [0, 0, -1], "p1": [0, 0, 1],
                        "radius": 0.5, "rho": 1.0})
    cases = [((0, 0, 0), 1), ((0.4, 0, 0), 1), ((0.5, 0, 0), 0), ((0, 0, -1.0), 1),
             ((0, 0, 1.0), 1), ((0, 0, -1.001), 0), ((0, 0, 1.001), 0), ((0.4, 0, 1.0), 1)]
    for p, want in cases:
        assert c.density(*p) == want, p


def test_collection_clamp_and_greedy():  # objects_test.go:95-112
    s = {"type": "sphere", "center": [0, 0, 0], "radius": 0.5, "rho": 0.8}
    coll = O.build_object({"type": "object_collection", "objects": [s, s]})
    assert coll.density(0, 0, 0) == 1.0
    greedy = O.build_object({"type": "object_collection", "objects": [s, s],
                             "greedy_dens_eval": True})
    assert greedy.density(0, 0, 0) == 0.8


def vg(rho, nx, ny, nz):
    return O.voxel_node(np.asarray(rho, dtype=np.float64), nx, ny, nz)


def test_voxel_layout():  # objects_test.go:119-147
    N = 3
    r = np.zeros(N ** 3)
    r[0 * 9 + 2 * 3 + 0] = 1.0
    g = vg(r, N, N, N)
    assert abs(g.density(1.0, -1.0, -1.0) - 1.0) <= 1e-9
    assert g.density(-1.0, 1.0, -1.0) <= 1e-9
    r2 = np.zeros(N ** 3)
    r2[1 * 9] = 1.0
    g2 = vg(r2, N, N, N)
    assert abs(g2.density(-1.0, -1.0, 0.0) - 1.0) <= 1e-9
    assert g2.density(-1.0, 0.0, -1.0) <= 1e-9


def test_voxel_outside_bounds_and_corners():  # objects_test.go:209-238
    g = vg(np.ones(8), 2, 2, 2)
    for p in [(1.001, 0, 0), (-1.001, 0, 0), (0, 1.001, 0), (0, -1.001, 0), (0, 0, 1.001),
              (0, 0, -1.001), (2, 2, 2)]:
        assert g.density(*p) == 0.0
    assert g.density(1, 1, 1) == 1.0 and g.density(-1, -1, -1) == 1.0


def test_voxel_trilinear():  # objects_test.go:240-268
    g = vg(np.ones(8), 2, 2, 2)
    for p in [(0, 0, 0), (0.5, -0.5, 0.3), (1, 1, 1), (-1, -1, -1)]:
        assert abs(g.density(*p) - 1.0) <= 1e-9
    r = np.zeros(8)
    r[0] = 1.0
    g2 = vg(r, 2, 2, 2)
    assert abs(g2.density(-1, -1, -1) - 1.0) <= 1e-9
    assert abs(g2.density(0, -1, -1) - 0.5) <= 1e-9


def test_voxel_non_cubic():  # objects_test.go:311-349
    nx, ny, nz = 2, 4, 3
    r = np.zeros(nx * ny * nz)
    r[2 * nx * ny + 1 * ny + 3] = 1.0
    g = vg(r, nx, ny, nz)
    assert abs(g.density(1, 1, 1) - 1.0) <= 1e-9
    assert g.density(-1, 1, 1) <= 1e-9 and g.density(1, -1, 1) <= 1e-9
    r2 = np.zeros(nx * ny * nz)
    r2[1] = 1.0
    g2 = vg(r2, nx, ny, nz)
    wy = 1 / (ny - 1) * 2 - 1
    assert abs(g2.density(-1, wy, -1) - 1.0) <= 1e-9
    assert g2.density(wy, -1, -1) <= 1e-9


def test_voxel_non_cubic_axis_separation():  # objects_test.go:433-461
    nx, ny, nz = 16, 32, 16
    s = sphere((0.5, 0, 0), 0.3, 1.0)
    r = np.zeros(nx * ny * nz)
    for k in range(nz):
        for i in range(nx):
            for j in range(ny):
                r[k * nx * ny + i * ny + j] = s.density(i / (nx - 1) * 2 - 1,
                                                        j / (ny - 1) * 2 - 1,
                                                        k / (nz - 1) * 2 - 1)
    g = vg(r, nx, ny, nz)
    assert g.density(0.5, 0, 0) >= 0.8
    assert g.density(0, 0.5, 0) <= 0.1


def test_voxel_raw_uint8_precision(tmp_path):  # objects_test.go:282-299
    p = tmp_path / "u8.raw"
    p.write_bytes(bytes([255, 128, 0]))
    rho = O._voxels_from_raw(str(p), [1, 1, 3], "uint8")
    assert np.allclose(rho, [1.0, 128 / 255, 0.0], atol=1e-10)


def test_voxel_round_trip(tmp_path):  # objects_test.go:157-205 (ExportToRaw + FromRaw)
    res = 64
    a = sphere((0.5, 0, 0), 0.15, 1.0)
    b = sphere((0, 0.3, 0), 0.08, 1.0)
    x = np.arange(res) / res * 2.0 - 1.0
    Z, X, Y = np.meshgrid(x, x, x, indexing="ij")
    rho = np.minimum(((X - 0.5) ** 2 + Y ** 2 + Z ** 2 < 0.15 ** 2).astype(float) +
                     (X ** 2 + (Y - 0.3) ** 2 + Z ** 2 < 0.08 ** 2).astype(float), 1.0)
    assert a.density(0.5, 0, 0) == 1.0 and b.density(0, 0.3, 0) == 1.0
    g = vg(rho.reshape(-1), res, res, res)
    # ExportToRaw samples at i/res*2-1 and stores byte(v/max*255)
    vol = np.array([[[g.density(i / res * 2 - 1, j / res * 2 - 1, k / res * 2 - 1)
                      for j in range(res)] for i in range(res)] for k in range(0, res, 1)])
    raw = (vol / vol.max() * 255).astype(np.uint8)
    p = tmp_path / "volume.raw"
    p.write_bytes(raw.tobytes())
    g2 = vg(O._voxels_from_raw(str(p), [res] * 3, "uint8"), res, res, res)
    assert g2.density(0.5, 0, 0) > 0.5 and g2.density(0, 0.5, 0) < 0.1
    assert g2.density(0, 0.3, 0) > 0.5 and g2.density(0.3, 0, 0) < 0.1


def tess(uc_objects, ucb, b):
    return O.build_object({
        "type": "tessellated_obj_coll",
        "uc": {"objects": {"type": "object_collection", "objects": uc_objects},
               "xmin": ucb[0], "xmax": ucb[1], "ymin": ucb[0], "ymax": ucb[1],
               "zmin": ucb[0], "zmax": ucb[1]},
        "xmin": b[0], "xmax": b[1], "ymin": b[0], "ymax": b[1], "zmin": b[0], "zmax": b[1]})


def test_tessellation_x_cylinder_seams():  # objects_test.go:356-381
    lat = tess([{"type": "cylinder", "p0": [0, 0.5, 0.5], "p1": [1, 0.5, 0.5],
                 "radius": 0.1, "rho": 1.0}], (0, 1), (-2, 2))
    for x in (-1.9, -1.5, -1.0, -0.5, 0.0, 0.5, 1.0, 1.5, 1.9):
        assert lat.density(x, 0.5, 0.5) != 0
    assert lat.density(1.0 - 1e-9, 0.5, 0.5) != 0 and lat.density(1.0 + 1e-9, 0.5, 0.5) != 0


# Strut endpoints of the Kelvin cell used by objects_test.go:387-428 (the
# MakeKelvin table of objects.go:588-626, test data).
KELVIN = [
    ((0.25, 0.00, 0.50), (0.50, 0.00, 0.75)), ((0.25, 1.00, 0.50), (0.50, 1.00, 0.75)),
    ((0.25, 0.00, 0.50), (0.50, 0.00, 0.25)), ((0.25, 1.00, 0.50), (0.50, 1.00, 0.25)),
    ((0.25, 0.00, 0.50), (0.00, 0.25, 0.50)), ((0.50, 0.00, 0.75), (0.75, 0.00, 0.50)),
    ((0.50, 1.00, 0.75), (0.75, 1.00, 0.50)), ((0.50, 0.00, 0.75), (0.50, 0.25, 1.00)),
    ((0.75, 0.00, 0.50), (0.50, 0.00, 0.25)), ((0.75, 1.00, 0.50), (0.50, 1.00, 0.25)),
    ((0.75, 0.00, 0.50), (1.00, 0.25, 0.50)), ((0.50, 0.00, 0.25), (0.50, 0.25, 0.00)),
    ((1.00, 0.50, 0.75), (0.75, 0.50, 1.00)), ((1.00, 0.75, 0.50), (0.75, 1.00, 0.50)),
    ((1.00, 0.50, 0.25), (0.75, 0.50, 0.00)), ((0.25, 1.00, 0.50), (0.00, 0.75, 0.50)),
    ((0.50, 1.00, 0.75), (0.50, 0.75, 1.00)), ((0.50, 1.00, 0.25), (0.50, 0.75, 0.00)),
    ((0.00, 0.25, 0.50), (0.00, 0.50, 0.75)), ((1.00, 0.25, 0.50), (1.00, 0.50, 0.75)),
    ((0.00, 0.25, 0.50), (0.00, 0.50, 0.25)), ((1.00, 0.25, 0.50), (1.00, 0.50, 0.25)),
    ((0.00, 0.50, 0.75), (0.25, 0.50, 1.00)), ((0.00, 0.50, 0.75), (0.00, 0.75, 0.50)),
    ((1.00, 0.50, 0.75), (1.00, 0.75, 0.50)), ((0.00, 0.75, 0.50), (0.00, 0.50, 0.25)),
    ((1.00, 0.75, 0.50), (1.00, 0.50, 0.25)), ((0.00, 0.50, 0.25), (0.25, 0.50, 0.00)),
    ((0.25, 0.50, 0.00), (0.50, 0.75, 0.00)), ((0.25, 0.50, 1.00), (0.50, 0.75, 1.00)),
    ((0.25, 0.50, 0.00), (0.50, 0.25, 0.00)), ((0.25, 0.50, 1.00), (0.50, 0.25, 1.00)),
    ((0.50, 0.75, 0.00), (0.75, 0.50, 0.00)), ((0.50, 0.75, 1.00), (0.75, 0.50, 1.00)),
    ((0.75, 0.50, 0.00), (0.50, 0.25, 0.00)), ((0.75, 0.50, 1.00), (0.50, 0.25, 1.00)),
]


def kelvin(rad=0.1):
    struts = [{"type": "cylinder", "p0": list(a), "p1": list(b), "radius": rad, "rho": 1.0}
              for a, b in KELVIN]
    return tess(struts, (0.0, 1.0), (-2.0, 2.0))


def test_kelvin_z_boundary():  # objects_test.go:387-408
    lat = kelvin()
    inside = lat.density(0.5, 0.1875, 0.9375)
    outside = lat.density(0.5, 0.1875, 1.0625)
    assert inside != 0 and outside != 0 and abs(inside - outside) <= 1e-9


def test_kelvin_y_face_struts():  # objects_test.go:413-428
    lat = kelvin()
    assert lat.density(0.375, 0.01, 0.625) != 0
    assert lat.density(0.375, -0.01, 0.625) != 0


# ---- deformations/deformations_test.go -------------------------------------
def test_affine_is_linear_map():  # deformations_test.go:11-30
    ident = O.build_deformation({"type": "affine",
                                 "matrix": [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]})
    assert ident.apply(3, 4, 5) == (3, 4, 5)
    sx = O.build_deformation({"type": "affine",
                              "matrix": [[2.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]})
    assert sx.apply(3, 4, 5) == (6, 4, 5)


def test_gaussian_is_displacement():  # deformations_test.go:34-56
    g = O.build_deformation({"type": "gaussian", "amplitudes": [0.5, 0.0, 0.0],
                             "sigmas": [1.0, 1.0, 1.0], "centers": [0.0, 0.0, 0.0]})
    x, y, z = g.apply(0, 0, 0)
    assert abs(x - 0.5) <= 1e-9 and y == 0 and z == 0
    g0 = O.build_deformation({"type": "gaussian", "amplitudes": [0.0, 0.0, 0.0],
                              "sigmas": [1.0, 1.0, 1.0], "centers": [0.0, 0.0, 0.0]})
    assert g0.apply(3, 4, 5) == (3, 4, 5)
