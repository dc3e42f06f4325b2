"""FP64 CPU oracle: ctypes wrapper and scene builder.

TEST INFRASTRUCTURE ONLY.  The oracle is what the HIP library is compared
with (tests/, ``__graft_entry__.smoke()``) and the timed CPU baseline in
``bench.py``.  The product never imports this module.

``build_object`` / ``build_deformation`` restate the reference's
map-to-object factories with their type rules:

* ``NewObject`` / ``FromMap``: objects/objects.go:42-61 (sphere), 100-117
  (cube), 148-169 (box), 209-245 (parallelepiped), 306-331 (cylinder),
  381-420 (collection), 479-510 (unit cell), 536-566 (tessellation),
  658-689 (top-level types), 715-787 + 892-958 (voxel grid), 983-1012
  (gyroid);
* ``NewDeformation`` / ``FromMap``: deformations/deformations.go:49-76,
  101-124, 150-161, 184-195, 234-254, 286-300, 308-343.

Numbers follow the Go decoders: a YAML integer stays an ``int`` (so
``rho: 1`` fails where the reference type-asserts ``float64``), and
every JSON number is a ``float64`` (``numbers="json"``).  Where the
reference panics (an int inside a cube centre, a vector longer than 3,
too few deformation parameters) this builder raises ``ValueError``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []),
                       check=True)
    return LIB_PATH


def _bind(L):
    """Declares the argument and result types of every exported function."""
    P = ctypes.c_void_p
    D = ctypes.c_double
    DP = ctypes.POINTER(ctypes.c_double)
    I = ctypes.c_int
    sig = {
        "orc_sphere": (P, [DP, D, D]),
        "orc_cube": (P, [DP, D, D]),
        "orc_box": (P, [DP, DP, D]),
        "orc_parallelepiped": (P, [DP, DP, DP, DP, D]),
        "orc_cylinder": (P, [DP, DP, D, D]),
        "orc_gyroid": (P, [DP, D, D, D]),
        "orc_collection": (P, [I]),
        "orc_collection_add": (I, [P, P]),
        "orc_unit_cell": (P, [P, DP]),
        "orc_tessellated": (P, [P, DP]),
        "orc_voxel_grid": (P, [DP, I, I, I]),
        "orc_node_free": (None, [P]),
        "orc_density": (D, [P, D, D, D]),
        "orc_min_feature_size": (D, [P]),
        "orc_def_gaussian": (P, [DP, DP, DP]),
        "orc_def_affine": (P, [DP]),
        "orc_def_linear": (P, [DP]),
        "orc_def_rigid": (P, [DP]),
        "orc_def_sigmoid": (P, [D, D, D, I]),
        "orc_def_tanh": (P, [D, D, I, I]),
        "orc_def_composed": (P, []),
        "orc_def_composed_add": (I, [P, P]),
        "orc_deform_free": (None, [P]),
        "orc_deform_apply": (None, [P, DP]),
        "orc_scene_density": (D, [P, D, D, D]),
        "orc_integrate_simple": (D, [P, DP, DP, D, D, D]),
        "orc_integrate_hierarchical": (D, [P, DP, DP, D, D, D]),
        "orc_camera_from_angles": (None, [D, D, D, DP, DP]),
        "orc_pixel_ray": (None, [DP, DP, I, D, I, I, DP, DP]),
        "orc_render": (None, [P, DP, DP, I, I, D, D, D, I, I, I, I, DP]),
        "orc_export_volume": (None, [P, I, I, DP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def build_mutant(mutant: int, out_dir: str):
    """Compiles xrp_oracle.c with -DORC_MUTANT=mutant (a planted mistake,
    listed at the top of that file) into out_dir and returns the loaded
    library.  Used only by the tests that show the oracle's pins reject
    each planted mistake."""
    path = os.path.join(out_dir, f"liboracle_mutant{mutant}.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                    "-std=c11", "-D_GNU_SOURCE", f"-DORC_MUTANT={int(mutant)}",
                    "-shared", "-o", path, os.path.join(HERE, "xrp_oracle.c"),
                    "-lm", "-lpthread"], check=True)
    return _bind(ctypes.CDLL(path))


class using_library:
    """Routes every oracle call made inside the block to library L (a
    mutant from build_mutant); objects built inside must be dropped inside."""

    def __init__(self, L):
        self.L = L

    def __enter__(self):
        global _lib
        lib()
        self.saved, _lib = _lib, self.L
        return self.L

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _vec(vals: Sequence[float], n: int = 3):
    arr = (ctypes.c_double * n)(*[float(v) for v in vals])
    return arr


# ---------------------------------------------------------------------------
# Go type rules
# ---------------------------------------------------------------------------
class _Num:
    """Go's view of a decoded number: YAML ints stay ints, JSON is float64."""

    def __init__(self, numbers: str):
        if numbers not in ("yaml", "json"):
            raise ValueError("numbers must be 'yaml' or 'json'")
        self.json = numbers == "json"

    def is_float(self, v: Any) -> bool:
        if isinstance(v, bool):
            return False
        if isinstance(v, float):
            return True
        return self.json and isinstance(v, int)

    def is_int(self, v: Any) -> bool:
        return (not self.json) and isinstance(v, int) and not isinstance(v, bool)

    def float64(self, v: Any, msg: str) -> float:
        """data[k].(float64)"""
        if self.is_float(v):
            return float(v)
        raise ValueError(msg)

    def to_float64(self, v: Any, msg: str) -> float:
        """objects.ToFloat64 / deformations.toFloat64: int or float64."""
        if self.is_float(v) or self.is_int(v):
            return float(v)
        raise ValueError(msg)

    def to_vec(self, v: Any, msg: str) -> List[float]:
        """objects.ToVec: ints and floats copied, other types left 0."""
        if not isinstance(v, list):
            raise ValueError(msg)
        if len(v) > 3:
            raise ValueError("index out of range: vector has more than 3 entries")
        out = [0.0, 0.0, 0.0]
        for i, e in enumerate(v):
            if self.is_float(e) or self.is_int(e):
                out[i] = float(e)
        return out


class Node:
    """Owned pointer to an oracle object tree."""

    def __init__(self, ptr: int, owned: bool = True):
        self.ptr = ptr
        self.owned = owned

    def release(self) -> int:
        self.owned = False
        return self.ptr

    def density(self, x: float, y: float, z: float) -> float:
        return lib().orc_density(self.ptr, x, y, z)

    def min_feature_size(self) -> float:
        return lib().orc_min_feature_size(self.ptr)

    def __del__(self):
        if getattr(self, "owned", False) and self.ptr:
            lib().orc_node_free(self.ptr)
            self.ptr = None


class Deform:
    def __init__(self, ptr: int, owned: bool = True):
        self.ptr = ptr
        self.owned = owned

    def release(self) -> int:
        self.owned = False
        return self.ptr

    def apply(self, x: float, y: float, z: float):
        p = _vec([x, y, z])
        lib().orc_deform_apply(self.ptr, p)
        return p[0], p[1], p[2]

    def __del__(self):
        if getattr(self, "owned", False) and self.ptr:
            lib().orc_deform_free(self.ptr)
            self.ptr = None


_CHILD_TYPES = ("sphere", "cube", "box", "cylinder", "parallelepiped", "gyroid",
                "tessellated_obj_coll", "voxel_grid")
_TOP_TYPES = ("sphere", "cube", "box", "cylinder", "parallelepiped", "gyroid",
              "object_collection", "tessellated_obj_coll", "voxel_grid")


def _bounds(d: Dict, N: _Num) -> List[float]:
    out = []
    for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
        out.append(N.to_float64(d.get(k), f"{k} is not a float64"))
    return out


def _voxels_from_raw(path: str, res: Sequence[int], dtype: str) -> np.ndarray:
    """objects.go:892-958"""
    sizes = {"uint8": 1, "uint16": 2, "uint32": 4, "float32": 4, "float64": 8}
    if dtype not in sizes:
        raise ValueError(f"unsupported data type: {dtype}")
    with open(path, "rb") as fh:
        data = fh.read()
    expected = res[0] * res[1] * res[2] * sizes[dtype]
    if len(data) != expected:
        raise ValueError(f"file size ({len(data)}) does not match expected size "
                         f"({expected}) for type {dtype}")
    if dtype == "uint8":
        return np.frombuffer(data, np.uint8).astype(np.float64) / 255.0
    if dtype == "uint16":
        return np.frombuffer(data, "<u2").astype(np.float64) / 65535.0
    if dtype == "uint32":
        return np.frombuffer(data, "<u4").astype(np.float64) / 4294967295.0
    if dtype == "float32":
        return np.frombuffer(data, "<f4").astype(np.float64)
    return np.frombuffer(data, "<f8").astype(np.float64)


def _object(d: Any, N: _Num, allowed: Sequence[str], base_dir: str) -> Node:
    L = lib()
    if not isinstance(d, dict):
        raise ValueError("object is not a map")
    t = d.get("type")
    if t not in allowed:
        if allowed is _TOP_TYPES:
            raise ValueError(f"unknown object type `{t}`")
        raise ValueError("unknown object type")
    if t == "sphere":
        c = d.get("center")
        if not isinstance(c, list):
            raise ValueError("center is not a Vec3")
        if len(c) > 3:
            raise ValueError("index out of range: center has more than 3 entries")
        cv = [0.0, 0.0, 0.0]
        for i, e in enumerate(c):
            cv[i] = N.to_float64(e, f"center[{i}] is not a float64")
        r = N.float64(d.get("radius"), "radius is not a float64")
        rho = N.float64(d.get("rho"), "rho is not a float64")
        return Node(L.orc_sphere(_vec(cv), r, rho))
    if t == "cube":
        c = d.get("center")
        if not isinstance(c, list):
            raise ValueError("center is not a Vec3")
        if len(c) > 3:
            raise ValueError("index out of range: center has more than 3 entries")
        cv = [0.0, 0.0, 0.0]
        for i, e in enumerate(c):
            if not N.is_float(e):  # val.(float64) panics on anything else
                raise ValueError("interface conversion: cube center entry is not float64")
            cv[i] = float(e)
        side = N.float64(d.get("side"), "side is not a float64")
        rho = N.float64(d.get("rho"), "rho is not a float64")
        return Node(L.orc_cube(_vec(cv), side, rho))
    if t == "box":
        c = N.to_vec(d.get("center"), "center is not a Vec3")
        s = N.to_vec(d.get("sides"), "sides is not a Vec3")
        rho = N.to_float64(d.get("rho"), "rho is not a float64")
        return Node(L.orc_box(_vec(c), _vec(s), rho))
    if t == "parallelepiped":
        o = N.to_vec(d.get("origin"), "origin is not a Vec3")
        v0 = N.to_vec(d.get("v0"), "v0 is not a Vec3")
        v1 = N.to_vec(d.get("v1"), "v1 is not a Vec3")
        v2 = N.to_vec(d.get("v2"), "v2 is not a Vec3")
        rho = N.to_float64(d.get("rho"), "rho is not a float64")
        return Node(L.orc_parallelepiped(_vec(o), _vec(v0), _vec(v1), _vec(v2), rho))
    if t == "cylinder":
        p0 = N.to_vec(d.get("p0"), "p0 is not a Vec3")
        p1 = N.to_vec(d.get("p1"), "p0 is not a Vec3")
        r = N.float64(d.get("radius"), "radius is not a float64")
        rho = 1.0 if "rho" not in d else N.to_float64(d["rho"], "rho is not a float64")
        return Node(L.orc_cylinder(_vec(p0), _vec(p1), r, rho))
    if t == "gyroid":
        c = d.get("center")
        if not isinstance(c, list):
            raise ValueError("center is not a Vec3")
        if len(c) > 3:
            raise ValueError("index out of range: center has more than 3 entries")
        cv = [0.0, 0.0, 0.0]
        for i, e in enumerate(c):
            cv[i] = N.to_float64(e, f"center[{i}] is not a float64")
        sc = N.float64(d.get("scale"), "scale is not a float64")
        th = N.float64(d.get("thickness"), "thickness is not a float64")
        rho = N.float64(d.get("rho"), "rho is not a float64")
        return Node(L.orc_gyroid(_vec(cv), sc, th, rho))
    if t == "object_collection":
        return _collection(d, N, base_dir)
    if t == "tessellated_obj_coll":
        ucd = d.get("uc")
        if not isinstance(ucd, dict):
            raise ValueError("uc is not a map")
        objs = ucd.get("objects")
        if not isinstance(objs, dict):
            raise ValueError("objects is not a map")
        coll = _collection(objs, N, base_dir)
        ucb = _bounds(ucd, N)
        uc = Node(L.orc_unit_cell(coll.release(), _vec(ucb, 6)))
        tb = _bounds(d, N)
        return Node(L.orc_tessellated(uc.release(), _vec(tb, 6)))
    if t == "voxel_grid":
        return _voxel(d, N, base_dir)
    raise ValueError("unknown object type")


def _collection(d: Dict, N: _Num, base_dir: str) -> Node:
    L = lib()
    greedy = d.get("greedy_dens_eval")
    greedy = bool(greedy) if isinstance(greedy, bool) else False
    objs = d.get("objects")
    if not isinstance(objs, list):
        raise ValueError("objects is not a list")
    coll = Node(L.orc_collection(1 if greedy else 0))
    for od in objs:
        child = _object(od, N, _CHILD_TYPES, base_dir)
        L.orc_collection_add(coll.ptr, child.release())
    return coll


def _voxel(d: Dict, N: _Num, base_dir: str) -> Node:
    L = lib()
    path = d.get("path")
    if isinstance(path, str):
        ext = path[path.rfind(".") + 1:].lower()
        if ext != "raw":
            raise ValueError("only raw files are supported")
        res = d.get("resolution")
        if not isinstance(res, list):
            raise ValueError("resolution must be provided for raw files as a list of 3 integers")
        if len(res) != 3:
            raise ValueError("resolution must be a list of 3 integers")
        r3 = []
        for i, v in enumerate(res):
            if N.is_int(v) or N.is_float(v):
                r3.append(int(v))
            else:
                raise ValueError(f"resolution[{i}] has unsupported type (need int-like)")
        dtype = d.get("dtype") if isinstance(d.get("dtype"), str) else "uint8"
        full = path if os.path.isabs(path) else os.path.join(base_dir, path)
        rho = _voxels_from_raw(full, r3, dtype)
        return voxel_node(rho, r3[0], r3[1], r3[2])
    for k in ("nx", "ny", "nz"):
        if not N.is_int(d.get(k)):
            raise ValueError(f"{k} is not an int")
    raise ValueError("path is not a string")


def voxel_node(rho: np.ndarray, nx: int, ny: int, nz: int) -> Node:
    """VoxelGrid with flat [z][x][y] densities (objects.go:691-698)."""
    arr = np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(-1))
    if arr.size != nx * ny * nz:
        raise ValueError("voxel count does not match nx*ny*nz")
    ptr = lib().orc_voxel_grid(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               nx, ny, nz)
    return Node(ptr)


def build_object(data: Dict, numbers: str = "yaml", base_dir: str = ".") -> Node:
    """objects.NewObject (objects.go:658-689)."""
    return _object(data, _Num(numbers), _TOP_TYPES, base_dir)


def _flist(v: Any, N: _Num, msg: str, need: int) -> List[float]:
    if not isinstance(v, list):
        raise ValueError(msg)
    out = []
    for e in v:
        if not N.is_float(e):
            raise ValueError("interface conversion: deformation entry is not float64")
        out.append(float(e))
    if len(out) < need:
        raise ValueError(f"index out of range: need {need} entries")
    return out


def build_deformation(data: Dict, numbers: str = "yaml") -> Deform:
    """deformations.NewDeformation (deformations.go:308-343)."""
    N = _Num(numbers)
    L = lib()
    if not isinstance(data, dict) or data.get("type") is None:
        raise ValueError("deformation type is nil")
    t = data["type"]
    if t == "gaussian":
        a = _flist(data.get("amplitudes"), N, "amplitudes must be a list", 3)
        s = _flist(data.get("sigmas"), N, "sigmas must be a list", 3)
        c = _flist(data.get("centers"), N, "centers must be a list", 3)
        return Deform(L.orc_def_gaussian(_vec(a[:3]), _vec(s[:3]), _vec(c[:3])))
    if t == "affine":
        m = data.get("matrix")
        if not isinstance(m, list):
            raise ValueError("matrix must be a list")
        if len(m) != 3:
            raise ValueError("matrix must have 3 rows")
        flat = []
        for row in m:
            if not isinstance(row, list):
                raise ValueError("matrix row must be a list")
            if len(row) != 3:
                raise ValueError("matrix row must have 3 elements")
            for e in row:
                if not N.is_float(e):
                    raise ValueError("interface conversion: matrix entry is not float64")
                flat.append(float(e))
        return Deform(L.orc_def_affine(_vec(flat, 9)))
    if t == "linear":
        s = _flist(data.get("strains"), N, "strains must be a list", 6)
        return Deform(L.orc_def_linear(_vec(s[:6], 6)))
    if t == "rigid":
        dsp = _flist(data.get("displacements"), N, "displacements must be a list", 3)
        return Deform(L.orc_def_rigid(_vec(dsp[:3])))
    if t == "sigmoid":
        amp = N.to_float64(data.get("amplitude"), "amplitude must be a float")
        cen = N.to_float64(data.get("center"), "center must be a float")
        ls = N.to_float64(data.get("lengthscale"), "lengthscale must be a float")
        dr = data.get("direction")
        if not isinstance(dr, str):
            raise ValueError("direction must be a string")
        if dr not in ("x", "y", "z"):
            raise ValueError("Invalid direction")
        return Deform(L.orc_def_sigmoid(amp, cen, ls, "xyz".index(dr)))
    if t == "tanh_shear":
        # extension (no deformations.go counterpart): p[direction] +=
        # amplitude * tanh(rate * p[source])
        amp = N.to_float64(data.get("amplitude"), "amplitude must be a float")
        rate = N.to_float64(data.get("rate"), "rate must be a float")
        dr, src = data.get("direction"), data.get("source")
        if not isinstance(dr, str) or not isinstance(src, str):
            raise ValueError("direction and source must be strings")
        if dr not in ("x", "y", "z") or src not in ("x", "y", "z") or dr == src:
            raise ValueError("Invalid direction")
        return Deform(L.orc_def_tanh(amp, rate, "xyz".index(dr), "xyz".index(src)))
    if t == "composed":
        lst = data.get("deformations")
        if not isinstance(lst, list):
            raise ValueError("deformations must be a list")
        comp = Deform(L.orc_def_composed())
        for sub in lst:
            child = build_deformation(sub, numbers)
            L.orc_def_composed_add(comp.ptr, child.release())
        return comp
    raise ValueError(f"unknown deformation type {t}")


# ---------------------------------------------------------------------------
# Render core
# ---------------------------------------------------------------------------
class _SceneC(ctypes.Structure):
    _fields_ = [("obj", ctypes.c_void_p), ("def_", ctypes.c_void_p),
                ("density_multiplier", ctypes.c_double),
                ("flat_field", ctypes.c_double)]


class Scene:
    """lat[0], df[0], density_multiplier, flat_field (main.go:37-41)."""

    def __init__(self, obj: Node, deform: Optional[Deform] = None,
                 density_multiplier: float = 1.0, flat_field: float = 0.0):
        self.obj = obj
        self.deform = deform
        self.c = _SceneC(obj.ptr, deform.ptr if deform else None,
                         density_multiplier, flat_field)

    @property
    def ref(self):
        return ctypes.byref(self.c)

    def density(self, x, y, z) -> float:
        return lib().orc_scene_density(self.ref, x, y, z)

    def integrate_simple(self, origin, direction, ds, smin, smax) -> float:
        return lib().orc_integrate_simple(self.ref, _vec(origin), _vec(direction),
                                          ds, smin, smax)

    def integrate_hierarchical(self, origin, direction, DS, smin, smax) -> float:
        return lib().orc_integrate_hierarchical(self.ref, _vec(origin),
                                                _vec(direction), DS, smin, smax)

    def render(self, azimuthal: Sequence[float], polar: Sequence[float], res: int,
               ds: float, R: float = 4.0, fov: float = 40.0,
               integrator: str = "simple", nthreads: Optional[int] = None,
               rows: Optional[Sequence[int]] = None) -> np.ndarray:
        """[view][i][j] float64 transmittance (main.go:429-471)."""
        nv = len(azimuthal)
        az = np.ascontiguousarray(azimuthal, dtype=np.float64)
        po = np.ascontiguousarray(polar, dtype=np.float64)
        out = np.full((nv, res, res), np.nan, dtype=np.float64)
        r0, r1 = (0, res) if rows is None else (int(rows[0]), int(rows[1]))
        nt = nthreads or os.cpu_count() or 1
        DP = ctypes.POINTER(ctypes.c_double)
        lib().orc_render(self.ref, az.ctypes.data_as(DP), po.ctypes.data_as(DP), nv,
                         res, ds, R, fov, 1 if integrator == "hierarchical" else 0,
                         nt, r0, r1, out.ctypes.data_as(DP))
        return out

    def export_volume(self, res: int, nthreads: Optional[int] = None) -> np.ndarray:
        out = np.zeros(res * res * res, dtype=np.float64)
        DP = ctypes.POINTER(ctypes.c_double)
        lib().orc_export_volume(self.ref, res, nthreads or os.cpu_count() or 1,
                                out.ctypes.data_as(DP))
        return out


def camera_from_angles(az: float, polar: float, R: float = 4.0):
    eye = _vec([0, 0, 0])
    cam = (ctypes.c_double * 16)()
    lib().orc_camera_from_angles(az, polar, R, eye, cam)
    return np.array(eye[:]), np.array(cam[:]).reshape(4, 4)


def pixel_ray(eye, cam, res: int, fov: float, i: int, j: int):
    o = _vec([0, 0, 0])
    d = _vec([0, 0, 0])
    c = (ctypes.c_double * 16)(*np.asarray(cam, dtype=np.float64).reshape(-1))
    lib().orc_pixel_ray(_vec(eye), c, res, fov, i, j, o, d)
    return np.array(o[:]), np.array(d[:])


def generate_camera_angles(num_images: int, job_num: int = 0, jobs_modulo: int = 1,
                           polar_angle: float = 90.0):
    """main.go:242-257 without the random out-of-plane branch."""
    out = []
    for i_img in range(job_num, num_images, jobs_modulo):
        dth = 360.0 / float(num_images)
        th = float(i_img) * dth + 90.0
        out.append((th, polar_angle))
    return out
