"""ctypes binding of libxrp's C ABI (include/xrp.h).

The shared library is built in-tree by ``xray_projection_render_amd.build``.
Nothing here falls back to a CPU implementation: when the library or a GPU
is missing the calls raise.
"""
from __future__ import annotations

import ctypes
import weakref
import json
import os
from typing import Dict, Iterable, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("XRP_LIB") or os.path.join(HERE, "libxrp.so")

XRP_OK = 0
STATUS_NAMES = {
    0: "XRP_OK", 1: "XRP_ERR_NULL", 2: "XRP_ERR_DIMS", 3: "XRP_ERR_DEVICE",
    4: "XRP_ERR_ALLOC", 10: "XRP_ERR_SCENE", 11: "XRP_ERR_DEFORMATION",
    12: "XRP_ERR_ARG", 13: "XRP_ERR_UNSUPPORTED", 14: "XRP_ERR_IO", 15: "XRP_ERR_INTERNAL",
}
INTEGRATORS = {"simple": 0, "hierarchical": 1}
PRECISIONS = {"auto": 0, "guarded": 1, "fp64": 2}
NUMBERS = {"json": 0, "yaml": 1}

EXPORTED = [
    "xrp_abi_version", "xrp_last_error", "xrp_default_params",
    "xrp_context_create", "xrp_context_destroy", "xrp_context_synchronize",
    "xrp_scene_create", "xrp_scene_create_voxels", "xrp_scene_destroy",
    "xrp_scene_min_feature_size", "xrp_scene_auto_ds", "xrp_scene_clip_box",
    "xrp_scene_layout", "xrp_camera_from_angles", "xrp_sample_count",
    "xrp_render", "xrp_render_device", "xrp_render_volume_projections",
]


class XrpError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code
        self.message = message


class RenderParams(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("resolution", ctypes.c_int32),
        ("ds", ctypes.c_double), ("R", ctypes.c_double),
        ("fov_deg", ctypes.c_double), ("flat_field", ctypes.c_double),
        ("density_multiplier", ctypes.c_double),
        ("integrator", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("audit", ctypes.c_int32), ("output_format", ctypes.c_int32),
        ("transparency", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class _OutStruct(ctypes.Structure):
    """Output structs open with struct_size, set here to the size this
    binding declares (the library writes no further)."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.struct_size = ctypes.sizeof(self)

    def as_dict(self) -> Dict:
        return {k: getattr(self, k) for k, _ in self._fields_
                if k not in ("reserved", "reserved0", "reserved1", "struct_size")}


class RenderStats(_OutStruct):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("reserved0", ctypes.c_int32),
        ("rays", ctypes.c_uint64), ("samples", ctypes.c_uint64),
        ("reevaluated", ctypes.c_uint64), ("audit_mismatch", ctypes.c_uint64),
        ("fine_samples", ctypes.c_uint64), ("kernel_ms", ctypes.c_double),
        ("position_error", ctypes.c_double), ("precision_used", ctypes.c_int32),
        ("block", ctypes.c_int32), ("staged_chunks", ctypes.c_uint64),
        ("staged_voxels", ctypes.c_uint64), ("fallback_chunks", ctypes.c_uint64),
        ("skipped_samples", ctypes.c_uint64), ("redo_rays", ctypes.c_uint64),
        ("deferred_rays", ctypes.c_uint64), ("two_pass", ctypes.c_int32),
        ("reserved1", ctypes.c_int32),
    ]


class ViewAngle(ctypes.Structure):
    _fields_ = [("azimuthal_deg", ctypes.c_double), ("polar_deg", ctypes.c_double)]


class Camera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_double * 3), ("cam2world", ctypes.c_double * 16),
                ("fov_deg", ctypes.c_double), ("R", ctypes.c_double)]


_lib = None


class BatchItem(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("views", ctypes.POINTER(ViewAngle)),
                ("n_views", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("params", RenderParams)]


class TileShape(ctypes.Structure):
    _fields_ = [("wave_rows", ctypes.c_int32), ("wave_cols", ctypes.c_int32),
                ("groups_per_cu", ctypes.c_int32), ("reserved", ctypes.c_int32)]

    def as_tuple(self) -> Tuple[int, int, int]:
        return (self.wave_rows, self.wave_cols, self.groups_per_cu)


class BatchStats(_OutStruct):
    _fields_ = [("struct_size", ctypes.c_uint32), ("reserved0", ctypes.c_int32),
                ("kernel_ms", ctypes.c_double), ("rays", ctypes.c_uint64),
                ("samples", ctypes.c_uint64), ("reevaluated", ctypes.c_uint64),
                ("tiles", ctypes.c_uint64), ("launches", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class SurrogateConfig(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("log2_table", ctypes.c_int32),
                ("n_min", ctypes.c_int32), ("n_max", ctypes.c_int32), ("wrap", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("seed", ctypes.c_uint64)]


class TrainConfig(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("batch", ctypes.c_int32),
                ("samples", ctypes.c_int64), ("epochs", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("lr_table", ctypes.c_double),
                ("lr_mlp", ctypes.c_double)]


class TrainStats(_OutStruct):
    _fields_ = [("struct_size", ctypes.c_uint32), ("reserved0", ctypes.c_int32),
                ("first_loss", ctypes.c_double), ("final_loss", ctypes.c_double),
                ("label_ms", ctypes.c_double), ("train_ms", ctypes.c_double),
                ("steps", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class SurrogateStats(_OutStruct):
    _fields_ = [("struct_size", ctypes.c_uint32), ("reserved0", ctypes.c_int32),
                ("surrogate_ms", ctypes.c_double), ("check_ms", ctypes.c_double),
                ("fallback_ms", ctypes.c_double), ("max_sampled_error", ctypes.c_double),
                ("tiles", ctypes.c_uint64), ("tiles_failed", ctypes.c_uint64),
                ("check_pixels", ctypes.c_uint64), ("network_samples", ctypes.c_uint64),
                ("exact_samples", ctypes.c_uint64)]


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libxrp.so not found at {LIB_PATH}; build it with "
            "`python -m xray_projection_render_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    DP = ctypes.POINTER(ctypes.c_double)
    FP = ctypes.POINTER(ctypes.c_float)
    PP = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "xrp_abi_version": (I, []),
        "xrp_last_error": (ctypes.c_char_p, []),
        "xrp_default_params": (None, [ctypes.POINTER(RenderParams)]),
        "xrp_context_create": (I, [I, PP]),
        "xrp_context_destroy": (None, [P]),
        "xrp_context_synchronize": (I, [P]),
        "xrp_scene_create": (I, [ctypes.c_char_p, ctypes.c_char_p, I, ctypes.c_char_p, PP]),
        "xrp_scene_create_voxels": (I, [FP, I, I, I, ctypes.c_char_p, I, PP]),
        "xrp_scene_destroy": (None, [P]),
        "xrp_scene_min_feature_size": (I, [P, DP]),
        "xrp_scene_auto_ds": (I, [P, DP]),
        "xrp_scene_clip_box": (I, [P, DP, DP, ctypes.POINTER(I)]),
        "xrp_scene_layout": (I, [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)]),
        "xrp_camera_from_angles": (I, [D, D, D, DP, DP]),
        "xrp_sample_count": (I, [D, D, D, ctypes.POINTER(ctypes.c_int64)]),
        "xrp_render": (I, [P, P, ctypes.POINTER(RenderParams), ctypes.POINTER(ViewAngle),
                           I, P, ctypes.POINTER(RenderStats)]),
        "xrp_render_device": (I, [P, P, ctypes.POINTER(RenderParams),
                                  ctypes.POINTER(ViewAngle), I, P,
                                  ctypes.POINTER(RenderStats)]),
        "xrp_render_device_noalloc": (I, [P, P, ctypes.POINTER(RenderParams),
                                          ctypes.POINTER(ViewAngle), I, P,
                                          ctypes.POINTER(RenderStats)]),
        "xrp_render_volume_projections": (I, [P, FP, I, I, I, ctypes.POINTER(Camera), I, I,
                                              D, D, FP]),
        "xrp_export_volume": (I, [P, P, I, D, ctypes.POINTER(ctypes.c_uint8)]),
        "xrp_scene_load_files": (I, [ctypes.c_char_p, ctypes.c_char_p, PP]),
        "xrp_batch_create": (I, [P, ctypes.POINTER(BatchItem), I, PP]),
        "xrp_batch_set_shape": (I, [P, ctypes.POINTER(TileShape)]),
        "xrp_batch_tune": (I, [P, I, ctypes.POINTER(TileShape)]),
        "xrp_batch_run": (I, [P, ctypes.POINTER(BatchStats)]),
        "xrp_batch_output": (I, [P, I, ctypes.POINTER(FP)]),
        "xrp_batch_read": (I, [P, I, FP]),
        "xrp_batch_latencies": (I, [P, DP]),
        "xrp_batch_destroy": (None, [P]),
        "xrp_surrogate_default_config": (None, [ctypes.POINTER(SurrogateConfig)]),
        "xrp_train_default_config": (None, [ctypes.POINTER(TrainConfig)]),
        "xrp_surrogate_create": (I, [P, P, ctypes.POINTER(SurrogateConfig), PP]),
        "xrp_surrogate_train": (I, [P, ctypes.POINTER(TrainConfig), ctypes.POINTER(TrainStats)]),
        "xrp_surrogate_render_device": (I, [P, ctypes.POINTER(RenderParams),
                                            ctypes.POINTER(ViewAngle), I, D, D, P,
                                            ctypes.POINTER(SurrogateStats)]),
        "xrp_surrogate_render": (I, [P, ctypes.POINTER(RenderParams), ctypes.POINTER(ViewAngle),
                                     I, D, D, FP, ctypes.POINTER(SurrogateStats)]),
        "xrp_surrogate_param_count": (ctypes.c_int64, [P]),
        "xrp_surrogate_get_params": (I, [P, FP]),
        "xrp_surrogate_set_params": (I, [P, FP]),
        "xrp_surrogate_predict": (I, [P, FP, ctypes.c_int64, FP]),
        "xrp_surrogate_gradients": (I, [P, FP, FP, ctypes.c_int64, FP, DP]),
        "xrp_surrogate_destroy": (None, [P]),
        "xrp_diag_mfma_bf16": (I, [P, FP, FP, FP]),
        "xrp_diag_l2_bandwidth": (I, [P, DP]),
        "xrp_diag_sincos_error": (I, [P, DP, DP]),
        "xrp_diag_deform_error": (I, [P, DP, DP, DP]),
        "RenderProjections": (ctypes.POINTER(ctypes.c_char), [ctypes.c_char_p]),
        "FreeString": (None, [ctypes.POINTER(ctypes.c_char)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code: int) -> None:
    if code != XRP_OK:
        msg = lib().xrp_last_error()
        raise XrpError(code, msg.decode() if msg else "")


def _dp(arr: np.ndarray):
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def camera_from_angles(az: float, polar: float, R: float = 4.0) -> Tuple[np.ndarray, np.ndarray]:
    eye = np.zeros(3)
    cam = np.zeros(16)
    check(lib().xrp_camera_from_angles(az, polar, R, _dp(eye), _dp(cam)))
    return eye, cam.reshape(4, 4)


def sample_count(smin: float, smax: float, ds: float) -> int:
    n = ctypes.c_int64()
    check(lib().xrp_sample_count(smin, smax, ds, ctypes.byref(n)))
    return n.value


class Context:
    """One device, one HIP stream and the cached device buffers."""

    def __init__(self, device: int = 0):
        p = ctypes.c_void_p()
        check(lib().xrp_context_create(device, ctypes.byref(p)))
        self.ptr = p
        self._owned = weakref.WeakSet()  # batches and surrogates living on it

    def close(self) -> None:
        """Destroys what lives on the context (batches, surrogates), then the
        context itself."""
        if getattr(self, "ptr", None):
            for obj in list(getattr(self, "_owned", ())):
                obj.close()
            lib().xrp_context_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """An object tree (+ optional deformation) parsed by libxrp."""

    def __init__(self, ptr):
        self.ptr = ptr

    @classmethod
    def from_text(cls, object_text: str, deformation_text: Optional[str] = None,
                  numbers: str = "json", base_dir: Optional[str] = None) -> "Scene":
        p = ctypes.c_void_p()
        check(lib().xrp_scene_create(
            object_text.encode(), deformation_text.encode() if deformation_text else None,
            NUMBERS[numbers], base_dir.encode() if base_dir else None, ctypes.byref(p)))
        return cls(p)

    @classmethod
    def from_dict(cls, obj: Dict, deformation: Optional[Dict] = None,
                  numbers: str = "yaml", base_dir: Optional[str] = None) -> "Scene":
        """Dicts keep Python int/float, so YAML-decoded data passes through
        with yaml.v3 number semantics (numbers="yaml")."""
        return cls.from_text(json.dumps(obj), json.dumps(deformation) if deformation else None,
                             numbers, base_dir)

    @classmethod
    def from_files(cls, object_path: str, deformation_path: Optional[str] = None) -> "Scene":
        """load_object / load_deformation (main.go:50-120): YAML or JSON by extension."""
        p = ctypes.c_void_p()
        check(lib().xrp_scene_load_files(object_path.encode(),
                                          deformation_path.encode() if deformation_path else None,
                                          ctypes.byref(p)))
        return cls(p)

    @classmethod
    def from_voxels(cls, rho: np.ndarray, deformation: Optional[Dict] = None,
                    numbers: str = "yaml") -> "Scene":
        """rho shaped (nz, nx, ny): flat index z*nx*ny + x*ny + y."""
        arr = np.ascontiguousarray(rho, dtype=np.float32)
        nz, nx, ny = arr.shape
        p = ctypes.c_void_p()
        check(lib().xrp_scene_create_voxels(
            arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), nx, ny, nz,
            json.dumps(deformation).encode() if deformation else None,
            NUMBERS[numbers], ctypes.byref(p)))
        return cls(p)

    def min_feature_size(self) -> float:
        v = ctypes.c_double()
        check(lib().xrp_scene_min_feature_size(self.ptr, ctypes.byref(v)))
        return v.value

    def auto_ds(self) -> float:
        v = ctypes.c_double()
        check(lib().xrp_scene_auto_ds(self.ptr, ctypes.byref(v)))
        return v.value

    def clip_box(self):
        lo, hi = np.zeros(3), np.zeros(3)
        e = ctypes.c_int()
        check(lib().xrp_scene_clip_box(self.ptr, _dp(lo), _dp(hi), ctypes.byref(e)))
        return lo, hi, bool(e.value)

    def layout(self):
        n, d, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().xrp_scene_layout(self.ptr, ctypes.byref(n), ctypes.byref(d),
                                     ctypes.byref(v)))
        return n.value, d.value, bool(v.value)

    def close(self) -> None:
        if getattr(self, "ptr", None):
            lib().xrp_scene_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_params(resolution: int, ds: float, R: float = 4.0, fov: float = 40.0,
                flat_field: float = 0.0, density_multiplier: float = 1.0,
                integrator: str = "simple", precision: str = "auto",
                audit: bool = False, output: str = "f32",
                transparency: bool = False) -> RenderParams:
    p = RenderParams()
    lib().xrp_default_params(ctypes.byref(p))
    p.resolution = resolution
    p.ds = ds
    p.R = R
    p.fov_deg = fov
    p.flat_field = flat_field
    p.density_multiplier = density_multiplier
    p.integrator = INTEGRATORS[integrator]
    p.precision = PRECISIONS[precision]
    p.audit = 1 if audit else 0
    p.output_format = {"f32": 0, "gray8_alpha8": 1}[output]
    p.transparency = 1 if transparency else 0
    return p


def _views(angles: Sequence[Tuple[float, float]]):
    arr = (ViewAngle * len(angles))()
    for k, (az, po) in enumerate(angles):
        arr[k].azimuthal_deg = az
        arr[k].polar_deg = po
    return arr


def render(ctx: Context, scene: Scene, angles: Sequence[Tuple[float, float]],
           resolution: int, ds: float, **kw) -> Tuple[np.ndarray, Dict]:
    """Transmittance images [view][i][j] (float32) into host memory, or
    (gray, alpha) byte pairs [view][i][j][2] with output="gray8_alpha8"."""
    p = make_params(resolution, ds, **kw)
    if p.output_format == 1:
        out = np.empty((len(angles), resolution, resolution, 2), dtype=np.uint8)
    else:
        out = np.empty((len(angles), resolution, resolution), dtype=np.float32)
    st = RenderStats()
    check(lib().xrp_render(ctx.ptr, scene.ptr, ctypes.byref(p), _views(angles),
                           len(angles), out.ctypes.data_as(ctypes.c_void_p),
                           ctypes.byref(st)))
    return out, st.as_dict()


def render_device(ctx: Context, scene: Scene, angles: Sequence[Tuple[float, float]],
                  resolution: int, ds: float, out_ptr: int, **kw) -> Dict:
    """Render into device memory at out_ptr (e.g. a torch tensor's data_ptr())."""
    p = make_params(resolution, ds, **kw)
    st = RenderStats()
    check(lib().xrp_render_device(ctx.ptr, scene.ptr, ctypes.byref(p), _views(angles),
                                  len(angles), ctypes.c_void_p(out_ptr), ctypes.byref(st)))
    return st.as_dict()


def render_device_noalloc(ctx: Context, scene: Scene, angles: Sequence[Tuple[float, float]],
                          resolution: int, ds: float, out_ptr: int, **kw) -> Dict:
    """render_device without any allocation in the library (its per-call data
    goes into a fixed per-GPU scratch region; xrp_render_device_noalloc)."""
    p = make_params(resolution, ds, **kw)
    st = RenderStats()
    check(lib().xrp_render_device_noalloc(ctx.ptr, scene.ptr, ctypes.byref(p), _views(angles),
                                          len(angles), ctypes.c_void_p(out_ptr),
                                          ctypes.byref(st)))
    return st.as_dict()


def export_volume(ctx: Context, scene: Scene, res: int,
                  density_multiplier: float = 1.0) -> np.ndarray:
    """volume.raw bytes of main.go:549-635, shaped (res, res, res) = (z, x, y)."""
    out = np.empty((res, res, res), dtype=np.uint8)
    check(lib().xrp_export_volume(ctx.ptr, scene.ptr, res, density_multiplier,
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
    return out


def render_projections_json(params: Dict) -> Dict:
    """The library's RenderProjections (api.go:82-180) on a parameter dict."""
    L = lib()
    ptr = L.RenderProjections(json.dumps(params).encode())
    try:
        return json.loads(ctypes.string_at(ptr).decode())
    finally:
        L.FreeString(ptr)


def render_volume_projections(ctx: Context, volume: np.ndarray,
                              cameras: Iterable[Tuple[np.ndarray, np.ndarray, float, float]],
                              image_res: int, ds: float, flat_field: float = 0.0) -> np.ndarray:
    """volume shaped (nz, nx, ny); cameras = (eye, cam2world 4x4, fov_deg, R)."""
    vol = np.ascontiguousarray(volume, dtype=np.float32)
    nz, nx, ny = vol.shape
    cams = list(cameras)
    arr = (Camera * len(cams))()
    for k, (eye, cam, fov, R) in enumerate(cams):
        arr[k].eye[:] = [float(v) for v in eye]
        arr[k].cam2world[:] = [float(v) for v in np.asarray(cam).reshape(-1)]
        arr[k].fov_deg = fov
        arr[k].R = R
    out = np.empty((len(cams), image_res, image_res), dtype=np.float32)
    check(lib().xrp_render_volume_projections(
        ctx.ptr, vol.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), nx, ny, nz, arr,
        len(cams), image_res, ds, flat_field,
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return out


class Batch:
    """Many scenes rendered by one graph-captured persistent launch
    (xrp_batch_*).  items: (scene, angles, resolution, ds, render kwargs)."""

    def __init__(self, ctx: Context, items: Sequence[Tuple]):
        self._keep = []
        arr = (BatchItem * len(items))()
        self.shapes = []
        for k, it in enumerate(items):
            scene, angles, res, ds = it[:4]
            kw = it[4] if len(it) > 4 else {}
            views = _views(angles)
            self._keep.append((scene, views))
            arr[k].scene = scene.ptr
            arr[k].views = ctypes.cast(views, ctypes.POINTER(ViewAngle))
            arr[k].n_views = len(angles)
            arr[k].params = make_params(res, ds, **kw)
            self.shapes.append((len(angles), res, res))
        ptr = ctypes.c_void_p()
        check(lib().xrp_batch_create(ctx.ptr, arr, len(items), ctypes.byref(ptr)))
        self.ptr = ptr
        self.ctx = ctx
        ctx._owned.add(self)

    def set_shape(self, wave_rows: int, wave_cols: int, groups_per_cu: int) -> None:
        t = TileShape(wave_rows, wave_cols, groups_per_cu, 0)
        check(lib().xrp_batch_set_shape(self.ptr, ctypes.byref(t)))

    def tune(self, repeats: int = 2) -> Tuple[int, int, int]:
        t = TileShape()
        check(lib().xrp_batch_tune(self.ptr, repeats, ctypes.byref(t)))
        return t.as_tuple()

    def run(self) -> Dict:
        st = BatchStats()
        check(lib().xrp_batch_run(self.ptr, ctypes.byref(st)))
        return st.as_dict()

    def output_ptr(self, item: int) -> int:
        fp = ctypes.POINTER(ctypes.c_float)()
        check(lib().xrp_batch_output(self.ptr, item, ctypes.byref(fp)))
        return ctypes.cast(fp, ctypes.c_void_p).value

    def read(self, item: int) -> np.ndarray:
        out = np.empty(self.shapes[item], dtype=np.float32)
        check(lib().xrp_batch_read(self.ptr, item,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def latencies(self) -> np.ndarray:
        out = np.zeros(len(self.shapes), dtype=np.float64)
        check(lib().xrp_batch_latencies(self.ptr, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))))
        return out

    def close(self) -> None:
        if getattr(self, "ptr", None):
            lib().xrp_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SURROGATE_LEVELS = 16
SURROGATE_HIDDEN = 64


def split_surrogate_params(flat: np.ndarray, log2_table: int) -> Dict[str, np.ndarray]:
    """Flat surrogate parameters -> named arrays (tables[16][T][2], W1[32][64],
    b1[64], W2[64][64], b2[64], w3[64], b3)."""
    T = 1 << log2_table
    o = 0
    out = {}
    for name, shape in [("tables", (SURROGATE_LEVELS, T, 2)), ("W1", (32, 64)), ("b1", (64,)),
                        ("W2", (64, 64)), ("b2", (64,)), ("w3", (64,)), ("b3", ())]:
        n = int(np.prod(shape)) if shape else 1
        out[name] = flat[o:o + n].reshape(shape) if shape else flat[o]
        o += n
    assert o == flat.size
    return out


def join_surrogate_params(p: Dict[str, np.ndarray]) -> np.ndarray:
    return np.concatenate([np.asarray(p[k], dtype=np.float32).reshape(-1)
                           for k in ("tables", "W1", "b1", "W2", "b2", "w3", "b3")])


class Surrogate:
    """Learned density surrogate of a scene (xrp_surrogate_*)."""

    def __init__(self, ctx: Context, scene: Scene, log2_table: int = 19, n_min: int = 16,
                 n_max: int = 512, wrap: int = -1, seed: int = 1):
        c = SurrogateConfig()
        lib().xrp_surrogate_default_config(ctypes.byref(c))
        c.log2_table, c.n_min, c.n_max, c.wrap, c.seed = log2_table, n_min, n_max, wrap, seed
        ptr = ctypes.c_void_p()
        check(lib().xrp_surrogate_create(ctx.ptr, scene.ptr, ctypes.byref(c), ctypes.byref(ptr)))
        self.ptr, self.ctx, self.scene, self.log2_table = ptr, ctx, scene, log2_table
        ctx._owned.add(self)

    def train(self, samples: int = 1 << 24, batch: int = 65536, epochs: int = 4,
              lr_table: float = 1e-2, lr_mlp: float = 1e-3) -> Dict:
        c = TrainConfig()
        lib().xrp_train_default_config(ctypes.byref(c))
        c.samples, c.batch, c.epochs, c.lr_table, c.lr_mlp = samples, batch, epochs, lr_table, lr_mlp
        st = TrainStats()
        check(lib().xrp_surrogate_train(self.ptr, ctypes.byref(c), ctypes.byref(st)))
        return st.as_dict()

    def render_device(self, angles, resolution: int, ds: float, ds_surrogate: float,
                      tolerance: float, out_ptr: int, **kw) -> Dict:
        p = make_params(resolution, ds, **kw)
        st = SurrogateStats()
        check(lib().xrp_surrogate_render_device(self.ptr, ctypes.byref(p), _views(angles),
                                                len(angles), ds_surrogate, tolerance,
                                                ctypes.c_void_p(out_ptr), ctypes.byref(st)))
        return st.as_dict()

    def render(self, angles, resolution: int, ds: float, ds_surrogate: float,
               tolerance: float, **kw):
        """Images [view][i][j] (float32, host) and the render stats."""
        p = make_params(resolution, ds, **kw)
        st = SurrogateStats()
        out = np.empty((len(angles), resolution, resolution), dtype=np.float32)
        check(lib().xrp_surrogate_render(self.ptr, ctypes.byref(p), _views(angles),
                                         len(angles), ds_surrogate, tolerance,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         ctypes.byref(st)))
        return out, st.as_dict()

    def params(self) -> np.ndarray:
        n = lib().xrp_surrogate_param_count(self.ptr)
        out = np.empty(n, dtype=np.float32)
        check(lib().xrp_surrogate_get_params(self.ptr, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float))))
        return out

    def set_params(self, flat: np.ndarray) -> None:
        a = np.ascontiguousarray(flat, dtype=np.float32)
        assert a.size == lib().xrp_surrogate_param_count(self.ptr)
        check(lib().xrp_surrogate_set_params(self.ptr, a.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float))))

    def predict(self, u: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 3)
        out = np.empty(a.shape[0], dtype=np.float32)
        FPt = ctypes.POINTER(ctypes.c_float)
        check(lib().xrp_surrogate_predict(self.ptr, a.ctypes.data_as(FPt), a.shape[0],
                                          out.ctypes.data_as(FPt)))
        return out

    def gradients(self, u: np.ndarray, targets: np.ndarray):
        a = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        g = np.empty(lib().xrp_surrogate_param_count(self.ptr), dtype=np.float32)
        loss = ctypes.c_double()
        FPt = ctypes.POINTER(ctypes.c_float)
        check(lib().xrp_surrogate_gradients(self.ptr, a.ctypes.data_as(FPt),
                                            t.ctypes.data_as(FPt), a.shape[0],
                                            g.ctypes.data_as(FPt), ctypes.byref(loss)))
        return g, loss.value

    def close(self) -> None:
        if getattr(self, "ptr", None):
            lib().xrp_surrogate_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def diag_sincos_error(ctx: Context) -> Tuple[float, float]:
    """Largest absolute errors of the hardware sine and cosine over every FP32
    argument in [-0.5, 0.5] revolutions (xrp_diag_sincos_error)."""
    a, b = ctypes.c_double(), ctypes.c_double()
    check(lib().xrp_diag_sincos_error(ctx.ptr, ctypes.byref(a), ctypes.byref(b)))
    return float(a.value), float(b.value)


def diag_deform_error(ctx: Context) -> Tuple[float, float, float]:
    """Largest absolute errors of the deformations' hardware tanh, logistic
    function and e^-x over every FP32 argument of their ranges
    (xrp_diag_deform_error)."""
    a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    check(lib().xrp_diag_deform_error(ctx.ptr, ctypes.byref(a), ctypes.byref(b),
                                      ctypes.byref(c)))
    return float(a.value), float(b.value), float(c.value)


def diag_l2_bandwidth(ctx: Context) -> float:
    """Measured L2 read bandwidth of the device, GB/s (xrp_diag_l2_bandwidth)."""
    v = ctypes.c_double()
    check(lib().xrp_diag_l2_bandwidth(ctx.ptr, ctypes.byref(v)))
    return float(v.value)


def diag_mfma_bf16(ctx: Context, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    FPt = ctypes.POINTER(ctypes.c_float)
    A = np.ascontiguousarray(a, dtype=np.float32)
    B = np.ascontiguousarray(b, dtype=np.float32)
    D = np.empty((16, 16), dtype=np.float32)
    check(lib().xrp_diag_mfma_bf16(ctx.ptr, A.ctypes.data_as(FPt), B.ctypes.data_as(FPt),
                                   D.ctypes.data_as(FPt)))
    return D
