"""MI355X compute point-cloud rasteriser — Python mirror of the reference API.

The product is ``libscr.so`` (hand-written HIP kernels for gfx950 behind the C
ABI in ``include/scr.h``).  This module is the host-side mirror of the
reference's Swift interface (``ComputeRasteriser``,
``ComputeRasteriserPointCloud``, ``ComputeRasteriserConfiguration``,
``PackedPointCloud``, ``GPUPacker``) over ctypes, so that tests and the
benchmark read like the reference's own.  There is no CPU fallback: if the
library is missing or no GPU is present, constructing a ``Context`` raises.

Matrices are numpy ``float32`` arrays of shape (4, 4) indexed ``m[column, row]``
(the memory layout of ``simd_float4x4``); ``mat()`` converts from the usual
row-major math notation.
"""
from __future__ import annotations

import atexit
import ctypes as C
import dataclasses
import math
import weakref
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

_PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = _PKG_DIR / "libscr.so"

# ---------------------------------------------------------------- constants
THREADS_PER_GROUP = 128
DEFAULT_POINTS_PER_BATCH = 10240
MODE_HIGH_QUALITY_AVERAGE = 0
MODE_NEAREST_POINT = 1
POINT_SIZE_SCREEN_SPACE = 0
POINT_SIZE_WORLD_SPACE = 1

PLANE_DEPTH, PLANE_RED, PLANE_GREEN, PLANE_BLUE, PLANE_COUNT = 0, 1, 2, 3, 4
PLANE_NEAREST_DEPTH, PLANE_NEAREST_INDEX, PLANE_RGBA8, PLANE_DEPTH_F32 = 5, 6, 7, 8
PLANE_DIFF_SUM_R, PLANE_DIFF_SUM_G, PLANE_DIFF_SUM_B = 9, 10, 11
POOL_XYZ_LOW, POOL_XYZ_MED, POOL_XYZ_HIGH, POOL_COLORS, POOL_LEVELS, POOL_BATCHES = range(6)

ERRORS = {
    -1: "SCR_ERR_INVALID_ARGUMENT",
    -2: "SCR_ERR_NO_FREE_SLOTS",
    -3: "SCR_ERR_BATCH_TOO_LARGE",
    -4: "SCR_ERR_POINTS_PER_BATCH",
    -5: "SCR_ERR_HIP",
    -6: "SCR_ERR_OUT_OF_MEMORY",
    -7: "SCR_ERR_SOURCE_TOO_SMALL",
    -8: "SCR_ERR_UNSUPPORTED",
    -9: "SCR_ERR_SLOT_RANGE",
    -10: "SCR_ERR_NO_DEVICE",
    -11: "SCR_ERR_PLY_INVALID_HEADER",
    -12: "SCR_ERR_PLY_UNSUPPORTED_FORMAT",
    -13: "SCR_ERR_PLY_MISSING_VERTEX_ELEMENT",
    -14: "SCR_ERR_PLY_MISSING_PROPERTY",
    -15: "SCR_ERR_PLY_UNSUPPORTED_PROPERTY",
    -16: "SCR_ERR_PLY_TRUNCATED_BODY",
    -17: "SCR_ERR_PLY_INVALID_ASCII_VALUE",
}

# RasterBatch — Pipelines/ComputeRasteriserTypes.h:11-33 (64 B)
RASTER_BATCH_DTYPE = np.dtype(
    [
        ("state", "<i4"),
        ("minX", "<f4"), ("minY", "<f4"), ("minZ", "<f4"),
        ("maxX", "<f4"), ("maxY", "<f4"), ("maxZ", "<f4"),
        ("numPoints", "<u4"),
        ("firstPoint", "<u4"),
        ("fileIndex", "<u4"),
        ("lodCum01", "<u4"), ("lodCum23", "<u4"), ("lodCum45", "<u4"), ("lodCum67", "<u4"),
        ("padding7", "<u4"), ("padding8", "<u4"),
    ]
)
# VisibleBatch — ComputeRasteriserTypes.h:52-57 (16 B)
VISIBLE_BATCH_DTYPE = np.dtype(
    [("batchIndex", "<u4"), ("level", "<i4"), ("lodThreshold", "<f4"), ("activePoints", "<u4")]
)
# RasterFile — ComputeRasteriserTypes.h:35-40 (4 x float4x4, 256 B)
RASTER_FILE_DTYPE = np.dtype(
    [("transform", "<f4", (4, 4)), ("transformFrustum", "<f4", (4, 4)), ("world", "<f4", (4, 4)),
     ("prevTransform", "<f4", (4, 4))]
)
assert RASTER_BATCH_DTYPE.itemsize == 64 and VISIBLE_BATCH_DTYPE.itemsize == 16
assert RASTER_FILE_DTYPE.itemsize == 256


class ScrError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


# ---------------------------------------------------------------- C structs
class CConfig(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("depthTolerance", C.c_float), ("backgroundColor", C.c_float * 4),
        ("enableFrustumCulling", C.c_int32), ("lodBias", C.c_int32), ("enableCLOD", C.c_int32),
        ("enableLODDither", C.c_int32), ("colorizeChunks", C.c_int32), ("colorizeOverdraw", C.c_int32),
        ("pointSizeMode", C.c_int32), ("minimumPointSize", C.c_float), ("maximumPointSize", C.c_float),
        ("pointSizeScale", C.c_float), ("applyTint", C.c_int32), ("tintAlphaIsCoverage", C.c_int32),
        ("motionBlur", C.c_float), ("holeFillIterations", C.c_int32), ("motionBlurSamples", C.c_int32),
        ("motionBlurMaxSpread", C.c_float), ("applyDisplacement", C.c_int32), ("reserved", C.c_int32 * 1),
    ]


class CCompositeOptions(C.Structure):
    _fields_ = [("mode", C.c_int32), ("blend", C.c_int32), ("depthCompare", C.c_int32), ("writeDepth", C.c_int32)]


COMPOSITE_OVER, COMPOSITE_DEPTH = 0, 1
BLEND_REPLACE, BLEND_OVER = 0, 1
COMPARE_ALWAYS, COMPARE_GREATER, COMPARE_GREATER_EQUAL, COMPARE_LESS, COMPARE_LESS_EQUAL = range(5)


class CCamera(C.Structure):
    _fields_ = [("viewMatrix", C.c_float * 16), ("projectionMatrix", C.c_float * 16)]


class CCloudInfo(C.Structure):
    _fields_ = [
        ("maxResidentBatches", C.c_int32), ("pointsPerBatch", C.c_int32), ("residentBatchCount", C.c_int32),
        ("freeSlotCount", C.c_int32), ("pointCount", C.c_int64), ("nFiles", C.c_int32), ("hasTints", C.c_int32),
    ]


class CCloudView(C.Structure):
    """scr_cloud_view: device pointers of a cloud's pool for user kernels."""
    _fields_ = [("batches", C.c_void_p), ("xyzLow", C.c_void_p), ("xyzMed", C.c_void_p), ("xyzHigh", C.c_void_p),
                ("colors", C.c_void_p), ("levels", C.c_void_p), ("tints", C.c_void_p),
                ("pointsPerBatch", C.c_uint32), ("slots", C.c_uint32), ("stream", C.c_void_p),
                ("displacements", C.c_void_p)]


class CPackOptions(C.Structure):
    _fields_ = [("keyBits", C.c_int32), ("lodLevels", C.c_int32), ("coarseVoxelDivisions", C.c_int32),
                ("inputOnDevice", C.c_int32)]


class CSceneParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("halfExtent", C.c_float), ("numBoxes", C.c_int32), ("terrainFraction", C.c_float),
        ("amplitude", C.c_float), ("baseWavelength", C.c_float), ("octaves", C.c_int32), ("boxMinEdge", C.c_float),
        ("boxMaxEdge", C.c_float), ("colorNoise", C.c_float), ("quantizeColors", C.c_int32),
    ]


class CPassTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("totalMs", C.c_double), ("launches", C.c_int32)]


class CMVShape(C.Structure):
    _fields_ = [("threads", C.c_int32), ("pointsPerThread", C.c_int32), ("splits", C.c_int32),
                ("waveTiles", C.c_int32)]


class CMVTuning(C.Structure):
    _fields_ = [("depth", CMVShape), ("color", CMVShape), ("nearest", CMVShape)]


_VP = C.c_void_p
_I = C.c_int
_SZ = C.c_size_t
_SIGNATURES = {
    "scr_version": (C.c_char_p, []),
    "scr_config_default": (None, [C.POINTER(CConfig)]),
    "scr_ctx_create": (_I, [_I, C.POINTER(_VP)]),
    "scr_ctx_destroy": (_I, [_VP]),
    "scr_last_error": (C.c_char_p, [_VP]),
    "scr_sync": (_I, [_VP]),
    "scr_profile_enable": (_I, [_VP, _I]),
    "scr_profile_collect": (_I, [_VP, _VP, _I, C.POINTER(_I)]),
    "scr_device_alloc": (_I, [_VP, _SZ, C.POINTER(_VP)]),
    "scr_device_free": (_I, [_VP, _VP]),
    "scr_memcpy_h2d": (_I, [_VP, _VP, _VP, _SZ]),
    "scr_memcpy_d2h": (_I, [_VP, _VP, _VP, _SZ]),
    "scr_memcpy_d2d": (_I, [_VP, _VP, _VP, _SZ]),
    "scr_measure_atomic_rate": (_I, [_VP, _I, C.POINTER(C.c_double)]),
    "scr_check_fast_reciprocal": (_I, [_VP, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "scr_ply_parse": (_I, [_VP, _SZ, _SZ, _VP, _VP, C.POINTER(_SZ), C.c_char_p, _SZ]),
    "scr_cloud_create": (_I, [_VP, _I, _I, _VP, _I, C.POINTER(_VP)]),
    "scr_cloud_destroy": (_I, [_VP]),
    "scr_cloud_get_info": (_I, [_VP, C.POINTER(CCloudInfo)]),
    "scr_cloud_add_batches": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP, _I, _I, _VP]),
    "scr_cloud_remove_batches": (_I, [_VP, _VP, _I, _I]),
    "scr_cloud_commit": (_I, [_VP]),
    "scr_cloud_clear": (_I, [_VP]),
    "scr_cloud_set_model": (_I, [_VP, _VP]),
    "scr_cloud_update_files": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "scr_cloud_set_tints": (_I, [_VP, _VP, _SZ, _I]),
    "scr_cloud_set_displacements": (_I, [_VP, _VP, _SZ, _I]),
    "scr_cloud_device_view": (_I, [_VP, _I, _VP]),
    "scr_cloud_read_pool": (_I, [_VP, _I, _VP, _SZ]),
    "scr_cloud_read_visible": (_I, [_VP, _VP, _I, C.POINTER(_I)]),
    "scr_pack_options_default": (None, [C.POINTER(CPackOptions)]),
    "scr_pack": (_I, [_VP, _VP, _VP, _SZ, C.POINTER(CPackOptions), _VP]),
    "scr_pack_chunk": (_I, [_VP, _VP, _VP, _SZ, C.POINTER(CPackOptions), C.POINTER(C.c_int32)]),
    "scr_scene_params_default": (None, [C.POINTER(CSceneParams)]),
    "scr_gen_scene": (_I, [_VP, C.POINTER(CSceneParams), _SZ, _VP, _VP, _I]),
    "scr_scene_height": (C.c_float, [C.POINTER(CSceneParams), C.c_float, C.c_float]),
    "scr_frame_create": (_I, [_VP, _I, _I, C.POINTER(_VP)]),
    "scr_frame_destroy": (_I, [_VP]),
    "scr_frame_hole_fill": (_I, [_VP, _I]),
    "scr_composite_options_default": (None, [C.POINTER(CCompositeOptions)]),
    "scr_frame_composite": (_I, [_VP, C.POINTER(CCompositeOptions), _VP, _VP, _I]),
    "scr_frame_read": (_I, [_VP, _I, _VP, _SZ]),
    "scr_render": (_I, [_VP, C.POINTER(CConfig), C.POINTER(CCamera), _VP, _I, _VP]),
    "scr_render_camera": (_I, [_VP, C.POINTER(CConfig), C.POINTER(CCamera), C.c_uint64, _VP, _I, _VP]),
    "scr_pick": (_I, [_VP, C.POINTER(CConfig), C.POINTER(CCamera), _VP, _VP, C.c_float, C.c_float, _I,
                      C.POINTER(C.c_uint32), C.POINTER(_I)]),
    "scr_diff_render": (_I, [_VP, C.POINTER(CConfig), C.POINTER(CCamera), _VP, _VP, _VP, _VP]),
    "scr_diff_backward": (_I, [_VP, C.POINTER(CConfig), C.POINTER(CCamera), _VP, _VP, _VP, _VP]),
    "scr_multiview_create": (_I, [_VP, _VP, _I, _VP, _VP, _I, C.POINTER(_VP)]),
    "scr_multiview_destroy": (_I, [_VP]),
    "scr_multiview_render": (_I, [_VP, C.POINTER(CConfig), _VP, _I]),
    "scr_multiview_get_tuning": (_I, [_VP, C.POINTER(CMVTuning)]),
    "scr_multiview_set_tuning": (_I, [_VP, C.POINTER(CMVTuning)]),
    "scr_multiview_autotune": (_I, [_VP, C.POINTER(CConfig), _VP, _I, C.POINTER(CMVTuning)]),
    "scr_multiview_diff_render": (_I, [_VP, C.POINTER(CConfig), _VP, _VP, _VP]),
    "scr_multiview_diff_backward": (_I, [_VP, C.POINTER(CConfig), _VP, _VP, _VP]),
}

_lib: Optional[C.CDLL] = None


def load_library(path: Optional[Path] = None) -> C.CDLL:
    """Load libscr.so (or the library at ``path``, which then serves every
    object created afterwards) and bind every entry point; raises if it is
    missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------- math helpers
def mat(rows) -> np.ndarray:
    """Row-major 4x4 (math notation) → column-major storage m[column, row]."""
    return np.ascontiguousarray(np.asarray(rows, dtype=np.float64).T, dtype=np.float32)


IDENTITY = np.eye(4, dtype=np.float32)


def perspective_reverse_z(fovy_degrees: float, aspect: float, near: float, far: float) -> np.ndarray:
    """Right-handed reverse-Z perspective (view z = -near → ndc 1, z = -far → 0), column-major."""
    f = 1.0 / math.tan(math.radians(fovy_degrees) * 0.5)
    return mat([
        [f / aspect, 0, 0, 0],
        [0, f, 0, 0],
        [0, 0, near / (far - near), near * far / (far - near)],
        [0, 0, -1, 0],
    ])


def orthographic_reverse_z(height: float, aspect: float, near: float, far: float) -> np.ndarray:
    """Right-handed reverse-Z orthographic projection of a view volume
    `height` high and `height * aspect` wide (view z = -near → ndc 1,
    z = -far → 0; w = 1), column-major."""
    h = 2.0 / height
    return mat([
        [h / aspect, 0, 0, 0],
        [0, h, 0, 0],
        [0, 0, 1.0 / (far - near), far / (far - near)],
        [0, 0, 0, 1],
    ])


def look_at(eye, target, up) -> np.ndarray:
    """Right-handed view matrix looking down -Z, column-major."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    u = np.cross(right, fwd)
    return mat([
        [right[0], right[1], right[2], -right @ eye],
        [u[0], u[1], u[2], -u @ eye],
        [-fwd[0], -fwd[1], -fwd[2], fwd @ eye],
        [0, 0, 0, 1],
    ])


@dataclasses.dataclass
class Camera:
    """A camera: view and projection matrices (column-major float32 4x4)."""
    viewMatrix: np.ndarray
    projectionMatrix: np.ndarray

    @staticmethod
    def perspective(eye, target, up=(0.0, 1.0, 0.0), fov: float = 60.0, aspect: float = 1.0,
                    near: float = 0.1, far: float = 100.0) -> "Camera":
        return Camera(look_at(eye, target, up), perspective_reverse_z(fov, aspect, near, far))

    @staticmethod
    def orthographic(eye, target, up=(0.0, 1.0, 0.0), height: float = 2.0, aspect: float = 1.0,
                     near: float = 0.1, far: float = 100.0) -> "Camera":
        return Camera(look_at(eye, target, up), orthographic_reverse_z(height, aspect, near, far))

    def to_c(self) -> CCamera:
        c = CCamera()
        C.memmove(C.addressof(c.viewMatrix), np.ascontiguousarray(self.viewMatrix, np.float32).ctypes.data, 64)
        C.memmove(C.addressof(c.projectionMatrix), np.ascontiguousarray(self.projectionMatrix, np.float32).ctypes.data, 64)
        return c


# ---------------------------------------------------------------- configuration
@dataclasses.dataclass
class ComputeRasteriserConfiguration:
    """ComputeRasteriserTypes.swift:26-142 (defaults :97-119)."""
    mode: int = MODE_HIGH_QUALITY_AVERAGE
    depthTolerance: float = 0.01
    backgroundColor: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    enableFrustumCulling: bool = True
    lodBias: int = 0
    enableCLOD: bool = True
    enableLODDither: bool = True
    colorizeChunks: bool = False
    colorizeOverdraw: bool = False
    pointSizeMode: int = POINT_SIZE_SCREEN_SPACE
    minimumPointSize: float = 1.0
    maximumPointSize: float = 1.0
    pointSizeScale: float = 1.0
    applyTint: bool = False
    tintAlphaIsCoverage: bool = False
    motionBlur: float = 0.0
    holeFillIterations: int = 0
    motionBlurSamples: int = 8
    motionBlurMaxSpread: float = 64.0
    applyDisplacement: bool = False

    def to_c(self) -> CConfig:
        c = CConfig()
        c.mode = int(self.mode)
        c.depthTolerance = float(self.depthTolerance)
        for i in range(4):
            c.backgroundColor[i] = float(self.backgroundColor[i])
        c.enableFrustumCulling = int(bool(self.enableFrustumCulling))
        c.lodBias = int(self.lodBias)
        c.enableCLOD = int(bool(self.enableCLOD))
        c.enableLODDither = int(bool(self.enableLODDither))
        c.colorizeChunks = int(bool(self.colorizeChunks))
        c.colorizeOverdraw = int(bool(self.colorizeOverdraw))
        c.pointSizeMode = int(self.pointSizeMode)
        c.minimumPointSize = float(self.minimumPointSize)
        c.maximumPointSize = float(self.maximumPointSize)
        c.pointSizeScale = float(self.pointSizeScale)
        c.applyTint = int(bool(self.applyTint))
        c.tintAlphaIsCoverage = int(bool(self.tintAlphaIsCoverage))
        c.motionBlur = float(self.motionBlur)
        c.holeFillIterations = int(self.holeFillIterations)
        c.motionBlurSamples = int(self.motionBlurSamples)
        c.motionBlurMaxSpread = float(self.motionBlurMaxSpread)
        c.applyDisplacement = int(bool(self.applyDisplacement))
        return c


# ---------------------------------------------------------------- packed model
def identity_files(n: int = 1) -> np.ndarray:
    f = np.zeros(n, RASTER_FILE_DTYPE)
    for k in ("transform", "transformFrustum", "world", "prevTransform"):
        f[k] = IDENTITY
    return f


@dataclasses.dataclass
class PackedPointCloud:
    """ComputeRasteriserTypes.swift:269-320 (SoA point arrays + batch table)."""
    batches: np.ndarray          # RASTER_BATCH_DTYPE
    files: np.ndarray            # RASTER_FILE_DTYPE
    xyzLow: np.ndarray           # uint32
    xyzMed: np.ndarray
    xyzHigh: np.ndarray
    colors: np.ndarray           # uint32 RGBA8
    levels: np.ndarray           # uint8
    boundsMin: np.ndarray
    boundsMax: np.ndarray
    orderedPositions: Optional[np.ndarray] = None
    sourceIndices: Optional[np.ndarray] = None

    @property
    def pointCount(self) -> int:
        return int(self.colors.shape[0])

    @property
    def batchCount(self) -> int:
        return int(self.batches.shape[0])


def lod_cumulative_counts(batch) -> list:
    """RasterBatch.lodCumulativeCounts — ComputeRasteriserTypes.swift:196-203."""
    words = [int(batch["lodCum01"]), int(batch["lodCum23"]), int(batch["lodCum45"]), int(batch["lodCum67"])]
    out = []
    for w in words:
        out += [w & 0xFFFF, (w >> 16) & 0xFFFF]
    return out


def set_lod_cumulative_counts(batch, counts) -> None:
    """RasterBatch.setLODCumulativeCounts — ComputeRasteriserTypes.swift:214-223."""
    if len(counts) != 8:
        raise ValueError(f"expected 8 cumulative level counts, got {len(counts)}")
    for c in counts:
        if not 0 <= int(c) <= 65535:
            raise ValueError(f"cumulative count {c} does not fit uint16")
    batch["lodCum01"] = int(counts[0]) | (int(counts[1]) << 16)
    batch["lodCum23"] = int(counts[2]) | (int(counts[3]) << 16)
    batch["lodCum45"] = int(counts[4]) | (int(counts[5]) << 16)
    batch["lodCum67"] = int(counts[6]) | (int(counts[7]) << 16)


# ---------------------------------------------------------------- runtime objects
_LIVE_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_live_contexts() -> None:
    """Release every context's objects while the interpreter and the HIP
    runtime are both still intact (renderers first, then frames and clouds)."""
    for ctx in list(_LIVE_CONTEXTS):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    """One GPU, one HIP stream (scr_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.scr_ctx_create(device, C.byref(h))
        if rc != 0:
            raise ScrError(rc, "scr_ctx_create failed (no usable GPU?)")
        self.handle = h
        # objects holding device resources of this context; closed before it
        self._owned = weakref.WeakSet()
        _LIVE_CONTEXTS.add(self)

    def _adopt(self, obj) -> None:
        self._owned.add(obj)

    def check(self, rc: int) -> None:
        if rc != 0:
            raise ScrError(rc, (self.lib.scr_last_error(self.handle) or b"").decode())

    def sync(self) -> None:
        self.check(self.lib.scr_sync(self.handle))

    def measure_atomic_rate(self, add: bool = False) -> float:
        """Global atomics per second (atomicMax, or atomicAdd with add=True) in
        the pattern of the draw passes' tile flushes."""
        out = C.c_double(0.0)
        self.check(self.lib.scr_measure_atomic_rate(self.handle, 1 if add else 0, C.byref(out)))
        return float(out.value)

    def check_fast_reciprocal(self) -> tuple:
        """(mismatches, smallest mismatching float's bits) of the draw passes'
        three-instruction reciprocal against 1.0f / w over [2^-126, 2^126)."""
        n, first = C.c_uint64(0), C.c_uint32(0)
        self.check(self.lib.scr_check_fast_reciprocal(self.handle, C.byref(n), C.byref(first)))
        return int(n.value), int(first.value)

    def profile(self, enable: bool = True) -> None:
        """Bracket every kernel launch with HIP events (per-pass timing)."""
        self.check(self.lib.scr_profile_enable(self.handle, int(enable)))

    def collect_profile(self) -> dict:
        """{pass name: (total ms, launches)} since the last collect."""
        buf = (CPassTime * 64)()
        n = C.c_int()
        self.check(self.lib.scr_profile_collect(self.handle, buf, 64, C.byref(n)))
        return {buf[i].name.decode(): (buf[i].totalMs, buf[i].launches) for i in range(min(n.value, 64))}

    def close(self) -> None:
        if getattr(self, "handle", None):
            # multi-view renderers refer to frames and clouds: release them first
            owned = sorted(list(getattr(self, "_owned", ())), key=lambda o: getattr(o, "_close_order", 1))
            for o in owned:
                o.close()
            self.lib.scr_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Raw device allocation owned by a Context (inputs kept resident in HBM)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx.check(ctx.lib.scr_device_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value
        ctx._adopt(self)

    def close(self) -> None:
        self.free()

    @classmethod
    def from_array(cls, ctx: Context, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.nbytes)
        ctx.check(ctx.lib.scr_memcpy_h2d(ctx.handle, b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        self.ctx.check(self.ctx.lib.scr_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            self.ctx.lib.scr_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ComputeRasteriserPointCloud:
    """ComputeRasteriserPointCloud.swift: a slot pool of packed points on the GPU."""

    def __init__(self, ctx: Context, maxResidentBatches: int, pointsPerBatch: int = DEFAULT_POINTS_PER_BATCH,
                 files: Optional[np.ndarray] = None):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        f = None if files is None else np.ascontiguousarray(files, RASTER_FILE_DTYPE)
        ctx.check(self.lib.scr_cloud_create(ctx.handle, int(maxResidentBatches), int(pointsPerBatch), _ptr(f),
                                            0 if f is None else len(f), C.byref(h)))
        self.handle = h
        self.worldMatrix = IDENTITY.copy()
        ctx._adopt(self)

    @classmethod
    def from_packed(cls, ctx: Context, packed: PackedPointCloud) -> "ComputeRasteriserPointCloud":
        """init(context:packed:) — ComputeRasteriserPointCloud.swift:222-236."""
        nums = packed.batches["numPoints"] if packed.batchCount else np.array([DEFAULT_POINTS_PER_BATCH])
        ppb = max(1, int(nums.max()))
        cloud = cls(ctx, max(1, packed.batchCount), ppb, packed.files)
        if packed.batchCount > 0:
            cloud.addBatches(packed.xyzLow, packed.xyzMed, packed.xyzHigh, packed.colors, packed.levels,
                             packed.batches)
        return cloud

    # --- info
    def info(self) -> CCloudInfo:
        i = CCloudInfo()
        self.ctx.check(self.lib.scr_cloud_get_info(self.handle, C.byref(i)))
        return i

    @property
    def capacity(self):
        i = self.info()
        return (i.maxResidentBatches, i.pointsPerBatch)

    @property
    def residentBatchCount(self) -> int:
        return self.info().residentBatchCount

    @property
    def pointCount(self) -> int:
        return self.info().pointCount

    @property
    def freeSlotCount(self) -> int:
        return self.info().freeSlotCount

    @property
    def batchCount(self) -> int:
        return self.info().maxResidentBatches

    @property
    def pointsPerBatch(self) -> int:
        return self.info().pointsPerBatch

    # --- slot pool
    def addBatches(self, xyzLow, xyzMed, xyzHigh, colors, levels, batches, commit: bool = True) -> list:
        arrs = [np.ascontiguousarray(a, dt) for a, dt in
                ((xyzLow, np.uint32), (xyzMed, np.uint32), (xyzHigh, np.uint32), (colors, np.uint32),
                 (levels, np.uint8))]
        b = np.ascontiguousarray(batches, RASTER_BATCH_DTYPE)
        n = min(len(a) for a in arrs)
        slots = np.zeros(len(b), np.int32)
        self.ctx.check(self.lib.scr_cloud_add_batches(self.handle, *[a.ctypes.data for a in arrs], n, b.ctypes.data,
                                                      len(b), int(commit), slots.ctypes.data))
        return [int(s) for s in slots]

    def removeBatches(self, slots, commit: bool = True) -> None:
        s = np.ascontiguousarray(slots, np.int32)
        self.ctx.check(self.lib.scr_cloud_remove_batches(self.handle, s.ctypes.data, len(s), int(commit)))

    def commitBatchUpdates(self) -> None:
        self.ctx.check(self.lib.scr_cloud_commit(self.handle))

    def clearAllBatches(self) -> None:
        self.ctx.check(self.lib.scr_cloud_clear(self.handle))

    def updateFiles(self, viewProjection, modelMatrix, frustumTransform=None, prevViewProjection=None) -> None:
        vp = np.ascontiguousarray(viewProjection, np.float32)
        m = np.ascontiguousarray(modelMatrix, np.float32)
        fr = None if frustumTransform is None else np.ascontiguousarray(frustumTransform, np.float32)
        pv = None if prevViewProjection is None else np.ascontiguousarray(prevViewProjection, np.float32)
        self.ctx.check(self.lib.scr_cloud_update_files(self.handle, vp.ctypes.data, m.ctypes.data, _ptr(fr), _ptr(pv)))

    def setTints(self, rgba: Optional[np.ndarray]) -> None:
        """Install the TintPass output (float4 per pool point), or drop it with None."""
        if rgba is None:
            self.ctx.check(self.lib.scr_cloud_set_tints(self.handle, None, 0, 0))
            return
        t = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
        self.ctx.check(self.lib.scr_cloud_set_tints(self.handle, t.ctypes.data, t.shape[0], 0))

    def setTintsDevice(self, ptr: int, nPoints: int) -> None:
        self.ctx.check(self.lib.scr_cloud_set_tints(self.handle, ptr, nPoints, 1))

    def setDisplacements(self, xyz: Optional[np.ndarray]) -> None:
        """displacementBuffer (ComputeRasteriserPointCloud.swift:91-94): one
        displacement per pool point, (n, 3) or (n, 4) (the 16-byte stride of a
        Metal float3 array), or None to drop it."""
        if xyz is None:
            self.ctx.check(self.lib.scr_cloud_set_displacements(self.handle, None, 0, 0))
            return
        d = np.zeros((np.asarray(xyz).reshape(len(xyz), -1).shape[0], 4), np.float32)
        d[:, :3] = np.asarray(xyz, np.float32).reshape(len(xyz), -1)[:, :3]
        self.ctx.check(self.lib.scr_cloud_set_displacements(self.handle, d.ctypes.data, d.shape[0], 0))

    def deviceView(self, withTints: bool = False, withDisplacements: bool = False) -> CCloudView:
        """The pool's device pointers and the context stream for a user HIP
        kernel (scr_device.h) — what TintPass / DisplacementPass bind for their
        user kernels (TintPass.swift:215-338, DisplacementPass.swift:120-172);
        withTints / withDisplacements make sure those buffers exist."""
        v = CCloudView()
        flags = (1 if withTints else 0) | (2 if withDisplacements else 0)
        self.ctx.check(self.lib.scr_cloud_device_view(self.handle, flags, C.byref(v)))
        return v

    def readPool(self, which: int) -> np.ndarray:
        mrb, ppb = self.capacity
        n = mrb * ppb
        if which == POOL_BATCHES:
            out = np.zeros(mrb, RASTER_BATCH_DTYPE)
        elif which == POOL_LEVELS:
            out = np.zeros(n, np.uint8)
        else:
            out = np.zeros(n, np.uint32)
        self.ctx.check(self.lib.scr_cloud_read_pool(self.handle, which, out.ctypes.data, out.nbytes))
        return out

    def readVisible(self) -> np.ndarray:
        cap = self.batchCount
        out = np.zeros(cap, VISIBLE_BATCH_DTYPE)
        n = C.c_int()
        self.ctx.check(self.lib.scr_cloud_read_visible(self.handle, out.ctypes.data, cap, C.byref(n)))
        return out[: n.value]

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.scr_cloud_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Frame:
    """Per-viewport GPU resources (HQS accumulators, nearest winner, output textures)."""

    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx = ctx
        self.width, self.height = int(width), int(height)
        h = C.c_void_p()
        ctx.check(ctx.lib.scr_frame_create(ctx.handle, self.width, self.height, C.byref(h)))
        self.handle = h
        ctx._adopt(self)

    def holeFill(self, iterations: int = 3) -> None:
        """HoleFillProcessor.update(_:iterations:) over the RGBA8 output, in place."""
        self.ctx.check(self.ctx.lib.scr_frame_hole_fill(self.handle, int(iterations)))

    def composite(self, targetRGBA8: int, targetDepth: Optional[int] = None, pitch: Optional[int] = None,
                  mode: int = COMPOSITE_DEPTH, blend: int = BLEND_OVER, depthCompare: int = COMPARE_GREATER,
                  writeDepth: bool = False) -> None:
        """ComputeRasteriser.draw (ComputeRasteriser.swift:672-683, Post.metal:23-50):
        the frame's colour and depth onto a caller's device target (uint32
        RGBA8 texels and float32 depths, ``pitch`` texels a row), enqueued on
        the context stream; order the target's producers and consumers
        against it with sync()."""
        o = CCompositeOptions(int(mode), int(blend), int(depthCompare), int(bool(writeDepth)))
        self.ctx.check(self.ctx.lib.scr_frame_composite(self.handle, C.byref(o), targetRGBA8, targetDepth,
                                                        self.width if pitch is None else int(pitch)))

    def read(self, which: int) -> np.ndarray:
        dt = np.float32 if which == PLANE_DEPTH_F32 else (
            np.int64 if PLANE_DIFF_SUM_R <= which <= PLANE_DIFF_SUM_B else np.uint32)
        out = np.zeros(self.width * self.height, dt)
        self.ctx.check(self.ctx.lib.scr_frame_read(self.handle, which, out.ctypes.data, out.nbytes))
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.ctx.lib.scr_frame_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ComputeRasteriser:
    """ComputeRasteriser.swift: configuration, clouds, resize, update+encode, pick."""

    def __init__(self, ctx: Context, configuration: Optional[ComputeRasteriserConfiguration] = None):
        self.ctx = ctx
        self.configuration = configuration or ComputeRasteriserConfiguration()
        self.pointClouds: list = []
        self.frame: Optional[Frame] = None

    def addPointCloud(self, cloud: ComputeRasteriserPointCloud) -> ComputeRasteriserPointCloud:
        self.pointClouds.append(cloud)
        return cloud

    def removePointCloud(self, cloud: ComputeRasteriserPointCloud) -> None:
        self.pointClouds.remove(cloud)

    def resize(self, width: int, height: int) -> None:
        if self.frame is None or (self.frame.width, self.frame.height) != (int(width), int(height)):
            self.frame = Frame(self.ctx, width, height)

    def render(self, camera: Camera, cameraKey: int = 0) -> None:
        """update(renderContext:camera:viewport:index:) + encode(_:) for one
        frame.  Motion blur sweeps from the previous frame of the camera
        `cameraKey` names (the reference keys it by the camera object, so the
        eyes of a stereo pair pass different keys)."""
        if self.frame is None:
            raise RuntimeError("resize() first")
        for cl in self.pointClouds:
            m = np.ascontiguousarray(cl.worldMatrix, np.float32)
            self.ctx.check(self.ctx.lib.scr_cloud_set_model(cl.handle, m.ctypes.data))
        handles = (C.c_void_p * max(1, len(self.pointClouds)))(*[c.handle.value for c in self.pointClouds])
        cfg = self.configuration.to_c()
        cam = camera.to_c()
        self.ctx.check(self.ctx.lib.scr_render_camera(self.ctx.handle, C.byref(cfg), C.byref(cam), int(cameraKey),
                                                      handles, len(self.pointClouds), self.frame.handle))

    def read(self, which: int) -> np.ndarray:
        return self.frame.read(which).reshape(self.frame.height, self.frame.width)

    def outputRGBA8(self) -> np.ndarray:
        """(H, W, 4) uint8 image (row 0 = top of the viewport)."""
        return self.read(PLANE_RGBA8).view(np.uint8).reshape(self.frame.height, self.frame.width, 4)

    def pickPointIndex(self, ndc, cloud: ComputeRasteriserPointCloud, camera: Camera,
                       searchRadius: int = 10) -> Optional[int]:
        """pickPointIndex(atNDC:in:camera:searchRadius:) — ComputeRasteriser.swift:609-670."""
        if self.frame is None:
            return None
        m = np.ascontiguousarray(cloud.worldMatrix, np.float32)
        self.ctx.check(self.ctx.lib.scr_cloud_set_model(cloud.handle, m.ctypes.data))
        cfg = self.configuration.to_c()
        cam = camera.to_c()
        out = C.c_uint32()
        found = C.c_int()
        self.ctx.check(self.ctx.lib.scr_pick(self.ctx.handle, C.byref(cfg), C.byref(cam), cloud.handle,
                                             self.frame.handle, float(ndc[0]), float(ndc[1]), int(searchRadius),
                                             C.byref(out), C.byref(found)))
        return int(out.value) if found.value else None


# ---------------------------------------------------------------- differentiable HQS
class DifferentiableHQS:
    """Differentiable HQS over one cloud (scr_diff_render / scr_diff_backward).

    One frame per view keeps the forward's depth / count planes for the
    backward.  batched=True renders and differentiates all views with one
    launch per phase (scr_multiview_diff_*); the results are the same.
    Pointers are device addresses: params float4 per pool point, images 3
    floats per pixel, dLdImage likewise, grad 4 doubles per pool point
    (accumulated)."""

    def __init__(self, ctx: Context, cloud: "ComputeRasteriserPointCloud", sizes,
                 configuration: Optional[ComputeRasteriserConfiguration] = None, batched: bool = False):
        self.ctx = ctx
        self.cloud = cloud
        self.configuration = configuration or ComputeRasteriserConfiguration()
        self._mv = None
        if batched:
            self._mv = MultiViewRasteriser(ctx, [cloud], sizes, [0] * len(sizes), self.configuration)
            self.frames = self._mv.frames
        else:
            self.frames = [Frame(ctx, w, h) for (w, h) in sizes]

    @property
    def batched(self) -> bool:
        return self._mv is not None

    @property
    def poolPoints(self) -> int:
        return self.cloud.batchCount * self.cloud.pointsPerBatch

    def _prep(self, camera: Camera):
        m = np.ascontiguousarray(self.cloud.worldMatrix, np.float32)
        self.ctx.check(self.ctx.lib.scr_cloud_set_model(self.cloud.handle, m.ctypes.data))
        return self.configuration.to_c(), camera.to_c()

    def forward(self, view: int, camera: Camera, params_ptr: int, image_ptr: int) -> None:
        cfg, cam = self._prep(camera)
        self.ctx.check(self.ctx.lib.scr_diff_render(self.ctx.handle, C.byref(cfg), C.byref(cam), self.cloud.handle,
                                                    params_ptr, self.frames[view].handle, image_ptr))

    def backward(self, view: int, camera: Camera, dLdImage_ptr: int, grad_ptr: int) -> None:
        cfg, cam = self._prep(camera)
        self.ctx.check(self.ctx.lib.scr_diff_backward(self.ctx.handle, C.byref(cfg), C.byref(cam),
                                                      self.cloud.handle, self.frames[view].handle, dLdImage_ptr,
                                                      grad_ptr))

    def forward_all(self, cameras, params_ptr: int, image_ptrs) -> None:
        if self._mv is not None:
            self._mv.configuration = self.configuration
            self._mv.diffRender(cameras, params_ptr, image_ptrs)
        else:
            for v, cam in enumerate(cameras):
                self.forward(v, cam, params_ptr, image_ptrs[v])

    def backward_all(self, cameras, dLdImage_ptrs, grad_ptr: int) -> None:
        """Views add to grad in view order."""
        if self._mv is not None:
            self._mv.configuration = self.configuration
            self._mv.diffBackward(cameras, dLdImage_ptrs, grad_ptr)
        else:
            for v, cam in enumerate(cameras):
                self.backward(v, cam, dLdImage_ptrs[v], grad_ptr)

    def read(self, view: int, which: int) -> np.ndarray:
        f = self.frames[view]
        return f.read(which).reshape(f.height, f.width)

    def sums(self, view: int) -> np.ndarray:
        """(H, W, 3) int64 fixed-point sums of the last forward of this view."""
        return np.stack([self.read(view, PLANE_DIFF_SUM_R + k) for k in range(3)], axis=-1)

    def close(self) -> None:
        if self._mv is not None:
            self._mv.close()
            self._mv = None
        else:
            for f in getattr(self, "frames", []):
                f.close()


# ---------------------------------------------------------------- multi-view
MV_RESOLUTIONS = ((256, 256), (512, 384), (1024, 768), (1920, 1080))


def sphere_views(centre, radius: float, count: int, seed: int, resolutions=MV_RESOLUTIONS,
                 fov: float = 60.0, near_frac: float = 0.01, far_frac: float = 4.0, up=(0.0, 0.0, 1.0)):
    """``count`` cameras with eyes uniform on the sphere of ``radius`` around
    ``centre`` looking at it, each with a resolution drawn uniformly from
    ``resolutions``.  Returns (cameras, sizes).  Deterministic in ``seed``."""
    rng = np.random.default_rng(seed)
    centre = np.asarray(centre, np.float64)
    cams, sizes = [], []
    for _ in range(count):
        w, h = resolutions[int(rng.integers(len(resolutions)))]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        eye = centre + radius * d
        u = np.asarray(up, np.float64)
        if abs(float(d @ u)) > 0.99:       # looking along the up axis
            u = np.array([0.0, 1.0, 0.0]) if abs(u[1]) < 0.5 else np.array([1.0, 0.0, 0.0])
        cams.append(Camera.perspective(tuple(eye), tuple(centre), tuple(u), fov=fov, aspect=w / h,
                                       near=radius * near_frac, far=radius * far_frac))
        sizes.append((w, h))
    return cams, sizes


MV_SHAPES = ((128, 8), (256, 4), (256, 8), (256, 16), (512, 4), (512, 8))
MV_SPLITS = (1, 2, 4)


def _tuning_to_dict(t: CMVTuning) -> dict:
    return {ph: {"threads": getattr(t, ph).threads, "pointsPerThread": getattr(t, ph).pointsPerThread,
                 "splits": getattr(t, ph).splits, "waveTiles": getattr(t, ph).waveTiles}
            for ph in ("depth", "color", "nearest")}


class MultiViewRasteriser:
    """Renders many cameras in one batched launch per phase (scr_multiview_*).

    View v renders ``clouds[viewClouds[v]]`` into its own frame of size
    ``sizes[v]``; the planes are those ComputeRasteriser.render would give
    that camera alone (order-independent integer accumulation)."""

    def __init__(self, ctx: Context, clouds, sizes, viewClouds,
                 configuration: Optional[ComputeRasteriserConfiguration] = None):
        self.ctx = ctx
        self.clouds = list(clouds)
        self.frames = [Frame(ctx, w, h) for (w, h) in sizes]
        self.viewClouds = [int(i) for i in viewClouds]
        if len(self.viewClouds) != len(self.frames):
            raise ValueError("one cloud index per view")
        self.configuration = configuration or ComputeRasteriserConfiguration()
        ch = (C.c_void_p * len(self.clouds))(*[c.handle.value for c in self.clouds])
        fh = (C.c_void_p * len(self.frames))(*[f.handle.value for f in self.frames])
        vc = np.ascontiguousarray(self.viewClouds, np.int32)
        h = C.c_void_p()
        ctx.check(ctx.lib.scr_multiview_create(ctx.handle, ch, len(self.clouds), fh, vc.ctypes.data,
                                               len(self.frames), C.byref(h)))
        self.handle = h
        ctx._adopt(self)

    _close_order = 0

    @property
    def nViews(self) -> int:
        return len(self.frames)

    def _cams(self, cameras):
        if len(cameras) != self.nViews:
            raise ValueError("one camera per view")
        for cl in self.clouds:
            m = np.ascontiguousarray(cl.worldMatrix, np.float32)
            self.ctx.check(self.ctx.lib.scr_cloud_set_model(cl.handle, m.ctypes.data))
        arr = (CCamera * self.nViews)()
        for i, cam in enumerate(cameras):
            arr[i] = cam.to_c()
        return arr

    def render(self, cameras, useGraph: bool = True) -> None:
        cams = self._cams(cameras)
        cfg = self.configuration.to_c()
        self.ctx.check(self.ctx.lib.scr_multiview_render(self.handle, C.byref(cfg), cams, 1 if useGraph else 0))

    def autotune(self, cameras, reps: int = 2) -> dict:
        cams = self._cams(cameras)
        cfg = self.configuration.to_c()
        t = CMVTuning()
        self.ctx.check(self.ctx.lib.scr_multiview_autotune(self.handle, C.byref(cfg), cams, int(reps), C.byref(t)))
        return _tuning_to_dict(t)

    @property
    def tuning(self) -> dict:
        t = CMVTuning()
        self.ctx.check(self.ctx.lib.scr_multiview_get_tuning(self.handle, C.byref(t)))
        return _tuning_to_dict(t)

    @tuning.setter
    def tuning(self, d: dict) -> None:
        t = CMVTuning()
        self.ctx.check(self.ctx.lib.scr_multiview_get_tuning(self.handle, C.byref(t)))
        for ph, sh in d.items():
            dst = getattr(t, ph)
            dst.threads = int(sh.get("threads", dst.threads))
            dst.pointsPerThread = int(sh.get("pointsPerThread", dst.pointsPerThread))
            dst.splits = int(sh.get("splits", dst.splits))
            dst.waveTiles = int(sh.get("waveTiles", dst.waveTiles))
        self.ctx.check(self.ctx.lib.scr_multiview_set_tuning(self.handle, C.byref(t)))

    def read(self, view: int, which: int) -> np.ndarray:
        f = self.frames[view]
        return f.read(which).reshape(f.height, f.width)

    def diffRender(self, cameras, params_ptr: int, image_ptrs) -> None:
        """scr_multiview_diff_render: image_ptrs[v] = device address of view v's image."""
        cams = self._cams(cameras)
        cfg = self.configuration.to_c()
        imgs = (C.c_void_p * self.nViews)(*[int(p) for p in image_ptrs])
        self.ctx.check(self.ctx.lib.scr_multiview_diff_render(self.handle, C.byref(cfg), cams, params_ptr, imgs))

    def diffBackward(self, cameras, dLdImage_ptrs, grad_ptr: int) -> None:
        """scr_multiview_diff_backward: grad (4 doubles per pool point) += all views' gradients."""
        cams = self._cams(cameras)
        cfg = self.configuration.to_c()
        dls = (C.c_void_p * self.nViews)(*[int(p) for p in dLdImage_ptrs])
        self.ctx.check(self.ctx.lib.scr_multiview_diff_backward(self.handle, C.byref(cfg), cams, dls, grad_ptr))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.ctx.lib.scr_multiview_destroy(self.handle)
            self.handle = None
        for f in getattr(self, "frames", []):
            f.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- packer
def _float4(a: np.ndarray, fill: float) -> np.ndarray:
    a = np.asarray(a, np.float32)
    if a.ndim == 2 and a.shape[1] == 4:
        return np.ascontiguousarray(a)
    out = np.full((a.shape[0], 4), fill, np.float32)
    out[:, : a.shape[1]] = a
    return out


class GPUPacker:
    """GPUPacker.swift: Morton radix-sort packer on the GPU.

    keyBits 30 sorts on the reference's 10-bit-per-axis Morton keys, 63 on
    21-bit-per-axis keys.  Equal keys keep input order (stable LSD sort).
    """

    def __init__(self, ctx: Context, lodLevels: int = 4, coarseVoxelDivisions: int = 64, keyBits: int = 30):
        self.ctx = ctx
        self.lodLevels = lodLevels
        self.coarseVoxelDivisions = coarseVoxelDivisions
        self.keyBits = keyBits

    def _opts(self, on_device: bool) -> CPackOptions:
        o = CPackOptions()
        o.keyBits, o.lodLevels, o.coarseVoxelDivisions = self.keyBits, self.lodLevels, self.coarseVoxelDivisions
        o.inputOnDevice = int(on_device)
        return o

    def pack(self, cloud: "ComputeRasteriserPointCloud", positions, colors, sourceIndices: bool = False):
        """replacePackedPointCloud(packer:queue:positions:colors:count:) — GPUPacker.swift:507-519."""
        pos, col = _float4(positions, 0.0), _float4(colors, 1.0)
        n = pos.shape[0]
        src = np.zeros(max(n, 1), np.uint32) if sourceIndices else None
        o = self._opts(False)
        self.ctx.check(self.ctx.lib.scr_pack(cloud.handle, pos.ctypes.data, col.ctypes.data, n, C.byref(o),
                                             _ptr(src)))
        return src[:n] if sourceIndices else None

    def packDevice(self, cloud: "ComputeRasteriserPointCloud", positions: "DeviceBuffer", colors: "DeviceBuffer",
                   count: int) -> None:
        o = self._opts(True)
        self.ctx.check(self.ctx.lib.scr_pack(cloud.handle, positions.ptr, colors.ptr, int(count), C.byref(o), None))

    def addRawChunksGPU(self, cloud: "ComputeRasteriserPointCloud", chunks) -> list:
        """addRawChunksGPU(packer:queue:chunks:) — GPUPacker.swift:545-577; returns first slots."""
        placed = []
        for positions, colors in chunks:
            pos, col = _float4(positions, 0.0), _float4(colors, 1.0)
            if pos.shape[0] == 0:
                continue
            first = C.c_int32(-1)
            o = self._opts(False)
            rc = self.ctx.lib.scr_pack_chunk(cloud.handle, pos.ctypes.data, col.ctypes.data, pos.shape[0],
                                             C.byref(o), C.byref(first))
            if rc == -2:  # no contiguous run free: skipped, as the reference does
                continue
            self.ctx.check(rc)
            placed.append(int(first.value))
        return placed


# ---------------------------------------------------------------- PLY input
class PLYPointCloudLoaderError(ScrError):
    """PLYPointCloudLoaderError (PLYPointCloudLoader.swift:4-31): ``code`` is
    the SCR_ERR_PLY_* value, the message the reference's description."""


class PLYPointCloudLoader:
    """PLYPointCloudLoader (PLYPointCloudLoader.swift:33-54), parsed natively
    by ``scr_ply_parse``.  ``parse``/``load`` return the loader's intermediate
    arrays — float32 (n, 4) positions (x, y, z, 0) and colours (r, g, b, 1) —
    and ``packInto`` hands them to the GPU packer, where the reference calls
    its CPU pack.  Needs no GPU for parsing."""

    @staticmethod
    def parse(data) -> tuple:
        lib = load_library()
        buf = np.frombuffer(data, np.uint8) if not isinstance(data, np.ndarray) else data.view(np.uint8).ravel()
        ptr = buf.ctypes.data if buf.size else None
        n = C.c_size_t(0)
        msg = C.create_string_buffer(256)
        rc = lib.scr_ply_parse(ptr, buf.size, 0, None, None, C.byref(n), msg, 256)
        if rc != 0:
            raise PLYPointCloudLoaderError(rc, msg.value.decode())
        pos = np.empty((n.value, 4), np.float32)
        col = np.empty((n.value, 4), np.float32)
        rc = lib.scr_ply_parse(ptr, buf.size, n.value, pos.ctypes.data, col.ctypes.data, C.byref(n), msg, 256)
        if rc != 0:
            raise PLYPointCloudLoaderError(rc, msg.value.decode())
        return pos, col

    @staticmethod
    def load(path) -> tuple:
        """load(url:) — the file is memory-mapped, as the reference's Data(contentsOf:options: .mappedIfSafe)."""
        size = Path(path).stat().st_size
        data = np.memmap(path, np.uint8, mode="r") if size else np.zeros(0, np.uint8)
        return PLYPointCloudLoader.parse(data)

    @staticmethod
    def packInto(packer: "GPUPacker", cloud: "ComputeRasteriserPointCloud", source) -> int:
        """Parses ``source`` (bytes, or a path) and packs it into ``cloud``; returns the point count."""
        if isinstance(source, (bytes, bytearray, memoryview, np.ndarray)):
            pos, col = PLYPointCloudLoader.parse(source)
        else:
            pos, col = PLYPointCloudLoader.load(source)
        packer.pack(cloud, pos, col)
        return pos.shape[0]


# ---------------------------------------------------------------- synthetic scenes
def scene_params(**kw) -> CSceneParams:
    p = CSceneParams()
    load_library().scr_scene_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def gen_scene(ctx: Context, count: int, on_device: bool = True, **kw):
    """Synthetic LiDAR-like scene (float4 positions, float4 colours).

    on_device=True returns two DeviceBuffers (inputs stay in HBM); otherwise
    two (count, 4) float32 numpy arrays."""
    p = scene_params(**kw)
    if on_device:
        pos, col = DeviceBuffer(ctx, count * 16), DeviceBuffer(ctx, count * 16)
        ctx.check(ctx.lib.scr_gen_scene(ctx.handle, C.byref(p), count, pos.ptr, col.ptr, 1))
        return pos, col
    pos = np.zeros((count, 4), np.float32)
    col = np.zeros((count, 4), np.float32)
    ctx.check(ctx.lib.scr_gen_scene(ctx.handle, C.byref(p), count, pos.ctypes.data, col.ctypes.data, 0))
    return pos, col


def scene_height(x: float, y: float, **kw) -> float:
    """Terrain height of the generated scene at (x, y)."""
    p = scene_params(**kw)
    return float(load_library().scr_scene_height(C.byref(p), float(x), float(y)))


def catmull_rom(points: np.ndarray, frames: int) -> tuple:
    """Uniform Catmull-Rom spline through ``points`` (K, 3), end points
    duplicated; ``frames`` samples evenly spaced in the spline parameter.
    Returns (positions, unit tangents), each (frames, 3) float64."""
    p = np.asarray(points, np.float64)
    ext = np.concatenate([p[:1], p, p[-1:]], axis=0)
    segs = len(p) - 1
    pos, tan = [], []
    for f in range(frames):
        u = f / max(frames - 1, 1) * segs
        i = min(int(u), segs - 1)
        t = u - i
        p0, p1, p2, p3 = ext[i], ext[i + 1], ext[i + 2], ext[i + 3]
        t2, t3 = t * t, t * t * t
        pos.append(0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t2 +
                          (-p0 + 3 * p1 - 3 * p2 + p3) * t3))
        d = 0.5 * ((-p0 + p2) + 2 * (2 * p0 - 5 * p1 + 4 * p2 - p3) * t + 3 * (-p0 + 3 * p1 - 3 * p2 + p3) * t2)
        n = np.linalg.norm(d)
        tan.append(d / n if n > 0 else np.array([1.0, 0.0, 0.0]))
    return np.array(pos), np.array(tan)


def flythrough_cameras(frames: int = 120, pathSeed: int = 0x5C0003, width: int = 2560, height: int = 1440,
                       controlPoints: int = 8, minAltitude: float = 50.0, maxAltitude: float = 400.0,
                       fov: float = 60.0, **scene) -> list:
    """Fly-through over the generated scene (``scene``: scene_params fields):
    a Catmull-Rom spline through ``controlPoints`` points drawn from
    ``pathSeed`` at ``minAltitude``..``maxAltitude`` metres above the terrain;
    each frame looks 200 m ahead along the tangent and 30 m down."""
    prm = scene_params(**scene)
    rng = np.random.default_rng(pathSeed)
    half = float(prm.halfExtent)
    ctrl = []
    for _ in range(controlPoints):
        x, y = rng.uniform(-0.8 * half, 0.8 * half, 2)
        z = scene_height(x, y, **scene) + rng.uniform(minAltitude, maxAltitude)
        ctrl.append((x, y, z))
    pos, tan = catmull_rom(np.array(ctrl), frames)
    cams = []
    for p, t in zip(pos, tan):
        target = p + 200.0 * t + np.array([0.0, 0.0, -30.0])
        cams.append(Camera.perspective(tuple(p), tuple(target), (0.0, 0.0, 1.0), fov=fov, aspect=width / height,
                                       near=1.0, far=4.0 * half * 2.0))
    return cams


# ---------------------------------------------------------------- benchmark workloads
# The scenes and cameras the benchmark measures and the full-size parity
# tests render (synthetic inputs; no reference counterpart).

HEADLINE_SCENE = dict(seed=0x5C0003, halfExtent=4000.0, numBoxes=32000)  # 100M points, 8 km square
MULTIVIEW_SLOTS = (1_000_000, 5_000_000, 20_000_000, 50_000_000)


def overview_camera(width: int, height: int) -> Camera:
    """The headline frame's pose over the 8 km scene: every batch in view."""
    return Camera(look_at((0.0, -7200.0, 5200.0), (0.0, 300.0, 0.0), (0.0, 0.0, 1.0)),
                  perspective_reverse_z(55.0, width / height, 10.0, 30000.0))


def generate_packed_cloud(ctx: Context, count: int, keyBits: int = 63, **scene) -> "ComputeRasteriserPointCloud":
    """A generated scene, generated and packed on the GPU (inputs never leave HBM)."""
    pos, col = gen_scene(ctx, count, on_device=True, **scene)
    cloud = ComputeRasteriserPointCloud(ctx, 1, DEFAULT_POINTS_PER_BATCH)
    GPUPacker(ctx, keyBits=keyBits).packDevice(cloud, pos, col, count)
    ctx.sync()
    pos.free()
    col.free()
    return cloud


def cloud_bounds(cloud: "ComputeRasteriserPointCloud"):
    """Centre and diagonal of the resident batches' boxes."""
    b = cloud.readPool(POOL_BATCHES)
    b = b[b["state"] != 0]
    lo = np.array([b["minX"].min(), b["minY"].min(), b["minZ"].min()], np.float64)
    hi = np.array([b["maxX"].max(), b["maxY"].max(), b["maxZ"].max()], np.float64)
    return (lo + hi) * 0.5, float(np.linalg.norm(hi - lo))


def multiview_workload(ctx: Context, n_views: int = 64, slots=MULTIVIEW_SLOTS):
    """Batched multi-view workload: clouds of 1M/5M/20M/50M points of the 2 km
    scene (seeds 0x5C0003..6); view v renders cloud v mod 4 from an eye
    uniform on the sphere of 1.5x the cloud's diagonal at a resolution drawn
    from {256x256, 512x384, 1024x768, 1920x1080}, screen-space points 1-3 px.
    Returns (clouds, cameras, sizes, viewClouds, configuration)."""
    clouds, centres, diags = [], [], []
    for i, n in enumerate(slots):
        c = generate_packed_cloud(ctx, n, seed=0x5C0003 + i, halfExtent=1000.0, numBoxes=2000)
        ctr, d = cloud_bounds(c)
        clouds.append(c)
        centres.append(ctr)
        diags.append(d)
    cams, sizes, which = [], [], []
    for v in range(n_views):
        k = v % len(clouds)
        cc, ss = sphere_views(centres[k], 1.5 * diags[k], 1, 0x5C0004 * 1000 + v)
        cams += cc
        sizes += ss
        which.append(k)
    cfg = ComputeRasteriserConfiguration(minimumPointSize=1.0, maximumPointSize=3.0, pointSizeScale=1500.0)
    return clouds, cams, sizes, which, cfg


def version() -> str:
    return load_library().scr_version().decode()
