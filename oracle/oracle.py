"""ctypes wrapper of the CPU oracle (libscr_oracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, never by the product path.  The oracle is
a scalar C restatement of the reference's passes (see scr_oracle.h for the
arithmetic contract and scr_oracle.c for the file:line citations).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libscr_oracle.so"

RASTER_BATCH_DTYPE = np.dtype(
    [("state", "<i4"), ("minX", "<f4"), ("minY", "<f4"), ("minZ", "<f4"), ("maxX", "<f4"), ("maxY", "<f4"),
     ("maxZ", "<f4"), ("numPoints", "<u4"), ("firstPoint", "<u4"), ("fileIndex", "<u4"), ("lodCum01", "<u4"),
     ("lodCum23", "<u4"), ("lodCum45", "<u4"), ("lodCum67", "<u4"), ("padding7", "<u4"), ("padding8", "<u4")]
)
VISIBLE_BATCH_DTYPE = np.dtype(
    [("batchIndex", "<u4"), ("level", "<i4"), ("lodThreshold", "<f4"), ("activePoints", "<u4")]
)
RASTER_FILE_DTYPE = np.dtype(
    [("transform", "<f4", (4, 4)), ("transformFrustum", "<f4", (4, 4)), ("world", "<f4", (4, 4)),
     ("prevTransform", "<f4", (4, 4))]
)


class OrcCloud(C.Structure):
    _fields_ = [
        ("batches", C.c_void_p), ("nSlots", C.c_int), ("files", C.c_void_p), ("nFiles", C.c_int),
        ("xyzLow", C.c_void_p), ("xyzMed", C.c_void_p), ("xyzHigh", C.c_void_p), ("colors", C.c_void_p),
        ("levels", C.c_void_p), ("tints", C.c_void_p), ("disp", C.c_void_p), ("prevDisp", C.c_void_p),
    ]


class OrcPlanes(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("red", C.c_void_p), ("green", C.c_void_p), ("blue", C.c_void_p),
                ("count", C.c_void_p), ("rgba8", C.c_void_p), ("depthF32", C.c_void_p)]


class OrcPackOut(C.Structure):
    _fields_ = [
        ("batches", C.c_void_p), ("xyzLow", C.c_void_p), ("xyzMed", C.c_void_p), ("xyzHigh", C.c_void_p),
        ("colors", C.c_void_p), ("levels", C.c_void_p), ("sourceIndices", C.c_void_p),
        ("orderedPositions", C.c_void_p), ("boundsMin", C.c_float * 3), ("boundsMax", C.c_float * 3),
    ]


_lib = None


def build() -> Path:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True, env=dict(os.environ))
    return LIB_PATH


def use_library(path: Optional[Path] = None) -> Optional[Path]:
    """Serve every later call from the oracle library at ``path`` (a build of
    scr_oracle.c with other compile flags, e.g. a planted mistake for the
    mutation test), or from the default build with None; returns the path in
    use before."""
    global _lib, _path
    prev = _path
    _lib = None
    _path = None if path is None else Path(path)
    return prev


_path: Optional[Path] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = _path or LIB_PATH
        if path == LIB_PATH and not LIB_PATH.exists():
            build()
        L = C.CDLL(str(path))
        f = C.c_float
        u32 = C.c_uint32
        vp = C.c_void_p
        sig = {
            "orc_log2f": (f, [f]), "orc_expf": (f, [f]), "orc_hash_unit": (f, [u32]), "orc_depth_to_uint": (u32, [f]),
            "orc_uint_to_depth": (f, [u32]), "orc_f2u": (u32, [f]), "orc_f2i": (C.c_int32, [f]),
            "orc_mat4_mul": (None, [vp, vp, vp]), "orc_mat4_vec": (None, [vp, vp, vp]),
            "orc_transform_point": (None, [vp, vp, vp]),
            "orc_project_q": (None, [vp, vp, vp, vp, vp]),
            "orc_decode_point": (None, [u32, vp, vp, vp, vp, C.c_int, C.c_int, vp]),
            "orc_morton30": (u32, [u32, u32, u32]), "orc_morton63": (C.c_uint64, [u32, u32, u32]),
            "orc_inside_footprint": (C.c_int, [C.c_int, C.c_int, C.c_int]),
            "orc_precision_level": (C.c_int, [f]), "orc_lod_threshold": (f, [f, C.c_int]),
            "orc_intersects_frustum": (C.c_int, [vp, vp, vp]),
            "orc_pixel_size_on_screen": (f, [vp, vp, vp, vp, C.c_int, C.c_int]),
            "orc_footprint_radius": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, vp]),
            "orc_update_files": (None, [vp, C.c_int, vp, vp, vp, vp, vp]),
            "orc_cull": (C.c_int, [C.POINTER(OrcCloud), vp, vp, vp, C.c_int, C.c_int, vp]),
            "orc_render_hqs": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.POINTER(OrcPlanes), C.c_int]),
            "orc_last_hqs_pass_seconds": (None, [vp]),
            "orc_render_nearest": (C.c_int, [C.POINTER(OrcCloud), vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp,
                                             C.c_int]),
            "orc_nearest_cpu_reference": (None, [C.POINTER(OrcCloud), C.c_int, C.c_int, vp, vp, C.c_int, vp, vp]),
            "orc_pick_search": (C.c_int, [vp, C.c_int, C.c_int, f, f, C.c_int, C.POINTER(u32)]),
            "orc_pack": (C.c_int, [vp, C.c_int, vp, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(OrcPackOut)]),
            "orc_pack_gpu": (C.c_int, [vp, C.c_int, vp, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(OrcPackOut)]),
            "orc_gpu_lod_enabled": (C.c_int, [C.c_int, C.c_int]),
            "orc_morton_keys": (None, [vp, C.c_int, C.c_size_t, C.c_int, vp]),
            "orc_lod_levels": (None, [vp, C.c_size_t, vp, vp, C.c_int, C.c_int, vp]),
            "orc_cube_grid": (None, [C.c_int, f, vp, vp]),
            "orc_diff_fixed": (C.c_int64, [f]),
            "orc_diff_pixel": (f, [C.c_int64, u32, f]),
            "orc_render_diff": (C.c_int, [C.POINTER(OrcCloud), vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp,
                                          C.c_int]),
            "orc_diff_backward": (C.c_int, [C.POINTER(OrcCloud), vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp,
                                            vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32)


# ---------------------------------------------------------------- scalar helpers
def log2f(x: float) -> float:
    return lib().orc_log2f(x)


def hash_unit(x: int) -> float:
    return lib().orc_hash_unit(x)


def depth_to_uint(z: float) -> int:
    return lib().orc_depth_to_uint(z)


def morton30(x: int, y: int, z: int) -> int:
    return lib().orc_morton30(x, y, z)


def morton63(x: int, y: int, z: int) -> int:
    return lib().orc_morton63(x, y, z)


def mat_mul(a, b) -> np.ndarray:
    a, b = _f32(a), _f32(b)
    out = np.zeros((4, 4), np.float32)
    lib().orc_mat4_mul(a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def transform_point(m, p) -> np.ndarray:
    """m (column-major 4x4) applied to the point (p, 1) as the render passes do."""
    m = _f32(m).reshape(-1)
    p = _f32(p).reshape(-1)
    out = np.zeros(4, np.float32)
    lib().orc_transform_point(m.ctypes.data, p.ctypes.data, out.ctypes.data)
    return out


def project_q(m, mn, step, qf) -> np.ndarray:
    """The render passes' projection of an undisplaced point from its
    quantised coordinates qf (floats) under transform m, batch minimum mn and
    decode step per axis (contract C1)."""
    m = _f32(m).reshape(-1)
    args = [_f32(x).reshape(-1) for x in (mn, step, qf)]
    out = np.zeros(4, np.float32)
    lib().orc_project_q(m.ctypes.data, args[0].ctypes.data, args[1].ctypes.data, args[2].ctypes.data,
                        out.ctypes.data)
    return out


def decode_point(batch: np.ndarray, words, level: int, fused: bool = True) -> np.ndarray:
    """decodePoint of the single point whose (low, med, high) words are given,
    under the batch record ``batch`` (one RASTER_BATCH_DTYPE row; firstPoint ignored)."""
    b = np.array(batch, dtype=RASTER_BATCH_DTYPE).reshape(1)
    w = [np.array([x], dtype=np.uint32) for x in words]
    out = np.zeros(3, np.float32)
    lib().orc_decode_point(0, b.ctypes.data, w[0].ctypes.data, w[1].ctypes.data, w[2].ctypes.data, level,
                           1 if fused else 0, out.ctypes.data)
    return out


def expf(x: float) -> float:
    """Contract C8 e^x (the coverage resolve's exponential)."""
    return lib().orc_expf(x)


def lod_threshold(px: float, bias: int) -> float:
    return lib().orc_lod_threshold(px, bias)


def intersects_frustum(m, mn, mx) -> bool:
    """intersectsFrustum (Common.metal:119-153) of the box [mn, mx] under m."""
    m, mn, mx = _f32(m), _f32(mn), _f32(mx)   # held until the call returns
    return bool(lib().orc_intersects_frustum(m.ctypes.data, mn.ctypes.data, mx.ctypes.data))


def _file(world=None) -> np.ndarray:
    f = identity_files(1)
    if world is not None:
        f["world"] = _f32(world).reshape(4, 4)
    return f


def pixel_size_on_screen(mn, mx, view, proj, width: int, height: int, world=None) -> float:
    """pixelSizeOnScreen (Common.metal:155-172) of the batch box [mn, mx]."""
    b = np.zeros(1, RASTER_BATCH_DTYPE)
    b["minX"], b["minY"], b["minZ"] = _f32(mn)
    b["maxX"], b["maxY"], b["maxZ"] = _f32(mx)
    f = _file(world)
    view, proj = _f32(view), _f32(proj)
    return lib().orc_pixel_size_on_screen(b.ctypes.data, f.ctypes.data, view.ctypes.data, proj.ctypes.data, width,
                                          height)


def precision_level(px: float) -> int:
    return lib().orc_precision_level(px)


def footprint_radius(point, cfg, view, proj, width: int, height: int, world=None) -> int:
    """pointFootprintRadius (Common.metal:18-61) of a decoded point."""
    f = _file(world)
    c = config_from(cfg)
    point, view, proj = _f32(point), _f32(view), _f32(proj)
    return lib().orc_footprint_radius(point.ctypes.data, f.ctypes.data, view.ctypes.data, proj.ctypes.data, width,
                                      height, C.addressof(c))


# ---------------------------------------------------------------- fixtures
def cube_grid(n: int, extent: float = 1.0):
    """cubeGrid lattice (positions (N,3) float32, colors (N,4) float32)."""
    n = max(n, 2)
    pos = np.zeros((n ** 3, 3), np.float32)
    col = np.zeros((n ** 3, 4), np.float32)
    lib().orc_cube_grid(n, extent, pos.ctypes.data, col.ctypes.data)
    return pos, col


class Packed:
    """Output of orc_pack (mirrors PackedPointCloud)."""

    def __init__(self, n: int, nb: int):
        self.batches = np.zeros(nb, RASTER_BATCH_DTYPE)
        self.xyzLow = np.zeros(n, np.uint32)
        self.xyzMed = np.zeros(n, np.uint32)
        self.xyzHigh = np.zeros(n, np.uint32)
        self.colors = np.zeros(n, np.uint32)
        self.levels = np.zeros(n, np.uint8)
        self.sourceIndices = np.zeros(n, np.uint32)
        self.orderedPositions = np.zeros((n, 3), np.float32)
        self.boundsMin = np.zeros(3, np.float32)
        self.boundsMax = np.zeros(3, np.float32)
        self.files = identity_files(1)

    @property
    def pointCount(self) -> int:
        return int(self.colors.shape[0])

    @property
    def batchCount(self) -> int:
        return int(self.batches.shape[0])


def identity_files(n: int = 1) -> np.ndarray:
    f = np.zeros(n, RASTER_FILE_DTYPE)
    for k in ("transform", "transformFrustum", "world", "prevTransform"):
        f[k] = np.eye(4, dtype=np.float32)
    return f


def gpu_lod_enabled(lodLevels: int, coarseVoxelDivisions: int) -> bool:
    """GPUPacker.init's LOD gate (GPUPacker.swift:35, 86-99)."""
    return bool(lib().orc_gpu_lod_enabled(lodLevels, coarseVoxelDivisions))


def pack(positions, colors, pointsPerBatch: int = 10240, lodLevels: int = 4, coarseVoxelDivisions: int = 64,
         keyBits: int = 30, gpuRules: bool = False) -> Packed:
    """PackedPointCloudFixtures.pack (stable Morton order; keyBits 30 or 63);
    gpuRules: under GPUPacker's LOD gate, what GPUPacker.pack produces."""
    pos = _f32(positions).reshape(-1, 3)
    col = _f32(colors).reshape(-1, 4)
    n = pos.shape[0]
    nb = (n + pointsPerBatch - 1) // pointsPerBatch if n else 0
    out = Packed(n, nb)
    o = OrcPackOut()
    o.batches, o.xyzLow, o.xyzMed, o.xyzHigh = _p(out.batches), _p(out.xyzLow), _p(out.xyzMed), _p(out.xyzHigh)
    o.colors, o.levels, o.sourceIndices, o.orderedPositions = (_p(out.colors), _p(out.levels),
                                                               _p(out.sourceIndices), _p(out.orderedPositions))
    fn = lib().orc_pack_gpu if gpuRules else lib().orc_pack
    rc = fn(pos.ctypes.data, 3, col.ctypes.data, n, pointsPerBatch, lodLevels, coarseVoxelDivisions, keyBits,
            C.byref(o))
    if rc < 0:
        raise ValueError(f"orc_pack failed ({rc})")
    out.boundsMin[:] = list(o.boundsMin)
    out.boundsMax[:] = list(o.boundsMax)
    return out


def morton_keys(positions, keyBits: int = 30) -> np.ndarray:
    pos = _f32(positions).reshape(-1, 3)
    keys = np.zeros(pos.shape[0], np.uint64)
    lib().orc_morton_keys(pos.ctypes.data, 3, pos.shape[0], keyBits, keys.ctypes.data)
    return keys


def lod_levels(sorted_positions, bmin, bmax, lodLevels: int = 4, coarse: int = 64) -> np.ndarray:
    sp = _f32(sorted_positions).reshape(-1, 3)
    out = np.zeros(sp.shape[0], np.uint8)
    bmin, bmax = _f32(bmin), _f32(bmax)
    lib().orc_lod_levels(sp.ctypes.data, sp.shape[0], bmin.ctypes.data, bmax.ctypes.data, lodLevels,
                         coarse, out.ctypes.data)
    return out


# ---------------------------------------------------------------- clouds and frames
class Cloud:
    """Pool-layout arrays of one cloud, as the rasteriser reads them."""

    def __init__(self, batches, files, xyzLow, xyzMed, xyzHigh, colors, levels, tints=None, disp=None,
                 prevDisp=None):
        self.batches = np.ascontiguousarray(batches, RASTER_BATCH_DTYPE)
        self.files = np.ascontiguousarray(files, RASTER_FILE_DTYPE)
        self.xyzLow = np.ascontiguousarray(xyzLow, np.uint32)
        self.xyzMed = np.ascontiguousarray(xyzMed, np.uint32)
        self.xyzHigh = np.ascontiguousarray(xyzHigh, np.uint32)
        self.colors = np.ascontiguousarray(colors, np.uint32)
        self.levels = np.ascontiguousarray(levels, np.uint8)
        self.tints = None if tints is None else np.ascontiguousarray(tints, np.float32).reshape(-1, 4)
        # displacements, float4 per pool point (xyz used); prevDisp: the
        # previous frame's (None: the current ones)
        self.disp = None if disp is None else np.ascontiguousarray(disp, np.float32).reshape(-1, 4)
        self.prevDisp = None if prevDisp is None else np.ascontiguousarray(prevDisp, np.float32).reshape(-1, 4)
        self.model = np.eye(4, dtype=np.float32)

    @staticmethod
    def from_packed(packed, ppb: Optional[int] = None) -> "Cloud":
        """The pool a from_packed upload builds: slot i holds batch i at i*ppb."""
        nb = packed.batchCount
        if ppb is None:
            ppb = max(1, int(packed.batches["numPoints"].max())) if nb else 10240
        cap = max(1, nb) * ppb
        low = np.zeros(cap, np.uint32)
        med = np.zeros(cap, np.uint32)
        high = np.zeros(cap, np.uint32)
        colors = np.zeros(cap, np.uint32)
        levels = np.zeros(cap, np.uint8)
        batches = np.zeros(max(1, nb), RASTER_BATCH_DTYPE)
        for i in range(nb):
            b = packed.batches[i].copy()
            s, n = int(b["firstPoint"]), int(b["numPoints"])
            d = i * ppb
            low[d:d + n] = packed.xyzLow[s:s + n]
            med[d:d + n] = packed.xyzMed[s:s + n]
            high[d:d + n] = packed.xyzHigh[s:s + n]
            colors[d:d + n] = packed.colors[s:s + n]
            levels[d:d + n] = packed.levels[s:s + n]
            b["firstPoint"] = d
            b["state"] = 1
            batches[i] = b
        return Cloud(batches, packed.files, low, med, high, colors, levels)

    def composed(self, vp, prev_vp=None) -> "Cloud":
        """Apply updateFiles(viewProjection, model, prevViewProjection) to a copy."""
        files = np.zeros_like(self.files)
        prev = None if prev_vp is None else _f32(prev_vp)
        vp, model = _f32(vp), _f32(self.model)
        lib().orc_update_files(self.files.ctypes.data, len(self.files), vp.ctypes.data, model.ctypes.data, None, None if prev is None else prev.ctypes.data,
                               files.ctypes.data)
        c = Cloud(self.batches, files, self.xyzLow, self.xyzMed, self.xyzHigh, self.colors, self.levels, self.tints,
                  self.disp, self.prevDisp)
        c.model = self.model
        return c

    def c_struct(self) -> OrcCloud:
        s = OrcCloud()
        s.batches, s.nSlots = self.batches.ctypes.data, len(self.batches)
        s.files, s.nFiles = self.files.ctypes.data, len(self.files)
        s.xyzLow, s.xyzMed, s.xyzHigh = self.xyzLow.ctypes.data, self.xyzMed.ctypes.data, self.xyzHigh.ctypes.data
        s.colors, s.levels = self.colors.ctypes.data, self.levels.ctypes.data
        s.tints = None if self.tints is None else self.tints.ctypes.data
        s.disp = None if self.disp is None else self.disp.ctypes.data
        s.prevDisp = None if self.prevDisp is None else self.prevDisp.ctypes.data
        return s


class Config(C.Structure):
    """scr_config (include/scr.h)."""
    _fields_ = [
        ("mode", C.c_int32), ("depthTolerance", C.c_float), ("backgroundColor", C.c_float * 4),
        ("enableFrustumCulling", C.c_int32), ("lodBias", C.c_int32), ("enableCLOD", C.c_int32),
        ("enableLODDither", C.c_int32), ("colorizeChunks", C.c_int32), ("colorizeOverdraw", C.c_int32),
        ("pointSizeMode", C.c_int32), ("minimumPointSize", C.c_float), ("maximumPointSize", C.c_float),
        ("pointSizeScale", C.c_float), ("applyTint", C.c_int32), ("tintAlphaIsCoverage", C.c_int32),
        ("motionBlur", C.c_float), ("holeFillIterations", C.c_int32), ("motionBlurSamples", C.c_int32),
        ("motionBlurMaxSpread", C.c_float), ("applyDisplacement", C.c_int32), ("reserved", C.c_int32 * 1),
    ]


def config_from(cfg) -> Config:
    """Copy any object with the configuration's field names into a Config."""
    c = Config()
    for name, _ in Config._fields_:
        if name == "reserved" or not hasattr(cfg, name):
            continue
        v = getattr(cfg, name)
        if name == "backgroundColor":
            for i in range(4):
                c.backgroundColor[i] = float(v[i])
        else:
            setattr(c, name, type(getattr(c, name))(v))
    return c


def cull(cloud: Cloud, cfg, view, proj, width: int, height: int) -> np.ndarray:
    view, proj = _f32(view), _f32(proj)   # held until the calls return
    vp = mat_mul(proj, view)
    cc = cloud.composed(vp)
    s = cc.c_struct()
    out = np.zeros(max(1, len(cc.batches)), VISIBLE_BATCH_DTYPE)
    c = config_from(cfg)
    n = lib().orc_cull(C.byref(s), C.addressof(c), view.ctypes.data, proj.ctypes.data, width, height,
                       out.ctypes.data)
    return out[:n]


def render_hqs(clouds: Sequence[Cloud], cfg, view, proj, width: int, height: int, threads: int = 1,
               prev_view_proj=None) -> dict:
    """HQS frame: returns dict of planes (H, W).  prev_view_proj is the
    previous frame's projection . view (motion blur; default: this frame's)."""
    view, proj = _f32(view), _f32(proj)   # held until the calls return
    vp = mat_mul(proj, view)
    composed = [c.composed(vp, prev_view_proj) for c in clouds]
    structs = (OrcCloud * max(1, len(composed)))(*[c.c_struct() for c in composed])
    n = width * height
    planes = {k: np.zeros(n, np.uint32) for k in ("depth", "red", "green", "blue", "count", "rgba8")}
    planes["depthF32"] = np.zeros(n, np.float32)
    pl = OrcPlanes(*[planes[k].ctypes.data for k in ("depth", "red", "green", "blue", "count", "rgba8", "depthF32")])
    c = config_from(cfg)
    rc = lib().orc_render_hqs(C.addressof(structs), len(composed), C.addressof(c), view.ctypes.data,
                              proj.ctypes.data, width, height, C.byref(pl), threads)
    if rc != 0:
        raise RuntimeError("orc_render_hqs failed")
    return {k: v.reshape(height, width) for k, v in planes.items()}


def last_hqs_pass_seconds() -> dict:
    """Wall-clock seconds of the phases of the process's last render_hqs."""
    out = np.zeros(5, np.float64)
    lib().orc_last_hqs_pass_seconds(out.ctypes.data)
    return dict(zip(("clear", "cull", "depth", "color", "resolve"), out.tolist()))


def render_nearest(cloud: Cloud, cfg, view, proj, width: int, height: int, threads: int = 1) -> dict:
    view, proj = _f32(view), _f32(proj)   # held until the calls return
    vp = mat_mul(proj, view)
    cc = cloud.composed(vp)
    s = cc.c_struct()
    n = width * height
    d = np.zeros(n, np.uint32)
    idx = np.zeros(n, np.uint32)
    rgba = np.zeros(n, np.uint32)
    df = np.zeros(n, np.float32)
    c = config_from(cfg)
    lib().orc_render_nearest(C.byref(s), C.addressof(c), view.ctypes.data, proj.ctypes.data, width,
                             height, d.ctypes.data, idx.ctypes.data, rgba.ctypes.data, df.ctypes.data, threads)
    return {"depth": d.reshape(height, width), "index": idx.reshape(height, width),
            "rgba8": rgba.reshape(height, width), "depthF32": df.reshape(height, width)}


def render_diff(cloud: Cloud, cfg, view, proj, width: int, height: int, params, threads: int = 1) -> dict:
    """Differentiable HQS forward: depth, count (H, W); sums int64 and image
    float32 (H, W, 3).  params: float32 (poolPoints, 4), rgb used."""
    view, proj = _f32(view), _f32(proj)   # held until the calls return
    vp = mat_mul(proj, view)
    cc = cloud.composed(vp)
    s = cc.c_struct()
    n = width * height
    prm = np.ascontiguousarray(params, np.float32)
    d = np.zeros(n, np.uint32)
    cnt = np.zeros(n, np.uint32)
    sums = np.zeros(n * 3, np.int64)
    img = np.zeros(n * 3, np.float32)
    c = config_from(cfg)
    rc = lib().orc_render_diff(C.byref(s), C.addressof(c), view.ctypes.data, proj.ctypes.data, width,
                               height, prm.ctypes.data, d.ctypes.data, cnt.ctypes.data, sums.ctypes.data,
                               img.ctypes.data, threads)
    if rc != 0:
        raise RuntimeError("orc_render_diff failed")
    return {"depth": d.reshape(height, width), "count": cnt.reshape(height, width),
            "sums": sums.reshape(height, width, 3), "image": img.reshape(height, width, 3)}


def diff_backward(cloud: Cloud, cfg, view, proj, width: int, height: int, depth, count, dLdImage, grad,
                  mask=None) -> None:
    """FP64 reference gradient: grad (poolPoints, 4) float64 += per-point sums."""
    view, proj = _f32(view), _f32(proj)   # held until the calls return
    vp = mat_mul(proj, view)
    cc = cloud.composed(vp)
    s = cc.c_struct()
    assert grad.dtype == np.float64 and grad.flags.c_contiguous
    d = np.ascontiguousarray(depth, np.uint32)
    cnt = np.ascontiguousarray(count, np.uint32)
    g = np.ascontiguousarray(dLdImage, np.float32)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    c = config_from(cfg)
    rc = lib().orc_diff_backward(C.byref(s), C.addressof(c), view.ctypes.data, proj.ctypes.data, width,
                                 height, d.ctypes.data, cnt.ctypes.data, g.ctypes.data, grad.ctypes.data, _p(m))
    if rc != 0:
        raise RuntimeError("orc_diff_backward failed")


def nearest_cpu_reference(cloud: Cloud, width: int, height: int, view=None, proj=None,
                          enableFrustumCulling: bool = False) -> dict:
    view = np.eye(4, dtype=np.float32) if view is None else _f32(view)
    proj = np.eye(4, dtype=np.float32) if proj is None else _f32(proj)
    vp = mat_mul(proj, view)
    cc = cloud.composed(vp)
    s = cc.c_struct()
    n = max(width, 0) * max(height, 0)
    d = np.zeros(max(n, 1), np.uint32)
    idx = np.zeros(max(n, 1), np.uint32)
    lib().orc_nearest_cpu_reference(C.byref(s), width, height, view.ctypes.data, proj.ctypes.data,
                                    int(enableFrustumCulling), d.ctypes.data, idx.ctypes.data)
    return {"depths": d[:n], "indices": idx[:n]}


def pick_search(indices, width: int, height: int, ndcX: float, ndcY: float, radius: int = 10):
    ind = np.ascontiguousarray(indices, np.uint32)
    out = C.c_uint32()
    found = lib().orc_pick_search(ind.ctypes.data, width, height, ndcX, ndcY, radius, C.byref(out))
    return int(out.value) if found else None


# ---------------------------------------------------------------- post
# Planted mistakes in the composite restatement, set only by the mutation
# test (tests/test_oracle_mutants.py), which requires a pin to fail for each;
# 0 is the oracle itself.
COMPOSITE_MUTANT = 0

COMPOSITE_OVER, COMPOSITE_DEPTH = 0, 1
BLEND_REPLACE, BLEND_OVER = 0, 1
COMPARE_ALWAYS, COMPARE_GREATER, COMPARE_GREATER_EQUAL, COMPARE_LESS, COMPARE_LESS_EQUAL = range(5)


def composite(rgba8, depth, target, targetDepth=None, mode: int = COMPOSITE_DEPTH, blend: int = BLEND_OVER,
              depthCompare: int = COMPARE_GREATER, writeDepth: bool = False):
    """ComputeRasteriser.draw's composite of the cloud's colour and reverse-Z
    depth onto a target (ComputeRasteriser.swift:672-683 draws a full-screen
    quad with Post.metal:23-50).  The fragment function returns the colour
    texel and, in the depth mode, discard_fragment() where depth <= 0
    (Post.metal:44-46) and the depth as the fragment depth; the pipeline's
    depth test (depthCompare against targetDepth, optional depth write) and
    colour blend are the render pass's: SOURCE-OVER with non-premultiplied
    source alpha, c = s.rgb * s.a + d.rgb * (1 - s.a), a = s.a + d.a * (1 - s.a),
    or REPLACE.  UNORM8 reads are byte / 255 (correctly rounded binary32),
    writes rint(saturate(x) * 255).  Returns (target, targetDepth) as new
    arrays; (H, W) uint32 and float32."""
    f32 = np.float32
    c = np.ascontiguousarray(rgba8, np.uint32)
    d = np.ascontiguousarray(depth, f32)
    t = np.array(target, np.uint32, copy=True)
    td = None if targetDepth is None else np.array(targetDepth, f32, copy=True)
    keep = np.ones(c.shape, bool)
    m = COMPOSITE_MUTANT
    if mode == COMPOSITE_DEPTH:
        keep = (d >= 0) if m == 16 else (d > 0)
        if depthCompare == COMPARE_GREATER:
            keep &= (d >= td) if m == 17 else (d > td)
        elif depthCompare == COMPARE_GREATER_EQUAL:
            keep &= d >= td
        elif depthCompare == COMPARE_LESS:
            keep &= d < td
        elif depthCompare == COMPARE_LESS_EQUAL:
            keep &= d <= td
        if writeDepth:
            td = np.where(keep | (m == 18), d, td).astype(f32)
    if blend == BLEND_REPLACE:
        out = c
    else:
        def ch(x, k):
            return ((x >> np.uint32(8 * k)) & np.uint32(0xFF)).astype(f32) / f32(255)

        def u8(x):
            return np.rint(np.clip(x.astype(f32), f32(0), f32(1)) * f32(255)).astype(np.uint32)

        sa = ch(c, 3)
        k = sa if m == 14 else (f32(1) - sa).astype(f32)
        out = u8(sa if m == 15 else (sa + (ch(t, 3) * k).astype(f32)).astype(f32)) << np.uint32(24)
        for j in range(3):
            v = ((ch(c, j) * sa).astype(f32) + (ch(t, j) * k).astype(f32)).astype(f32)
            out |= u8(v) << np.uint32(8 * j)
    t = np.where(keep, out, t).astype(np.uint32)
    return t, td


def hole_fill(rgba8: np.ndarray, iterations: int = 1) -> np.ndarray:
    """holeFillUpdate (HoleFill/Shaders.metal:10-48), ``iterations`` passes over
    an (H, W) uint32 RGBA8 image: a texel with alpha byte >= 128 is kept; an
    empty one takes the mean of its valid 8-neighbours' colours (bytes read as
    byte * fl(1/255), summed in the shader's row-major neighbour order in
    binary32, scaled by the correctly rounded 1/count) as rint(saturate(x) *
    255) bytes with alpha 255, or is kept when none is valid."""
    img = np.ascontiguousarray(rgba8, np.uint32)
    h, w = img.shape
    f32 = np.float32
    k = f32(1.0) / f32(255.0)
    for _ in range(iterations):
        valid = (img >> 24) >= 128
        s = np.zeros((h, w, 3), f32)
        n = np.zeros((h, w), f32)
        pad = np.zeros((h + 2, w + 2), np.uint32)
        pad[1:-1, 1:-1] = img
        vpad = np.zeros((h + 2, w + 2), bool)
        vpad[1:-1, 1:-1] = valid
        for oy in (-1, 0, 1):
            for ox in (-1, 0, 1):
                if ox == 0 and oy == 0:
                    continue
                q = pad[1 + oy:1 + oy + h, 1 + ox:1 + ox + w]
                ok = vpad[1 + oy:1 + oy + h, 1 + ox:1 + ox + w]
                for ch in range(3):
                    v = ((q >> (8 * ch)) & 0xFF).astype(f32) * k
                    s[:, :, ch] = np.where(ok, s[:, :, ch] + v, s[:, :, ch]).astype(f32)
                n = np.where(ok, n + f32(1), n).astype(f32)
        # texels without a valid neighbour (n = 0) are kept below; their
        # values here are discarded
        with np.errstate(divide="ignore", invalid="ignore"):
            r = (f32(1) / n).astype(f32)
            b = [np.rint(np.clip((s[:, :, ch] * r).astype(f32), 0, 1).astype(f32) * f32(255)).astype(np.uint32)
                 for ch in range(3)]
        filled = b[0] | (b[1] << 8) | (b[2] << 16) | np.uint32(0xFF000000)
        img = np.where(valid | (n == 0), img, filled).astype(np.uint32)
    return img
