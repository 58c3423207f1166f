"""Shared GPU-vs-oracle plumbing for the parity tests."""
import numpy as np

from _pkg import orc, scr
from fixtures import jittered_cloud

PLANES = {
    "depth": scr.PLANE_DEPTH, "red": scr.PLANE_RED, "green": scr.PLANE_GREEN, "blue": scr.PLANE_BLUE,
    "count": scr.PLANE_COUNT, "rgba8": scr.PLANE_RGBA8, "depthF32": scr.PLANE_DEPTH_F32,
}


def config1_camera(width=1280, height=720):
    """Cube-grid case: eye (2.5, 2, 3) → origin, up +Y, vFOV 60°, near 0.1, far 100."""
    return scr.Camera.perspective((2.5, 2.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), fov=60.0,
                                  aspect=width / height, near=0.1, far=100.0)


def gpu_cloud(ctx, packed, model=None):
    c = scr.ComputeRasteriserPointCloud.from_packed(ctx, packed)
    if model is not None:
        c.worldMatrix = np.asarray(model, np.float32)
    return c


def oracle_cloud(packed, model=None):
    c = orc.Cloud.from_packed(packed)
    if model is not None:
        c.model = np.asarray(model, np.float32)
    return c


def render_gpu_hqs(ctx, clouds, cfg, cam, w, h):
    ras = scr.ComputeRasteriser(ctx, cfg)
    for c in clouds:
        ras.addPointCloud(c)
    ras.resize(w, h)
    ras.render(cam)
    out = {k: ras.read(v) for k, v in PLANES.items()}
    out["depthF32"] = out["depthF32"].view(np.float32)
    return out


def render_gpu_nearest(ctx, clouds, cfg, cam, w, h):
    ras = scr.ComputeRasteriser(ctx, cfg)
    for c in clouds:
        ras.addPointCloud(c)
    ras.resize(w, h)
    ras.render(cam)
    return {"depth": ras.read(scr.PLANE_NEAREST_DEPTH), "index": ras.read(scr.PLANE_NEAREST_INDEX),
            "rgba8": ras.read(scr.PLANE_RGBA8), "depthF32": ras.read(scr.PLANE_DEPTH_F32).view(np.float32)}


def assert_planes_equal(gpu, ref, keys=None):
    keys = keys or ref.keys()
    for k in keys:
        g, r = np.asarray(gpu[k]), np.asarray(ref[k])
        if g.dtype == np.float32:
            g, r = g.view(np.uint32), r.view(np.uint32)
        bad = np.argwhere(g != r)
        assert bad.size == 0, f"plane {k}: {len(bad)} pixels differ, first at {bad[:5].tolist()} " \
                              f"gpu={g[tuple(bad[0])] if len(bad) else None} ref={r[tuple(bad[0])] if len(bad) else None}"


def multi_file_packed():
    """Batches spread over three RasterFiles with different world matrices
    (rotation + translation, uniform scale), as tiled datasets use them."""
    pos, col = jittered_cloud(6000, 0x99)
    packed = orc.pack(pos, col, pointsPerBatch=500)
    files = scr.identity_files(3)
    a = np.deg2rad(30.0)
    rot = np.eye(4, dtype=np.float32)
    rot[0, 0], rot[0, 1], rot[1, 0], rot[1, 1] = np.cos(a), np.sin(a), -np.sin(a), np.cos(a)  # [column][row]
    rot[3, :3] = (3.0, -1.0, 0.5)
    files[1]["world"] = rot
    sc = np.eye(4, dtype=np.float32) * np.float32(0.5)
    sc[3] = (-2.0, 2.0, 1.0, 1.0)
    files[2]["world"] = sc
    packed.files = files
    packed.batches["fileIndex"] = np.arange(packed.batchCount) % 3
    return packed


MODEL = np.eye(4, dtype=np.float32)
MODEL[3, :3] = (0.5, 1.0, -0.25)   # translation (column 3)
