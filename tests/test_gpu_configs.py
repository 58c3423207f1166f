"""Full-size parity on the generated workloads: the 20M-point scene (HQS and
nearest, 1920x1080, 1-3 px screen-space points) and frames of the 120-frame
CLOD fly-through over the 100M-point 8 km scene at 2560x1440.  Every plane is
compared bit for bit with the oracle (multithreaded on the host)."""
import os

import numpy as np
import pytest

from _pkg import orc, scr
from helpers import assert_planes_equal, render_gpu_hqs, render_gpu_nearest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _gen_packed(ctx, n, **scene):
    pos, col = scr.gen_scene(ctx, n, on_device=True, **scene)
    cloud = scr.ComputeRasteriserPointCloud(ctx, 1, scr.DEFAULT_POINTS_PER_BATCH)
    scr.GPUPacker(ctx, keyBits=63).packDevice(cloud, pos, col, n)
    ctx.sync()
    pos.free()
    col.free()
    return cloud


def _oracle_cloud(scr_cloud):
    return orc.Cloud(scr_cloud.readPool(scr.POOL_BATCHES), scr.identity_files(1),
                     scr_cloud.readPool(scr.POOL_XYZ_LOW), scr_cloud.readPool(scr.POOL_XYZ_MED),
                     scr_cloud.readPool(scr.POOL_XYZ_HIGH), scr_cloud.readPool(scr.POOL_COLORS),
                     scr_cloud.readPool(scr.POOL_LEVELS))


def test_config2_20m_hqs_and_nearest_full_size(ctx):
    cloud = _gen_packed(ctx, 20_000_000, seed=0x5C0002, halfExtent=1000.0, numBoxes=2000)
    oc = _oracle_cloud(cloud)
    W, H = 1920, 1080
    cam = scr.Camera.perspective((-1200.0, -1200.0, 600.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), fov=60.0,
                                 aspect=W / H, near=1.0, far=10000.0)
    cfg = scr.ComputeRasteriserConfiguration(minimumPointSize=1.0, maximumPointSize=3.0, pointSizeScale=1500.0,
                                             enableCLOD=False)
    gpu = render_gpu_hqs(ctx, [cloud], cfg, cam, W, H)
    ref = orc.render_hqs([oc], cfg, cam.viewMatrix, cam.projectionMatrix, W, H, threads=THREADS)
    assert ref["count"].sum() > 10_000_000
    assert_planes_equal(gpu, ref)
    cfg.mode = scr.MODE_NEAREST_POINT
    gpu = render_gpu_nearest(ctx, [cloud], cfg, cam, W, H)
    ref = orc.render_nearest(oc, cfg, cam.viewMatrix, cam.projectionMatrix, W, H)
    assert (ref["depth"] > 0).sum() > 300_000
    assert_planes_equal(gpu, ref)
    cloud.close()


def test_config3_100m_clod_flythrough_frames_full_size(ctx):
    scene = dict(seed=0x5C0003, halfExtent=4000.0, numBoxes=32000)
    cloud = _gen_packed(ctx, 100_000_000, **scene)
    oc = _oracle_cloud(cloud)
    W, H = 2560, 1440
    cams = scr.flythrough_cameras(120, pathSeed=0x5C0003, width=W, height=H, **scene)
    cfg = scr.ComputeRasteriserConfiguration(enableCLOD=True, enableLODDither=True, lodBias=0)
    levels = set()
    for f in (0, 60, 119):
        gpu = render_gpu_hqs(ctx, [cloud], cfg, cams[f], W, H)
        vis = cloud.readVisible()
        levels |= set(np.minimum(vis["level"], 2).tolist())
        ref = orc.render_hqs([oc], cfg, cams[f].viewMatrix, cams[f].projectionMatrix, W, H, threads=THREADS)
        assert ref["count"].sum() > 1_000_000, f
        assert_planes_equal(gpu, ref)
        # CLOD drew a prefix of each batch, not every point
        assert int(vis["activePoints"].astype(np.int64).sum()) < 100_000_000
    # the 12-, 8- and 4-byte position streams were all exercised
    assert levels == {0, 1, 2}, levels
    cloud.close()
