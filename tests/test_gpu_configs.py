"""Full-size parity on the generated workloads: the 20M-point scene (HQS and
nearest, 1920x1080, 1-3 px screen-space points) and its GPU pack against the
CPU pack; the 100M-point 8 km scene's headline 1920x1080 frames (HQS with
CLOD off and on, nearest) and frames of its 120-frame CLOD fly-through at
2560x1440; the 64-view batched workload over clouds of 1M/5M/20M/50M points.
Every plane is compared bit for bit with the oracle (multithreaded on the
host)."""
import os

import numpy as np
import pytest

from _pkg import orc, scr
from helpers import assert_planes_equal, render_gpu_hqs, render_gpu_nearest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def _gen_packed(ctx, n, **scene):
    return scr.generate_packed_cloud(ctx, n, **scene)


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


@pytest.fixture(scope="module")
def headline(ctx):
    """The 100M-point 8 km scene, GPU-generated and GPU-packed (63-bit keys),
    with its pool read back for the oracle."""
    cloud = _gen_packed(ctx, 100_000_000, **scr.HEADLINE_SCENE)
    yield cloud, _oracle_cloud(cloud)
    cloud.close()


def _frame_planes(frame, nearest=False):
    if nearest:
        keys = {"depth": scr.PLANE_NEAREST_DEPTH, "index": scr.PLANE_NEAREST_INDEX}
    else:
        keys = {"depth": scr.PLANE_DEPTH, "red": scr.PLANE_RED, "green": scr.PLANE_GREEN, "blue": scr.PLANE_BLUE,
                "count": scr.PLANE_COUNT}
    shape = (frame.height, frame.width)
    out = {k: frame.read(v).reshape(shape) for k, v in keys.items()}
    out["rgba8"] = frame.read(scr.PLANE_RGBA8).reshape(shape)
    out["depthF32"] = frame.read(scr.PLANE_DEPTH_F32).view(np.float32).reshape(shape)
    return out


@pytest.mark.parametrize("mode", ["hqs", "hqs_clod", "nearest"])
def test_headline_100m_1080p_frames_full_size(ctx, headline, mode):
    """The benchmark's frame: every point of the 100M cloud in view at
    1920x1080, 1-px points; HQS with CLOD off (every point drawn) and on,
    and nearest."""
    cloud, oc = headline
    W, H = 1920, 1080
    cam = scr.overview_camera(W, H)
    if mode == "nearest":
        cfg = scr.ComputeRasteriserConfiguration(mode=scr.MODE_NEAREST_POINT, enableCLOD=False)
        gpu = render_gpu_nearest(ctx, [cloud], cfg, cam, W, H)
        ref = orc.render_nearest(oc, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, threads=THREADS)
        assert (ref["depth"] > 0).sum() > 500_000
    else:
        cfg = scr.ComputeRasteriserConfiguration(enableCLOD=(mode == "hqs_clod"))
        gpu = render_gpu_hqs(ctx, [cloud], cfg, cam, W, H)
        ref = orc.render_hqs([oc], cfg, cam.viewMatrix, cam.projectionMatrix, W, H, threads=THREADS)
        drawn = int(cloud.readVisible()["activePoints"].astype(np.int64).sum())
        if mode == "hqs":
            assert drawn == 100_000_000 and ref["count"].sum() > 50_000_000
        else:
            assert 0 < drawn < 100_000_000
    assert_planes_equal(gpu, ref)


def test_config3_100m_clod_flythrough_frames_full_size(ctx, headline):
    scene = scr.HEADLINE_SCENE
    cloud, oc = headline
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


def test_config4_64_views_full_size(ctx):
    """The benchmark's batched workload: 64 cameras of uneven resolution over
    four resident clouds (1M/5M/20M/50M points), autotuned shapes, replayed
    from the captured graph; every view's planes against the oracle, HQS and
    nearest."""
    clouds, cams, sizes, which, cfg = scr.multiview_workload(ctx, 64)
    ocs = [_oracle_cloud(c) for c in clouds]
    mv = scr.MultiViewRasteriser(ctx, clouds, sizes, which, cfg)
    try:
        for nearest in (False, True):
            mv.configuration = scr.ComputeRasteriserConfiguration(
                mode=scr.MODE_NEAREST_POINT if nearest else scr.MODE_HIGH_QUALITY_AVERAGE,
                minimumPointSize=cfg.minimumPointSize, maximumPointSize=cfg.maximumPointSize,
                pointSizeScale=cfg.pointSizeScale)
            mv.autotune(cams, reps=1)
            mv.render(cams)
            mv.render(cams)           # a replay of the captured graph
            ctx.sync()
            covered = 0
            for v, cam in enumerate(cams):
                w, h = sizes[v]
                gpu = _frame_planes(mv.frames[v], nearest)
                if nearest:
                    ref = orc.render_nearest(ocs[which[v]], mv.configuration, cam.viewMatrix, cam.projectionMatrix,
                                             w, h, threads=THREADS)
                else:
                    ref = orc.render_hqs([ocs[which[v]]], mv.configuration, cam.viewMatrix, cam.projectionMatrix,
                                         w, h, threads=THREADS)
                covered += int((ref["depth"] > 0).sum())
                try:
                    assert_planes_equal(gpu, ref)
                except AssertionError as e:
                    raise AssertionError(f"view {v} ({w}x{h}, cloud {which[v]}, nearest={nearest}): {e}") from None
            assert covered > 1_000_000
    finally:
        mv.close()
        for c in clouds:
            c.close()


def test_config2_20m_gpu_pack_equals_cpu_pack(ctx):
    """GPUPacker on the 20M-point scene (4 LOD levels, 64 coarse divisions)
    against the CPU pack: permutation, batch table, every position word,
    colours and levels, with 63-bit keys and with 30-bit keys (the
    reference's); both take the one-kernel look-back sort at this size, and
    the 30-bit keys of a 2M-point prefix the split histogram / scan / scatter
    sort."""
    from test_gpu_pool_pack import _assert_pack_equal
    pos, col = scr.gen_scene(ctx, 20_000_000, on_device=False, seed=0x5C0002, halfExtent=1000.0, numBoxes=2000)
    for bits, n in ((63, 20_000_000), (30, 20_000_000), (30, 2_000_000)):
        ref = orc.pack(pos[:n, :3], col[:n], keyBits=bits)
        cloud = scr.ComputeRasteriserPointCloud(ctx, 1, scr.DEFAULT_POINTS_PER_BATCH)
        src = scr.GPUPacker(ctx, keyBits=bits).pack(cloud, pos[:n], col[:n], sourceIndices=True)
        assert cloud.residentBatchCount == ref.batchCount and cloud.pointCount == n
        assert np.array_equal(src, ref.sourceIndices), bits
        _assert_pack_equal(cloud, ref)
        cloud.close()
