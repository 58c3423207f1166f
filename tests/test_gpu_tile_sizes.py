"""Batches whose screen rectangle sits on either side of each LDS tile
capacity: single-view depth 9,984 pixels and nearest 4,992, multi-view depth
4,096, nearest 2,048 and colour 2,528 (at the default 256 threads),
translucent-defocus depth 4,096, colour 2,528, translucent-defocus colour
2,016.  A
rectangle within a kernel's capacity takes the one-tile-per-batch path, a
larger one the per-chunk tiles, so one patch size can take different paths in
the single-view, multi-view and defocus kernels; every one must equal the
oracle (single view) or the single-view planes (multi-view), bit for bit."""
import numpy as np
import pytest

from _pkg import orc, scr
from helpers import PLANES, assert_planes_equal, gpu_cloud, oracle_cloud, render_gpu_hqs, render_gpu_nearest

pytestmark = pytest.mark.gpu

W, H = 320, 240
# screen side of the patch in pixels; the batch rectangles measure about 42, 58, 82,
# 102 and 122 pixels across (areas 1,764 / 3,364 / 6,724 / 10,404 / 14,884)
SIDES = (40, 56, 80, 100, 120)
NEAREST_PLANES = {"depth": scr.PLANE_NEAREST_DEPTH, "index": scr.PLANE_NEAREST_INDEX, "rgba8": scr.PLANE_RGBA8}


def _camera():
    return scr.Camera.perspective((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), fov=60.0, aspect=W / H,
                                  near=0.1, far=20.0)


def _patch(side_px, seed):
    """4,096 points (one batch) on a square facing the camera whose projection
    is about side_px pixels across: at distance 3 with a 60-degree vertical
    field of view, H pixels span 2 * 3 * tan(30 deg) world units."""
    units_per_px = 2.0 * 3.0 * np.tan(np.radians(30.0)) / H
    s = 0.5 * side_px * units_per_px
    rng = np.random.default_rng(seed)
    n = 4096
    pos = np.stack([rng.uniform(-s, s, n), rng.uniform(-s, s, n), rng.uniform(-0.02, 0.02, n)], 1).astype(np.float32)
    col = np.concatenate([rng.uniform(0, 1, (n, 3)), np.ones((n, 1))], 1).astype(np.float32)
    packed = orc.pack(pos, col, pointsPerBatch=4096)
    assert packed.batchCount == 1
    return packed


@pytest.fixture(scope="module")
def patches():
    return [_patch(side, 0x7157 + side) for side in SIDES]


@pytest.mark.parametrize("k", range(len(SIDES)), ids=[f"{s}px" for s in SIDES])
def test_patch_single_view_matches_oracle(ctx, patches, k):
    packed, cam = patches[k], _camera()
    cfg = scr.ComputeRasteriserConfiguration()
    gpu = render_gpu_hqs(ctx, [gpu_cloud(ctx, packed)], cfg, cam, W, H)
    ref = orc.render_hqs([oracle_cloud(packed)], cfg, cam.viewMatrix, cam.projectionMatrix, W, H)
    rows, cols = np.nonzero(ref["count"])
    side = max(rows.max() - rows.min(), cols.max() - cols.min()) + 1
    assert abs(side - SIDES[k]) <= 3 and ref["count"].sum() > 2000
    assert_planes_equal(gpu, ref)
    cfg.mode = scr.MODE_NEAREST_POINT
    gpu = render_gpu_nearest(ctx, [gpu_cloud(ctx, packed)], cfg, cam, W, H)
    ref = orc.render_nearest(oracle_cloud(packed), cfg, cam.viewMatrix, cam.projectionMatrix, W, H)
    assert_planes_equal(gpu, ref)


@pytest.mark.parametrize("k", range(len(SIDES)), ids=[f"{s}px" for s in SIDES])
def test_patch_translucent_defocus_matches_oracle(ctx, patches, k):
    packed, cam = patches[k], _camera()
    cfg = scr.ComputeRasteriserConfiguration(applyTint=True, tintAlphaIsCoverage=True, minimumPointSize=1,
                                             maximumPointSize=2, pointSizeScale=6.0)
    g, o = gpu_cloud(ctx, packed), oracle_cloud(packed)
    rng = np.random.default_rng(0xDEF0 + k)
    t = rng.uniform(0.0, 1.0, (len(o.colors), 4)).astype(np.float32)
    t[rng.random(len(t)) < 0.1, 3] = -1.0
    g.setTints(t)
    o.tints = t
    gpu = render_gpu_hqs(ctx, [g], cfg, cam, W, H)
    ref = orc.render_hqs([o], cfg, cam.viewMatrix, cam.projectionMatrix, W, H)
    assert ref["count"].sum() > 2000
    assert_planes_equal(gpu, ref)


@pytest.mark.parametrize("mode", ["hqs", "nearest"])
def test_patches_multiview_match_single_view(ctx, patches, mode):
    cfg = scr.ComputeRasteriserConfiguration()
    planes = PLANES
    if mode == "nearest":
        cfg.mode = scr.MODE_NEAREST_POINT
        planes = NEAREST_PLANES
    clouds = [gpu_cloud(ctx, p) for p in patches]
    cams = [_camera() for _ in patches]
    want = []
    for cam, cloud in zip(cams, clouds):
        ras = scr.ComputeRasteriser(ctx, cfg)
        ras.addPointCloud(cloud)
        ras.resize(W, H)
        ras.render(cam)
        want.append({name: ras.read(p) for name, p in planes.items()})
    mv = scr.MultiViewRasteriser(ctx, clouds, [(W, H)] * len(clouds), list(range(len(clouds))), cfg)
    mv.render(cams, useGraph=False)
    for v, w in enumerate(want):
        for name, p in planes.items():
            got = mv.read(v, p)
            bad = np.argwhere(got != w[name])
            assert bad.size == 0, f"view {v} plane {name}: {len(bad)} pixels differ, first {bad[:3].tolist()}"
    mv.close()


def test_full_batch_on_one_pixel_sums_do_not_carry(ctx):
    """One batch of 65,535 white points (the most a batch holds) inside one
    pixel: each channel sums to 65,535 x 255 = 16,711,425 there, just under
    2^24, and the count to 65,535, just under 2^16 -- the largest values the
    colour tile's packed fields (red | green << 24 | count << 48, blue) hold.
    Single view against the oracle and a two-view multi-view launch against
    the single view, bit for bit."""
    n = 65535  # pointsPerBatch is at most 65,535
    units_per_px = 2.0 * 3.0 * np.tan(np.radians(30.0)) / H
    # half a pixel off the screen centre (a pixel corner) in x and y: a pixel
    # centre whichever way the axes point; the points spread over 0.2 pixel
    c = 0.5 * units_per_px
    rng = np.random.default_rng(0xFEED)
    pos = np.stack([rng.uniform(c - 0.1 * units_per_px, c + 0.1 * units_per_px, n),
                    rng.uniform(c - 0.1 * units_per_px, c + 0.1 * units_per_px, n), np.zeros(n)], 1).astype(np.float32)
    packed = orc.pack(pos, np.ones((n, 4), np.float32), pointsPerBatch=n)
    assert packed.batchCount == 1
    cam = _camera()
    cfg = scr.ComputeRasteriserConfiguration(enableCLOD=False)
    ref = orc.render_hqs([oracle_cloud(packed)], cfg, cam.viewMatrix, cam.projectionMatrix, W, H)
    assert np.count_nonzero(ref["count"]) == 1 and ref["count"].max() == n
    for ch in ("red", "green", "blue"):
        assert ref[ch].max() == n * 255
    cloud = gpu_cloud(ctx, packed)
    gpu = render_gpu_hqs(ctx, [cloud], cfg, cam, W, H)
    assert_planes_equal(gpu, ref)
    ras = scr.ComputeRasteriser(ctx, cfg)
    ras.addPointCloud(cloud)
    ras.resize(W, H)
    ras.render(cam)
    want = {name: ras.read(p) for name, p in PLANES.items()}
    mv = scr.MultiViewRasteriser(ctx, [cloud], [(W, H), (W, H)], [0, 0], cfg)
    mv.render([cam, _camera()], useGraph=False)
    for v in range(2):
        for name, p in PLANES.items():
            assert np.array_equal(mv.read(v, p), want[name]), (v, name)
    mv.close()
