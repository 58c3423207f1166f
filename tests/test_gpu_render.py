"""HIP render passes vs the CPU oracle, bit for bit (integer planes and the
RGBA8/R32F outputs).  Marked gpu: these call libscr.so through the C ABI."""
import numpy as np
import pytest

from _pkg import orc, scr
from fixtures import jittered_cloud, lattice, solid_cube
from helpers import (MODEL, PLANES, assert_planes_equal, config1_camera, gpu_cloud, multi_file_packed, oracle_cloud,
                     render_gpu_hqs, render_gpu_nearest)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cube64():
    """64^3 lattice on [-1,1]^3, colours = xyz, packed with 30-bit keys into 10,240-point batches."""
    pos, col = orc.cube_grid(64, extent=2.0)
    return orc.pack(pos, col, pointsPerBatch=10240, lodLevels=4, coarseVoxelDivisions=64, keyBits=30)


def test_measure_atomic_rate(ctx):
    # the atomic rate the bench reports the passes against: finite, and well
    # above a billion updates per second on any current GPU
    for add in (False, True):
        r = ctx.measure_atomic_rate(add=add)
        assert np.isfinite(r) and r > 1e9


def test_fast_reciprocal_is_correctly_rounded_on_its_range(ctx):
    # batches whose clip w is proven inside [2^-125, 2^125] take the
    # three-instruction reciprocal: it must equal 1.0f / w on every float there
    mismatches, first = ctx.check_fast_reciprocal()
    assert mismatches == 0, hex(first)


@pytest.mark.parametrize("clod", [True, False])
@pytest.mark.parametrize("dither", [True, False])
def test_config1_hqs_bit_exact(ctx, cube64, clod, dither):
    cfg = scr.ComputeRasteriserConfiguration(enableCLOD=clod, enableLODDither=dither)
    cam = config1_camera()
    gpu = render_gpu_hqs(ctx, [gpu_cloud(ctx, cube64)], cfg, cam, 1280, 720)
    ref = orc.render_hqs([oracle_cloud(cube64)], cfg, cam.viewMatrix, cam.projectionMatrix, 1280, 720)
    assert ref["count"].sum() > 10000
    assert_planes_equal(gpu, ref)


def test_config1_nearest_bit_exact(ctx, cube64):
    cfg = scr.ComputeRasteriserConfiguration(mode=scr.MODE_NEAREST_POINT)
    cam = config1_camera()
    gpu = render_gpu_nearest(ctx, [gpu_cloud(ctx, cube64)], cfg, cam, 1280, 720)
    ref = orc.render_nearest(oracle_cloud(cube64), cfg, cam.viewMatrix, cam.projectionMatrix, 1280, 720)
    assert (ref["depth"] > 0).sum() > 10000
    assert_planes_equal(gpu, ref)


def test_cull_visible_set_matches_oracle(ctx, cube64):
    cfg = scr.ComputeRasteriserConfiguration()
    cam = scr.Camera.perspective((0.3, 0.2, 1.6), (1.0, 0.5, -1.0), fov=50, aspect=1.5)  # partial view
    cloud = gpu_cloud(ctx, cube64)
    render_gpu_hqs(ctx, [cloud], cfg, cam, 300, 200)
    vis = np.sort(cloud.readVisible(), order="batchIndex")
    ref = orc.cull(oracle_cloud(cube64), cfg, cam.viewMatrix, cam.projectionMatrix, 300, 200)
    assert 0 < len(ref) < cube64.batchCount
    assert vis.tobytes() == ref.tobytes()


def test_single_column_chunks_bit_exact(ctx):
    """Points on a vertical line land in one pixel column, so every chunk's
    screen box is one pixel wide and goes through the LDS tile (the width-1
    case of the tile's row/column decode)."""
    n = 6000
    y = np.linspace(-0.5, 0.5, n, dtype=np.float32)
    pos = np.stack([np.zeros(n, np.float32), y, np.sin(y * 7).astype(np.float32) * 0.1], 1)
    col = np.stack([np.linspace(0, 1, n), np.linspace(1, 0, n), np.full(n, 0.5), np.ones(n)], 1).astype(np.float32)
    packed = orc.pack(pos, col, pointsPerBatch=4096)
    cam = scr.Camera.perspective((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), fov=60, aspect=16 / 9)
    cfg = scr.ComputeRasteriserConfiguration()
    gpu = render_gpu_hqs(ctx, [gpu_cloud(ctx, packed)], cfg, cam, 1280, 720)
    ref = orc.render_hqs([oracle_cloud(packed)], cfg, cam.viewMatrix, cam.projectionMatrix, 1280, 720)
    cols = np.nonzero(ref["count"].any(axis=0))[0]
    assert len(cols) == 1 and ref["count"].sum() > 1000
    assert_planes_equal(gpu, ref)
    cfg.mode = scr.MODE_NEAREST_POINT
    gpu = render_gpu_nearest(ctx, [gpu_cloud(ctx, packed)], cfg, cam, 1280, 720)
    ref = orc.render_nearest(oracle_cloud(packed), cfg, cam.viewMatrix, cam.projectionMatrix, 1280, 720)
    assert_planes_equal(gpu, ref)


def test_multiple_files_and_model_matrix(ctx):
    packed = multi_file_packed()
    cam = scr.Camera.perspective((12.0, 9.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    for kw in (dict(), dict(minimumPointSize=1, maximumPointSize=4, pointSizeScale=30.0)):
        cfg = scr.ComputeRasteriserConfiguration(**kw)
        gpu = render_gpu_hqs(ctx, [gpu_cloud(ctx, packed, MODEL)], cfg, cam, 200, 150)
        ref = orc.render_hqs([oracle_cloud(packed, MODEL)], cfg, cam.viewMatrix, cam.projectionMatrix, 200, 150)
        assert ref["count"].sum() > 1000
        assert_planes_equal(gpu, ref)
        cfg.mode = scr.MODE_NEAREST_POINT
        gpu = render_gpu_nearest(ctx, [gpu_cloud(ctx, packed, MODEL)], cfg, cam, 200, 150)
        ref = orc.render_nearest(oracle_cloud(packed, MODEL), cfg, cam.viewMatrix, cam.projectionMatrix, 200, 150)
        assert_planes_equal(gpu, ref)


@pytest.fixture(scope="module")
def jitter():
    pos, col = jittered_cloud(20000, 0x5EED)
    return orc.pack(pos, col, pointsPerBatch=777, lodLevels=4, coarseVoxelDivisions=8)


SIZE_CASES = [
    dict(minimumPointSize=1, maximumPointSize=3, pointSizeScale=25.0),
    dict(minimumPointSize=2, maximumPointSize=9, pointSizeScale=60.0),
    dict(pointSizeMode=scr.POINT_SIZE_WORLD_SPACE, minimumPointSize=1, maximumPointSize=6, pointSizeScale=0.05),
    dict(minimumPointSize=5, maximumPointSize=5),
]


@pytest.mark.parametrize("sizes", SIZE_CASES)
def test_point_sizes_hqs_and_nearest(ctx, jitter, sizes):
    cam = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    cfg = scr.ComputeRasteriserConfiguration(**sizes)
    gpu = render_gpu_hqs(ctx, [gpu_cloud(ctx, jitter)], cfg, cam, 320, 240)
    ref = orc.render_hqs([oracle_cloud(jitter)], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    assert ref["count"].sum() > 1000
    assert_planes_equal(gpu, ref)
    cfg.mode = scr.MODE_NEAREST_POINT
    gpu = render_gpu_nearest(ctx, [gpu_cloud(ctx, jitter)], cfg, cam, 320, 240)
    ref = orc.render_nearest(oracle_cloud(jitter), cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    assert_planes_equal(gpu, ref)


ORTHO_SIZE_CASES = [
    dict(),
    dict(minimumPointSize=1, maximumPointSize=4, pointSizeScale=3.0),
    dict(pointSizeMode=scr.POINT_SIZE_WORLD_SPACE, minimumPointSize=1, maximumPointSize=6, pointSizeScale=0.1),
]


@pytest.mark.parametrize("sizes", ORTHO_SIZE_CASES)
def test_orthographic_hqs_and_nearest(ctx, jitter, sizes):
    # an orthographic projection (w = 1 everywhere): constant footprint radius
    # in both size modes (Common.metal:42-43, 49-50)
    cam = scr.Camera.orthographic((11.0, 7.0, 13.0), (0, 0, 0), height=14.0, aspect=4 / 3, near=1.0, far=40.0)
    cfg = scr.ComputeRasteriserConfiguration(**sizes)
    gpu = render_gpu_hqs(ctx, [gpu_cloud(ctx, jitter)], cfg, cam, 320, 240)
    ref = orc.render_hqs([oracle_cloud(jitter)], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    assert ref["count"].sum() > 1000
    assert_planes_equal(gpu, ref)
    cfg.mode = scr.MODE_NEAREST_POINT
    gpu = render_gpu_nearest(ctx, [gpu_cloud(ctx, jitter)], cfg, cam, 320, 240)
    ref = orc.render_nearest(oracle_cloud(jitter), cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    assert (ref["depth"] > 0).sum() > 1000
    assert_planes_equal(gpu, ref)


def test_tints_mix_and_discard(ctx, jitter):
    cam = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    cfg = scr.ComputeRasteriserConfiguration(applyTint=True, minimumPointSize=1, maximumPointSize=2,
                                             pointSizeScale=20.0)
    g = gpu_cloud(ctx, jitter)
    o = oracle_cloud(jitter)
    n = len(o.colors)
    rng = np.random.default_rng(9)
    t = rng.uniform(-0.2, 1.2, (n, 4)).astype(np.float32)
    g.setTints(t)
    o.tints = t
    gpu = render_gpu_hqs(ctx, [g], cfg, cam, 320, 240)
    ref = orc.render_hqs([o], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    assert_planes_equal(gpu, ref)
    # applyTint with no tint buffer reads the zeroed stand-in
    g.setTints(None)
    o.tints = None
    gpu = render_gpu_hqs(ctx, [g], cfg, cam, 320, 240)
    ref = orc.render_hqs([o], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    assert_planes_equal(gpu, ref)


@pytest.mark.parametrize("flags", [dict(colorizeChunks=True), dict(colorizeOverdraw=True),
                                   dict(lodBias=-2), dict(lodBias=3), dict(enableFrustumCulling=False)])
def test_config_flags(ctx, jitter, flags):
    cam = scr.Camera.perspective((6.0, 3.0, 8.0), (0, 0, 0), fov=60, aspect=1.0)
    cfg = scr.ComputeRasteriserConfiguration(**flags)
    gpu = render_gpu_hqs(ctx, [gpu_cloud(ctx, jitter)], cfg, cam, 200, 200)
    ref = orc.render_hqs([oracle_cloud(jitter)], cfg, cam.viewMatrix, cam.projectionMatrix, 200, 200)
    assert_planes_equal(gpu, ref)


def _translation(x, y, z):
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = (x, y, z)  # column 3 = translation (m[column, row])
    return m


def test_multi_cloud_occlusion_independent_of_order(ctx):
    """MultiCloudDepthOrderTests.swift:131-162: near red cube wins in both insertion orders."""
    red = orc.pack(*solid_cube((1, 0, 0, 1)))
    blue = orc.pack(*solid_cube((0, 0, 1, 1)))
    near_m, far_m = _translation(0, 0, 1), _translation(0, 0, -1)
    cfg = scr.ComputeRasteriserConfiguration(depthTolerance=0.002, backgroundColor=(0, 0, 0, 1),
                                             enableFrustumCulling=False, enableCLOD=False, enableLODDither=False,
                                             minimumPointSize=8, maximumPointSize=64, pointSizeScale=32)
    cam = scr.Camera.perspective((0, 0, 5), (0, 0, 0), fov=45, aspect=1.0, near=0.1, far=100)
    outs = []
    for order in ((red, near_m, blue, far_m), (blue, far_m, red, near_m)):
        clouds = [gpu_cloud(ctx, order[0], order[1]), gpu_cloud(ctx, order[2], order[3])]
        outs.append(render_gpu_hqs(ctx, clouds, cfg, cam, 64, 64))
        ref = orc.render_hqs([oracle_cloud(order[0], order[1]), oracle_cloud(order[2], order[3])], cfg,
                             cam.viewMatrix, cam.projectionMatrix, 64, 64)
        assert_planes_equal(outs[-1], ref)
    assert_planes_equal(outs[0], outs[1])
    texel = outs[0]["rgba8"][32, 32]
    r, b = (texel & 0xFF) / 255, ((texel >> 16) & 0xFF) / 255
    assert r > 0.6 and b < 0.2


def test_batches_hidden_behind_a_wall(ctx):
    # a dense wall of points in front of a cube: the cube's batches lie behind
    # every accepted depth of their rectangles, so the colour pass skips them
    # whole; the planes must still equal the oracle's, which draws every point
    g = np.linspace(-1.5, 1.5, 600, dtype=np.float32)
    wx, wy = np.meshgrid(g, g)
    wall = np.stack([wx.ravel(), wy.ravel(), np.full(wx.size, 0.5, np.float32)], axis=1)
    cube, _ = lattice(24, extent=1.0)
    cube = cube + np.array([0.0, 0.0, -1.5], np.float32)
    pos = np.concatenate([wall, cube]).astype(np.float32)
    col = np.concatenate([np.tile([[0.2, 0.8, 0.2, 1.0]], (len(wall), 1)),
                          np.tile([[0.9, 0.1, 0.1, 1.0]], (len(cube), 1))]).astype(np.float32)
    packed = orc.pack(pos, col, pointsPerBatch=1024)
    cam = scr.Camera.perspective((0.0, 0.0, 5.0), (0, 0, 0), fov=30, aspect=1.0, near=0.5, far=20.0)
    cfg = scr.ComputeRasteriserConfiguration(enableCLOD=False, depthTolerance=0.01)
    gpu = render_gpu_hqs(ctx, [gpu_cloud(ctx, packed)], cfg, cam, 256, 256)
    ref = orc.render_hqs([oracle_cloud(packed)], cfg, cam.viewMatrix, cam.projectionMatrix, 256, 256)
    assert (ref["count"] > 0).all()
    assert ref["red"].sum() == 51 * ref["count"].sum()   # only wall points (red byte 51) are accepted
    assert_planes_equal(gpu, ref)


def test_visible_count_words_across_render_paths(ctx, cube64):
    """Each cloud keeps two visible-count words: a render whose front launch
    holds the cull counts into the word the previous front left zero, the
    other culls (pick, a render of more clouds than the front launch carries)
    reset and count into the current one.  Any interleaving must leave the
    last cull's visible list and the same planes."""
    cfg = scr.ComputeRasteriserConfiguration()
    cam = scr.Camera.perspective((0.3, 0.2, 1.6), (1.0, 0.5, -1.0), fov=50, aspect=1.5)  # partial view
    oc = oracle_cloud(cube64)
    ref_vis = orc.cull(oc, cfg, cam.viewMatrix, cam.projectionMatrix, 300, 200)
    ref = orc.render_hqs([oc], cfg, cam.viewMatrix, cam.projectionMatrix, 300, 200)
    assert 0 < len(ref_vis) < cube64.batchCount
    clouds = [gpu_cloud(ctx, cube64) for _ in range(5)]
    one = scr.ComputeRasteriser(ctx, cfg)
    one.addPointCloud(clouds[0])
    one.resize(300, 200)
    many = scr.ComputeRasteriser(ctx, cfg)
    for c in clouds:
        many.addPointCloud(c)
    many.resize(300, 200)

    def visible(c):
        return np.sort(c.readVisible(), order="batchIndex").tobytes()

    def render_one():
        one.render(cam)
        assert visible(clouds[0]) == ref_vis.tobytes()
        out = {k: one.read(v) for k, v in PLANES.items()}
        out["depthF32"] = out["depthF32"].view(np.float32)
        assert_planes_equal(out, ref)

    render_one()
    render_one()
    one.pickPointIndex((0.0, 0.0), clouds[0], cam)
    assert visible(clouds[0]) == ref_vis.tobytes()
    render_one()
    many.render(cam)
    for c in clouds:
        assert visible(c) == ref_vis.tobytes()
    render_one()
    many.render(cam)
    render_one()


def test_pick_round_trips_through_source_indices(ctx):
    """PickPointIndexTests.swift:49-114."""
    pos = [(-0.8, 0, 0), (0.0, 0, 0), (0.8, 0, 0)]
    col = [(1, 0, 0, 1), (0, 1, 0, 1), (0, 0, 1, 1)]
    packed = orc.pack(pos, col)
    cloud = gpu_cloud(ctx, packed)
    cfg = scr.ComputeRasteriserConfiguration(backgroundColor=(0, 0, 0, 1), enableFrustumCulling=False,
                                             enableCLOD=False, enableLODDither=False)
    ras = scr.ComputeRasteriser(ctx, cfg)
    ras.addPointCloud(cloud)
    ras.resize(64, 64)
    cam = scr.Camera.perspective((0, 0, 5), (0, 0, 0), fov=45, aspect=1.0)
    ras.render(cam)
    oracle_idx = orc.nearest_cpu_reference(oracle_cloud(packed), 64, 64, cam.viewMatrix, cam.projectionMatrix)
    vp = orc.mat_mul(cam.projectionMatrix, cam.viewMatrix).astype(np.float64)
    for i, p in enumerate(pos):
        clip = vp.T @ np.array([*p, 1.0])
        ndc = (clip[0] / clip[3], clip[1] / clip[3])
        picked = ras.pickPointIndex(ndc, cloud, cam)
        assert picked is not None
        assert packed.sourceIndices[picked] == i
        want = orc.pick_search(oracle_idx["indices"], 64, 64, np.float32(ndc[0]), np.float32(ndc[1]), 10)
        assert picked == want


def test_pick_empty_space_returns_none(ctx):
    packed = orc.pack([(0.0, 0.0, 0.0)], [(1, 1, 1, 1)])
    cloud = gpu_cloud(ctx, packed)
    ras = scr.ComputeRasteriser(ctx, scr.ComputeRasteriserConfiguration(enableFrustumCulling=False,
                                                                         enableCLOD=False, enableLODDither=False))
    ras.addPointCloud(cloud)
    ras.resize(64, 64)
    cam = scr.Camera.perspective((0, 0, 5), (0, 0, 0), fov=45, aspect=1.0)
    ras.render(cam)
    assert ras.pickPointIndex((0.95, 0.95), cloud, cam) is None


@pytest.mark.parametrize("offset", [5e5, 2e4])
def test_far_offset_close_camera_matches_oracle(ctx, offset):
    """A batch far from the origin seen from about a unit away: the
    transform's partial sums dwarf the clip w, so each point's float
    projection strays from the exact projection of the batch box by pixels.
    An occluder in front covers the middle of the far batch; the colour pass
    may skip a batch only where the cull pass's bounds prove every fragment
    hidden, so the frame must equal the oracle's."""
    o = np.float32(offset)
    g = np.linspace(0.0, 1.0, 24, dtype=np.float32)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    far = np.stack([X.ravel() + o, Y.ravel() + o, Z.ravel()], 1).astype(np.float32)
    h = np.linspace(0.15, 0.85, 70, dtype=np.float32)
    A, B = np.meshgrid(h, h, indexing="ij")
    occ = np.stack([A.ravel() + o, np.full(A.size, o - np.float32(0.3), np.float32), B.ravel()], 1)
    col_far = np.tile(np.array([[0.0, 0.2, 1.0, 1.0]], np.float32), (len(far), 1))
    col_occ = np.tile(np.array([[1.0, 0.1, 0.0, 1.0]], np.float32), (len(occ), 1))
    packs = [orc.pack(far, col_far, pointsPerBatch=4096), orc.pack(occ, col_occ, pointsPerBatch=4096)]
    cam = scr.Camera.perspective((o + 0.5, o - 1.4, 0.55), (o + 0.5, o + 0.5, 0.5), (0.0, 0.0, 1.0), fov=70,
                                 aspect=16 / 9, near=0.05, far=50.0)
    for sizes in (dict(), dict(minimumPointSize=1, maximumPointSize=4, pointSizeScale=3.0)):
        cfg = scr.ComputeRasteriserConfiguration(enableCLOD=False, **sizes)
        gpu = render_gpu_hqs(ctx, [gpu_cloud(ctx, p) for p in packs], cfg, cam, 640, 360)
        ref = orc.render_hqs([oracle_cloud(p) for p in packs], cfg, cam.viewMatrix, cam.projectionMatrix, 640, 360)
        assert ref["count"].sum() > 1000 and (ref["rgba8"] & 0xFF == 255).sum() > 100  # the occluder is seen
        assert_planes_equal(gpu, ref)
        cfg.mode = scr.MODE_NEAREST_POINT
        gpu = render_gpu_nearest(ctx, [gpu_cloud(ctx, packs[0])], cfg, cam, 640, 360)
        ref = orc.render_nearest(oracle_cloud(packs[0]), cfg, cam.viewMatrix, cam.projectionMatrix, 640, 360)
        assert_planes_equal(gpu, ref)


def test_pick_rejects_a_null_configuration(ctx):
    import ctypes as C
    packed = orc.pack([(0.0, 0.0, 0.0)], [(1, 1, 1, 1)])
    cloud = gpu_cloud(ctx, packed)
    ras = scr.ComputeRasteriser(ctx, scr.ComputeRasteriserConfiguration())
    ras.addPointCloud(cloud)
    ras.resize(16, 16)
    cam = scr.Camera.perspective((0, 0, 5), (0, 0, 0), fov=45, aspect=1.0).to_c()
    out, found = C.c_uint32(), C.c_int()
    rc = ctx.lib.scr_pick(ctx.handle, None, C.byref(cam), cloud.handle, ras.frame.handle, 0.0, 0.0, 4,
                          C.byref(out), C.byref(found))
    assert rc == -1  # SCR_ERR_INVALID_ARGUMENT


def test_empty_and_all_culled_frames(ctx, cube64):
    cfg = scr.ComputeRasteriserConfiguration(backgroundColor=(0.2, 0.4, 0.6, 1.0))
    bg = orc.render_hqs([], cfg, np.eye(4, dtype=np.float32), np.eye(4, dtype=np.float32), 16, 8)
    out = render_gpu_hqs(ctx, [], cfg, config1_camera(16, 8), 16, 8)
    assert_planes_equal(out, bg)
    # camera looking away from the cloud: zero visible batches
    cam = scr.Camera.perspective((0, 0, 5), (0, 0, 10), fov=45, aspect=2.0)
    cloud = gpu_cloud(ctx, cube64)
    out = render_gpu_hqs(ctx, [cloud], cfg, cam, 64, 32)
    assert len(cloud.readVisible()) == 0
    assert out["count"].sum() == 0
    assert np.all(out["rgba8"] == bg["rgba8"][0, 0])


@pytest.mark.parametrize("iterations", [1, 2, 3])
def test_hole_fill_on_a_sparse_frame(ctx, iterations):
    # a sparse cloud leaves holes between its 1-px points; the hole fill of
    # the frame's RGBA8 output must equal the oracle's on the same image
    pos, col = jittered_cloud(2500, 0x401E)
    packed = orc.pack(pos, col, pointsPerBatch=500)
    cam = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    ras = scr.ComputeRasteriser(ctx, scr.ComputeRasteriserConfiguration())
    ras.addPointCloud(gpu_cloud(ctx, packed))
    ras.resize(160, 120)
    ras.render(cam)
    before = ras.read(scr.PLANE_RGBA8)
    holes = (before >> 24) < 128
    assert 0.2 < holes.mean() < 0.98
    ras.frame.holeFill(iterations)
    after = ras.read(scr.PLANE_RGBA8)
    want = orc.hole_fill(before, iterations)
    assert (want != before).sum() > 100
    bad = np.argwhere(after != want)
    assert bad.size == 0, (bad[:4], after[tuple(bad[0])], want[tuple(bad[0])])
    ras.frame.holeFill(0)
    assert np.array_equal(ras.read(scr.PLANE_RGBA8), after)
    # configuration.holeFillIterations runs the same passes inside the frame's
    # encode (ComputeRasteriser.swift:298-320), in both modes
    ras.configuration = scr.ComputeRasteriserConfiguration(holeFillIterations=iterations)
    ras.render(cam)
    assert np.array_equal(ras.read(scr.PLANE_RGBA8), want)
    ras.configuration = scr.ComputeRasteriserConfiguration(mode=scr.MODE_NEAREST_POINT)
    ras.render(cam)
    near = ras.read(scr.PLANE_RGBA8)
    ras.configuration.holeFillIterations = iterations
    ras.render(cam)
    assert np.array_equal(ras.read(scr.PLANE_RGBA8), orc.hole_fill(near, iterations))


def _tints(n, seed):
    rng = np.random.default_rng(seed)
    t = rng.uniform(-0.2, 1.0, (n, 4)).astype(np.float32)
    t[rng.random(n) < 0.1, 3] = -1.0        # discard sentinel
    t[rng.random(n) < 0.1, 3] = 1.5         # fully defocused: coverage 0
    return t


def test_translucent_defocus_matches_oracle(ctx, jitter):
    # tintAlphaIsCoverage: the depth pass drops the discard sentinel, the
    # colour pass blends colour * (1 - tint alpha) with no depth test, the
    # resolve returns the weighted mean with alpha 1 - e^-sum(a)
    cam = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    cfg = scr.ComputeRasteriserConfiguration(applyTint=True, tintAlphaIsCoverage=True, minimumPointSize=1,
                                             maximumPointSize=3, pointSizeScale=20.0)
    g, o = gpu_cloud(ctx, jitter), oracle_cloud(jitter)
    t = _tints(len(o.colors), 21)
    g.setTints(t)
    o.tints = t
    gpu = render_gpu_hqs(ctx, [g], cfg, cam, 320, 240)
    ref = orc.render_hqs([o], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    alpha = ref["rgba8"] >> 24
    assert ref["count"].sum() > 10000 and ((alpha > 0) & (alpha < 255)).sum() > 500
    assert_planes_equal(gpu, ref)


@pytest.mark.parametrize("coverage", [False, True])
def test_motion_blur_matches_oracle(ctx, jitter, coverage):
    # the frame remembers the previous render's view-projection: the second
    # render sweeps each splat back toward where the first camera saw it
    camA = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    camB = scr.Camera.perspective((10.2, 7.6, 13.4), (0.3, 0, 0), fov=55, aspect=4 / 3)
    cfg = scr.ComputeRasteriserConfiguration(motionBlur=0.7, motionBlurSamples=6, motionBlurMaxSpread=9.0,
                                             applyTint=coverage, tintAlphaIsCoverage=coverage,
                                             minimumPointSize=1, maximumPointSize=2, pointSizeScale=20.0)
    g, o = gpu_cloud(ctx, jitter), oracle_cloud(jitter)
    if coverage:
        t = _tints(len(o.colors), 22)
        g.setTints(t)
        o.tints = t
    ras = scr.ComputeRasteriser(ctx, cfg)
    ras.addPointCloud(g)
    ras.resize(320, 240)
    ras.render(camA)            # first render: no previous pose, no sweep
    first = {k: ras.read(v) for k, v in PLANES.items()}
    first["depthF32"] = first["depthF32"].view(np.float32)
    ref = orc.render_hqs([o], cfg, camA.viewMatrix, camA.projectionMatrix, 320, 240)
    assert_planes_equal(first, ref)
    ras.render(camB)
    gpu = {k: ras.read(v) for k, v in PLANES.items()}
    gpu["depthF32"] = gpu["depthF32"].view(np.float32)
    prev = orc.mat_mul(camA.projectionMatrix, camA.viewMatrix)
    ref = orc.render_hqs([o], cfg, camB.viewMatrix, camB.projectionMatrix, 320, 240, prev_view_proj=prev)
    still = orc.render_hqs([o], cfg, camB.viewMatrix, camB.projectionMatrix, 320, 240)
    assert (ref["count"] != still["count"]).sum() > 1000     # the sweep moved energy
    assert_planes_equal(gpu, ref)


def _displacements(n, seed, nan_every=23):
    rng = np.random.default_rng(seed)
    d = np.zeros((n, 4), np.float32)
    d[:, :3] = rng.normal(scale=0.15, size=(n, 3)).astype(np.float32)
    d[::nan_every, int(seed) % 3] = np.nan                  # removed points
    d[5::41, :3] *= np.float32(200.0)                        # thrown off the screen
    return d


@pytest.mark.parametrize("sizes", [dict(), dict(minimumPointSize=1, maximumPointSize=3, pointSizeScale=25.0)])
def test_displacement_matches_oracle(ctx, jitter, sizes):
    """applyDisplacement (DepthPass/Shaders.metal:69-74, ColorPass/Shaders.metal:
    72-78): HQS planes of displaced points, NaN ones removed, bit for bit;
    nearest mode ignores the displacements; a cloud without a buffer reads
    zeros."""
    cam = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    g, o = gpu_cloud(ctx, jitter), oracle_cloud(jitter)
    d = _displacements(len(o.colors), 31)
    g.setDisplacements(d)
    o.disp = d
    cfg = scr.ComputeRasteriserConfiguration(applyDisplacement=True, **sizes)
    gpu = render_gpu_hqs(ctx, [g], cfg, cam, 320, 240)
    ref = orc.render_hqs([o], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    still = orc.render_hqs([oracle_cloud(jitter)], scr.ComputeRasteriserConfiguration(**sizes), cam.viewMatrix,
                           cam.projectionMatrix, 320, 240)
    assert ref["count"].sum() > 1000 and (ref["count"] != still["count"]).sum() > 1000
    assert_planes_equal(gpu, ref)
    cfg.mode = scr.MODE_NEAREST_POINT
    gpu = render_gpu_nearest(ctx, [g], cfg, cam, 320, 240)
    ref = orc.render_nearest(oracle_cloud(jitter), cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    assert_planes_equal(gpu, ref)
    # two clouds, one without a buffer (zeros), in one frame
    bare = gpu_cloud(ctx, jitter, MODEL)
    ob = oracle_cloud(jitter, MODEL)
    cfg = scr.ComputeRasteriserConfiguration(applyDisplacement=True, **sizes)
    gpu = render_gpu_hqs(ctx, [g, bare], cfg, cam, 320, 240)
    ref = orc.render_hqs([o, ob], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    assert_planes_equal(gpu, ref)


def test_displaced_motion_blur_and_defocus_per_camera(ctx, jitter):
    """Motion blur sweeps from the pose and displacements the same camera key
    last rendered (ComputeRasteriser.swift:255-266, 434-480): the previous
    view-projection per key in the frame, the previous displacements per key
    in the cloud (a copy made by each motion-blurred render; the current ones
    before the first); translucent defocus on the same displaced points."""
    camA = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    camB = scr.Camera.perspective((10.4, 7.5, 13.3), (0.2, 0, 0), fov=55, aspect=4 / 3)
    cfg = scr.ComputeRasteriserConfiguration(applyDisplacement=True, motionBlur=0.6, motionBlurSamples=6,
                                             motionBlurMaxSpread=10.0, minimumPointSize=1, maximumPointSize=2,
                                             pointSizeScale=20.0)
    g = gpu_cloud(ctx, jitter)
    n = g.capacity[0] * g.capacity[1]
    D0, D1 = _displacements(n, 7, 29), _displacements(n, 8, 31)
    ras = scr.ComputeRasteriser(ctx, cfg)
    ras.addPointCloud(g)
    ras.resize(320, 240)

    def check(cam, key, disp, prev_vp, prev_disp, what):
        ras.render(cam, cameraKey=key)
        gpu = {k: ras.read(v) for k, v in PLANES.items()}
        gpu["depthF32"] = gpu["depthF32"].view(np.float32)
        o = oracle_cloud(jitter)
        o.disp, o.prevDisp = disp, prev_disp
        ref = orc.render_hqs([o], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240, prev_view_proj=prev_vp)
        assert ref["count"].sum() > 1000, what
        assert_planes_equal(gpu, ref)
        return ref

    vpA = orc.mat_mul(camA.projectionMatrix, camA.viewMatrix)
    vpB = orc.mat_mul(camB.projectionMatrix, camB.viewMatrix)
    g.setDisplacements(D0)
    check(camA, 5, D0, None, None, "first render of key 5: no velocity")
    g.setDisplacements(D1)
    moved = check(camB, 5, D1, vpA, D0, "key 5 again: camera and displacement velocity")
    still = orc.render_hqs([oracle_cloud(jitter)], cfg, camB.viewMatrix, camB.projectionMatrix, 320, 240)
    assert (moved["count"] != still["count"]).sum() > 1000
    check(camB, 6, D1, None, None, "first render of key 6")
    check(camA, 5, D1, vpB, D1, "key 5's history is its own")
    # translucent defocus of the displaced points (no motion blur, no history)
    cfg = scr.ComputeRasteriserConfiguration(applyDisplacement=True, applyTint=True, tintAlphaIsCoverage=True,
                                             minimumPointSize=1, maximumPointSize=2, pointSizeScale=20.0)
    t = _tints(n, 5)
    g.setTints(t)
    ras.configuration = cfg
    ras.render(camA, cameraKey=9)
    gpu = {k: ras.read(v) for k, v in PLANES.items()}
    gpu["depthF32"] = gpu["depthF32"].view(np.float32)
    o = oracle_cloud(jitter)
    o.disp, o.tints = D1, t
    ref = orc.render_hqs([o], cfg, camA.viewMatrix, camA.projectionMatrix, 320, 240)
    assert_planes_equal(gpu, ref)


def test_motion_blur_history_is_per_camera(ctx, jitter):
    """Two cameras alternating into one frame (the two eyes of a stereo pair,
    ComputeRasteriser.swift:255-266): each sweeps from its own previous pose."""
    eyes = [scr.Camera.perspective((11.0 + dx, 7.0, 13.0), (dx, 0, 0), fov=55, aspect=4 / 3) for dx in (-0.3, 0.3)]
    moved = [scr.Camera.perspective((10.5 + dx, 7.4, 13.2), (0.2 + dx, 0, 0), fov=55, aspect=4 / 3)
             for dx in (-0.3, 0.3)]
    cfg = scr.ComputeRasteriserConfiguration(motionBlur=0.6, motionBlurSamples=6, motionBlurMaxSpread=10.0,
                                             minimumPointSize=1, maximumPointSize=2, pointSizeScale=20.0)
    g, o = gpu_cloud(ctx, jitter), oracle_cloud(jitter)
    ras = scr.ComputeRasteriser(ctx, cfg)
    ras.addPointCloud(g)
    ras.resize(320, 240)
    for k in (0, 1):
        ras.render(eyes[k], cameraKey=100 + k)
    for k in (0, 1):
        ras.render(moved[k], cameraKey=100 + k)
        gpu = {name: ras.read(v) for name, v in PLANES.items()}
        gpu["depthF32"] = gpu["depthF32"].view(np.float32)
        prev = orc.mat_mul(eyes[k].projectionMatrix, eyes[k].viewMatrix)
        ref = orc.render_hqs([o], cfg, moved[k].viewMatrix, moved[k].projectionMatrix, 320, 240, prev_view_proj=prev)
        assert_planes_equal(gpu, ref)


COMPOSITE_CASES = [
    dict(mode=scr.COMPOSITE_DEPTH, blend=scr.BLEND_OVER, depthCompare=scr.COMPARE_GREATER, writeDepth=False),
    dict(mode=scr.COMPOSITE_DEPTH, blend=scr.BLEND_OVER, depthCompare=scr.COMPARE_GREATER_EQUAL, writeDepth=True),
    dict(mode=scr.COMPOSITE_DEPTH, blend=scr.BLEND_REPLACE, depthCompare=scr.COMPARE_LESS, writeDepth=True),
    dict(mode=scr.COMPOSITE_DEPTH, blend=scr.BLEND_REPLACE, depthCompare=scr.COMPARE_LESS_EQUAL, writeDepth=False),
    dict(mode=scr.COMPOSITE_DEPTH, blend=scr.BLEND_OVER, depthCompare=scr.COMPARE_ALWAYS, writeDepth=True),
    dict(mode=scr.COMPOSITE_OVER, blend=scr.BLEND_OVER, depthCompare=scr.COMPARE_GREATER, writeDepth=False),
    dict(mode=scr.COMPOSITE_OVER, blend=scr.BLEND_REPLACE, depthCompare=scr.COMPARE_GREATER, writeDepth=False),
]


@pytest.mark.parametrize("opts", COMPOSITE_CASES)
@pytest.mark.parametrize("pitch_pad", [0, 37])
def test_composite_onto_a_target_matches_oracle(ctx, jitter, opts, pitch_pad):
    # ComputeRasteriser.draw: the frame's colour and depth onto a caller's
    # colour + depth target with random content (pitch wider than the frame
    # in one case: the padding must stay untouched)
    w, h = 320, 240
    cam = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    cfg = scr.ComputeRasteriserConfiguration(applyTint=True, tintAlphaIsCoverage=True, minimumPointSize=1,
                                             maximumPointSize=3, pointSizeScale=20.0)
    g = gpu_cloud(ctx, jitter)
    g.setTints(_tints(len(oracle_cloud(jitter).colors), 33))
    ras = scr.ComputeRasteriser(ctx, cfg)
    ras.addPointCloud(g)
    ras.resize(w, h)
    ras.render(cam)
    rgba, depth = ras.read(scr.PLANE_RGBA8), ras.read(scr.PLANE_DEPTH_F32).view(np.float32)
    alpha = rgba >> 24
    assert (depth > 0).mean() > 0.05 and ((alpha > 0) & (alpha < 255)).sum() > 500
    pitch = w + pitch_pad
    rng = np.random.default_rng(7 + pitch_pad)
    target = rng.integers(0, 2**32, (h, pitch), dtype=np.uint64).astype(np.uint32)
    # target depths straddle the cloud's, with exact ties on some texels
    tdepth = rng.uniform(0.0, float(depth.max()) * 1.2, (h, pitch)).astype(np.float32)
    tdepth[:, :w][::5] = depth[::5]
    tb = scr.DeviceBuffer.from_array(ctx, target)
    db = scr.DeviceBuffer.from_array(ctx, tdepth)
    ras.frame.composite(tb.ptr, db.ptr, pitch, **opts)
    got = tb.to_array(np.uint32, h * pitch).reshape(h, pitch)
    gotd = db.to_array(np.float32, h * pitch).reshape(h, pitch)
    want, wantd = orc.composite(rgba, depth, target[:, :w], tdepth[:, :w], **opts)
    assert (want != target[:, :w]).sum() > 1000
    bad = np.argwhere(got[:, :w] != want)
    assert bad.size == 0, (opts, bad[:4], got[tuple(bad[0])], want[tuple(bad[0])])
    assert np.array_equal(got[:, w:], target[:, w:])
    if opts["mode"] == scr.COMPOSITE_DEPTH:
        assert np.array_equal(gotd[:, :w].view(np.uint32), wantd.view(np.uint32))
    assert np.array_equal(gotd[:, w:].view(np.uint32), tdepth[:, w:].view(np.uint32))


def test_composite_rejects_bad_arguments(ctx, jitter):
    ras = scr.ComputeRasteriser(ctx, scr.ComputeRasteriserConfiguration())
    ras.addPointCloud(gpu_cloud(ctx, jitter))
    ras.resize(64, 48)
    ras.render(scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3))
    tb = scr.DeviceBuffer(ctx, 64 * 48 * 4)
    db = scr.DeviceBuffer(ctx, 64 * 48 * 4)
    with pytest.raises(scr.ScrError, match="depth target"):
        ras.frame.composite(tb.ptr, None)
    with pytest.raises(scr.ScrError, match="pitch"):
        ras.frame.composite(tb.ptr, db.ptr, pitch=63)
    with pytest.raises(scr.ScrError, match="compare"):
        ras.frame.composite(tb.ptr, db.ptr, depthCompare=9)
    with pytest.raises(scr.ScrError, match="mode"):
        ras.frame.composite(tb.ptr, db.ptr, mode=5)
    # the colour-only mode needs no depth target
    ras.frame.composite(tb.ptr, None, mode=scr.COMPOSITE_OVER)
    ctx.sync()
