"""Batched multi-view rendering: every view's planes must equal what the
per-view path (ComputeRasteriser.render, one camera at a time) gives that
camera, bit for bit, in every mode, for every draw shape, eager and replayed
from a captured graph.  The per-view path is itself pinned to the oracle by
test_gpu_render.py."""
import numpy as np
import pytest

from _pkg import scr
from helpers import MODEL, multi_file_packed

pytestmark = pytest.mark.gpu

HQS_PLANES = {"depth": scr.PLANE_DEPTH, "red": scr.PLANE_RED, "green": scr.PLANE_GREEN, "blue": scr.PLANE_BLUE,
              "count": scr.PLANE_COUNT, "rgba8": scr.PLANE_RGBA8, "depthF32": scr.PLANE_DEPTH_F32}
NEAREST_PLANES = {"depth": scr.PLANE_NEAREST_DEPTH, "index": scr.PLANE_NEAREST_INDEX, "rgba8": scr.PLANE_RGBA8,
                  "depthF32": scr.PLANE_DEPTH_F32}


def _scene_cloud(ctx, n, seed, ppb=10240):
    pos, col = scr.gen_scene(ctx, n, on_device=False, seed=seed, halfExtent=250.0, numBoxes=60)
    cap = (n + ppb - 1) // ppb + 1
    cloud = scr.ComputeRasteriserPointCloud(ctx, cap, ppb)
    scr.GPUPacker(ctx, keyBits=63).pack(cloud, pos, col)
    lo, hi = pos[:, :3].min(0), pos[:, :3].max(0)
    return cloud, (lo + hi) * 0.5, float(np.linalg.norm(hi - lo))


@pytest.fixture(scope="module")
def scene(ctx):
    clouds, centres, diags = [], [], []
    for i, n in enumerate((40_000, 130_000, 300_000)):
        c, ctr, d = _scene_cloud(ctx, n, 0x5C0004 + i)
        clouds.append(c)
        centres.append(ctr)
        diags.append(d)
    return clouds, centres, diags


def _views(scene, count, seed, resolutions=scr.MV_RESOLUTIONS):
    clouds, centres, diags = scene
    cams, sizes, which = [], [], []
    for v in range(count):
        k = v % len(clouds)
        c, s = scr.sphere_views(centres[k], 1.5 * diags[k], 1, seed * 1000 + v, resolutions=resolutions)
        cams += c
        sizes += s
        which.append(k)
    return cams, sizes, which


def _per_view(ctx, scene, cfg, cams, sizes, which, planes):
    clouds = scene[0]
    out = []
    for cam, (w, h), k in zip(cams, sizes, which):
        ras = scr.ComputeRasteriser(ctx, cfg)
        ras.addPointCloud(clouds[k])
        ras.resize(w, h)
        ras.render(cam)
        out.append({name: ras.read(p) for name, p in planes.items()})
    return out


def _batched(mv, planes):
    return [{name: mv.read(v, p) for name, p in planes.items()} for v in range(mv.nViews)]


def _assert_same(got, want):
    assert len(got) == len(want)
    for v, (g, w) in enumerate(zip(got, want)):
        for name in w:
            bad = np.argwhere(g[name] != w[name])
            assert bad.size == 0, f"view {v} plane {name}: {len(bad)} pixels differ, first {bad[:3].tolist()}"


def _drawn(frames, key="count"):
    return sum(int(np.count_nonzero(f[key])) for f in frames)


def test_multiview_hqs_variable_size_clod_matches_per_view(ctx, scene):
    cfg = scr.ComputeRasteriserConfiguration(enableCLOD=True, enableLODDither=True, pointSizeMode=0,
                                             minimumPointSize=1.0, maximumPointSize=3.0, pointSizeScale=1500.0)
    cams, sizes, which = _views(scene, 9, seed=1)
    mv = scr.MultiViewRasteriser(ctx, scene[0], sizes, which, cfg)
    want = _per_view(ctx, scene, cfg, cams, sizes, which, HQS_PLANES)
    assert _drawn(want) > 0
    mv.render(cams, useGraph=False)
    _assert_same(_batched(mv, HQS_PLANES), want)
    mv.render(cams, useGraph=True)
    _assert_same(_batched(mv, HQS_PLANES), want)
    mv.close()


def test_multiview_orthographic_and_perspective_views(ctx, scene):
    # views of both projections in one launch: w = 1 for the orthographic ones
    cfg = scr.ComputeRasteriserConfiguration(minimumPointSize=1.0, maximumPointSize=3.0, pointSizeScale=2.0)
    cams, sizes, which = _views(scene, 6, seed=7)
    clouds, centres, diags = scene
    for v in (1, 4):
        k = which[v]
        eye = centres[k] + np.array([0.6, 0.5, 0.7]) * diags[k]
        w, h = sizes[v]
        cams[v] = scr.Camera.orthographic(eye, centres[k], (0.0, 0.0, 1.0), height=0.8 * diags[k], aspect=w / h,
                                          near=0.05 * diags[k], far=4.0 * diags[k])
    mv = scr.MultiViewRasteriser(ctx, clouds, sizes, which, cfg)
    want = _per_view(ctx, scene, cfg, cams, sizes, which, HQS_PLANES)
    assert all(int(np.count_nonzero(want[v]["count"])) > 100 for v in (1, 4))
    mv.render(cams)
    _assert_same(_batched(mv, HQS_PLANES), want)
    mv.close()


def test_multiview_hqs_one_pixel_graph_replays(ctx, scene):
    """Graph replays, with other eager work and synchronisation between them,
    each reproduce the eager result; changing cameras between replays works
    through the re-uploaded view table."""
    cfg = scr.ComputeRasteriserConfiguration(minimumPointSize=1.0, maximumPointSize=1.0)
    cams_a, sizes, which = _views(scene, 8, seed=2)
    cams_b, _, _ = _views(scene, 8, seed=3)
    cams_b = [scr.Camera(b.viewMatrix, a.projectionMatrix) for a, b in zip(cams_a, cams_b)]
    want_a = _per_view(ctx, scene, cfg, cams_a, sizes, which, HQS_PLANES)
    want_b = _per_view(ctx, scene, cfg, cams_b, sizes, which, HQS_PLANES)
    mv = scr.MultiViewRasteriser(ctx, scene[0], sizes, which, cfg)
    for rep in range(3):
        mv.render(cams_a, useGraph=True)
        _assert_same(_batched(mv, HQS_PLANES), want_a)
        _per_view(ctx, scene, cfg, cams_b[:1], sizes[:1], which[:1], {"count": scr.PLANE_COUNT})
        mv.render(cams_b, useGraph=True)
        _assert_same(_batched(mv, HQS_PLANES), want_b)
    mv.close()


def test_multiview_nearest_matches_per_view(ctx, scene):
    cfg = scr.ComputeRasteriserConfiguration(mode=scr.MODE_NEAREST_POINT, minimumPointSize=1.0,
                                             maximumPointSize=3.0, pointSizeScale=1500.0)
    cams, sizes, which = _views(scene, 7, seed=4)
    want = _per_view(ctx, scene, cfg, cams, sizes, which, NEAREST_PLANES)
    mv = scr.MultiViewRasteriser(ctx, scene[0], sizes, which, cfg)
    for graph in (False, True, True):
        mv.render(cams, useGraph=graph)
        _assert_same(_batched(mv, NEAREST_PLANES), want)
    mv.close()


def test_multiview_every_shape_gives_the_same_planes(ctx, scene):
    small = ((256, 256), (512, 384))
    cams, sizes, which = _views(scene, 6, seed=5, resolutions=small)
    for mode, planes in ((scr.MODE_HIGH_QUALITY_AVERAGE, HQS_PLANES), (scr.MODE_NEAREST_POINT, NEAREST_PLANES)):
        for maxsize in (1.0, 3.0):
            cfg = scr.ComputeRasteriserConfiguration(mode=mode, maximumPointSize=maxsize, pointSizeScale=1500.0,
                                                     enableCLOD=True)
            want = _per_view(ctx, scene, cfg, cams, sizes, which, planes)
            mv = scr.MultiViewRasteriser(ctx, scene[0], sizes, which, cfg)
            for threads, ppt in scr.MV_SHAPES:
                for splits in scr.MV_SPLITS:
                    for wave in (0, 1):
                        sh = {"threads": threads, "pointsPerThread": ppt, "splits": splits, "waveTiles": wave}
                        mv.tuning = {"depth": sh, "color": sh, "nearest": sh}
                        mv.render(cams, useGraph=(splits == 2))
                        _assert_same(_batched(mv, planes), want)
            mv.close()


def test_multiview_autotune_installs_a_valid_shape(ctx, scene):
    cfg = scr.ComputeRasteriserConfiguration()
    cams, sizes, which = _views(scene, 4, seed=6, resolutions=((256, 256), (512, 384)))
    mv = scr.MultiViewRasteriser(ctx, scene[0], sizes, which, cfg)
    t = mv.autotune(cams, reps=1)
    assert t == mv.tuning
    for ph in ("depth", "color"):
        assert (t[ph]["threads"], t[ph]["pointsPerThread"]) in scr.MV_SHAPES
        assert t[ph]["splits"] in scr.MV_SPLITS and t[ph]["waveTiles"] in (0, 1)
    want = _per_view(ctx, scene, cfg, cams, sizes, which, HQS_PLANES)
    mv.render(cams)
    _assert_same(_batched(mv, HQS_PLANES), want)
    mv.close()


def test_multiview_empty_cloud_and_shared_cloud(ctx, scene):
    """A view over a cloud with no resident batches gets the cleared frame;
    several views may share one cloud."""
    empty = scr.ComputeRasteriserPointCloud(ctx, 3, 10240)
    clouds = [scene[0][0], empty]
    cams, sizes, _ = _views(scene, 3, seed=7, resolutions=((256, 256),))
    which = [0, 1, 0]
    cfg = scr.ComputeRasteriserConfiguration()
    mv = scr.MultiViewRasteriser(ctx, clouds, sizes, which, cfg)
    mv.render(cams)
    got = _batched(mv, HQS_PLANES)
    assert not got[1]["count"].any() and not got[1]["depth"].any()
    assert (got[1]["rgba8"] == 0).all()
    want = _per_view(ctx, (clouds,), cfg, cams, sizes, which, HQS_PLANES)
    _assert_same(got, want)
    mv.close()


def test_multiview_multi_file_clouds_with_model_matrices(ctx, scene):
    """Per-view composed file tables: clouds with several RasterFiles and a
    model matrix render the same batched as per view."""
    packed = multi_file_packed()
    tiled = scr.ComputeRasteriserPointCloud.from_packed(ctx, packed)
    tiled.worldMatrix = MODEL
    clouds = [tiled, scene[0][1]]
    cams, sizes, _ = _views(scene, 5, seed=8, resolutions=((256, 256), (512, 384)))
    cams[0] = scr.Camera.perspective((12.0, 9.0, 13.0), (0, 0, 0), fov=55, aspect=1.0)
    cams[2] = scr.Camera.perspective((-10.0, 8.0, 11.0), (0, 0, 0), fov=60, aspect=4 / 3)
    which = [0, 1, 0, 1, 0]
    cams[4] = scr.Camera.perspective((9.0, -12.0, 7.0), (0, 0, 0), fov=50, aspect=1.0)
    for kw in (dict(), dict(minimumPointSize=1, maximumPointSize=4, pointSizeScale=30.0)):
        cfg = scr.ComputeRasteriserConfiguration(**kw)
        want = _per_view(ctx, (clouds,), cfg, cams, sizes, which, HQS_PLANES)
        assert _drawn([want[0], want[2], want[4]]) > 500
        mv = scr.MultiViewRasteriser(ctx, clouds, sizes, which, cfg)
        mv.render(cams)
        _assert_same(_batched(mv, HQS_PLANES), want)
        mv.close()


def test_multiview_argument_errors(ctx, scene):
    clouds = scene[0]
    f = scr.Frame(ctx, 64, 64)
    lib = ctx.lib
    import ctypes as C
    h = C.c_void_p()
    ch = (C.c_void_p * 1)(clouds[0].handle.value)
    fh = (C.c_void_p * 2)(f.handle.value, f.handle.value)
    vc = np.zeros(2, np.int32)
    assert lib.scr_multiview_create(ctx.handle, ch, 1, fh, vc.ctypes.data, 2, C.byref(h)) == -1
    vc1 = np.array([3], np.int32)
    assert lib.scr_multiview_create(ctx.handle, ch, 1, fh, vc1.ctypes.data, 1, C.byref(h)) == -1
    mv = scr.MultiViewRasteriser(ctx, clouds[:1], [(64, 64)], [0])
    with pytest.raises(scr.ScrError):
        mv.tuning = {"depth": {"threads": 192}}
    with pytest.raises(scr.ScrError):
        mv.tuning = {"color": {"splits": 3}}
    mv.close()
    # the weighted-blend passes and the hole fill belong to scr_render
    for kw in (dict(holeFillIterations=1), dict(motionBlur=0.5), dict(applyTint=True, tintAlphaIsCoverage=True),
               dict(applyDisplacement=True)):
        mv = scr.MultiViewRasteriser(ctx, clouds[:1], [(64, 64)], [0], scr.ComputeRasteriserConfiguration(**kw))
        cam = scr.Camera.perspective((5.0, 4.0, 6.0), (0, 0, 0), fov=50, aspect=1.0)
        with pytest.raises(scr.ScrError) as e:
            mv.render([cam], useGraph=False)
        assert e.value.code == -8
        mv.close()
    f.close()
