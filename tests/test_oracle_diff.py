"""Pins for the oracle's differentiable HQS (no reference counterpart): the
fixed-point conversion against closed forms, the forward against linearity
and the HQS visibility it extends, and the FP64 backward against central
finite differences of the exact quadratic loss (steps on the 2^-16 grid, so
the differences carry no rounding of the parameters)."""
import numpy as np

from _pkg import orc, scr
from fixtures import jittered_cloud

GRID = 65536.0


def test_fixed_point_closed_forms():
    f = orc.lib().orc_diff_fixed
    assert f(0.5) == 32768 and f(1.0) == 65536 and f(-0.25) == -16384 and f(0.0) == 0
    assert f(1.0 / 3.0) == 21845                          # 21845.33 rounds down
    assert f(0.5 / GRID) == 0 and f(1.5 / GRID) == 2      # ties to even
    assert f(2.5 / GRID) == 2 and f(-1.5 / GRID) == -2
    assert f(float("nan")) == 0
    assert f(1e30) == int(np.float32(9.0e18)) and f(-1e30) == -int(np.float32(9.0e18))
    p = orc.lib().orc_diff_pixel
    assert p(3 * 32768, 3, 0.25) == 0.5 and p(0, 0, 0.25) == 0.25
    assert p(65536, 3, 0.0) == np.float32(1.0 / 3.0)


def _scene(n=3000, seed=0x51):
    pos, col = jittered_cloud(n, seed)
    packed = orc.pack(pos, col, pointsPerBatch=256)
    cloud = orc.Cloud.from_packed(packed)
    cam = scr.Camera.perspective((11.0, 8.0, 12.0), (0, 0, 0), fov=50, aspect=1.25)
    return packed, cloud, cam


def _grid_params(npts, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros((npts, 4), np.float32)
    p[:, :3] = np.rint(rng.uniform(-0.2, 1.2, (npts, 3)) * GRID) / GRID
    return p


def test_single_point_sum_and_image():
    packed = orc.pack([(0.0, 0.0, 0.5)], [(1, 1, 1, 1)], pointsPerBatch=1)
    cloud = orc.Cloud.from_packed(packed)
    cfg = scr.ComputeRasteriserConfiguration(enableFrustumCulling=False, enableCLOD=False)
    eye = np.eye(4, dtype=np.float32)
    params = np.array([[0.3, -0.7, 2.0, 0.0]], np.float32)
    out = orc.render_diff(cloud, cfg, eye, eye, 8, 8, params)
    assert out["count"][3, 4] == 1 and out["count"].sum() == 1
    assert list(out["sums"][3, 4]) == [int(np.rint(np.float32(0.3) * np.float32(GRID))), -45875, 131072]
    assert out["image"][3, 4, 0] == np.float32(int(np.rint(np.float32(0.3) * np.float32(GRID))) / GRID)
    assert out["image"][3, 4, 2] == 2.0
    assert np.all(out["image"][0, 0] == 0.0)


def test_visibility_is_hqs_and_sums_are_linear():
    packed, cloud, cam = _scene()
    cfg = scr.ComputeRasteriserConfiguration(minimumPointSize=1, maximumPointSize=3, pointSizeScale=25.0)
    W, H = 96, 72
    hqs = orc.render_hqs([cloud], cfg, cam.viewMatrix, cam.projectionMatrix, W, H)
    a = _grid_params(len(cloud.xyzLow), 1)
    b = _grid_params(len(cloud.xyzLow), 2)
    ra = orc.render_diff(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, a)
    rb = orc.render_diff(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, b, threads=3)
    rab = orc.render_diff(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, a + b)
    assert np.array_equal(ra["depth"], hqs["depth"]) and np.array_equal(ra["count"], hqs["count"])
    assert hqs["count"].sum() > 500
    assert np.array_equal(rab["sums"], ra["sums"] + rb["sums"])
    # a constant colour c gives sums = count * fixed(c) and image = c where covered
    c = np.zeros_like(a)
    c[:, :3] = (0.25, 0.5, 0.75)
    rc = orc.render_diff(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, c)
    cnt = rc["count"].astype(np.int64)
    for k, v in enumerate((16384, 32768, 49152)):
        assert np.array_equal(rc["sums"][..., k], cnt * v)
    cov = cnt > 0
    assert np.all(rc["image"][cov] == np.array([0.25, 0.5, 0.75], np.float32))


def _exact_image(sums, count):
    img = np.zeros(sums.shape, np.float64)
    cov = count > 0
    img[cov] = sums[cov] / (GRID * count[cov][:, None].astype(np.float64))
    return img


def test_backward_matches_central_differences_of_exact_loss():
    packed, cloud, cam = _scene(2500, 0x52)
    cfg = scr.ComputeRasteriserConfiguration(minimumPointSize=1, maximumPointSize=3, pointSizeScale=25.0,
                                             enableCLOD=False)
    W, H = 80, 64
    npts = len(cloud.xyzLow)
    theta = _grid_params(npts, 3)
    rng = np.random.default_rng(4)
    target = rng.uniform(0, 1, (H, W, 3))
    fwd = orc.render_diff(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, theta)
    img = _exact_image(fwd["sums"], fwd["count"])
    dLdI = (2.0 * (img - target)).astype(np.float32)
    dLdI[fwd["count"] == 0] = 0.0
    grad = np.zeros((npts, 4), np.float64)
    orc.diff_backward(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, fwd["depth"], fwd["count"], dLdI, grad)
    touched = np.nonzero(np.abs(grad[:, :3]).sum(1) > 0)[0]
    assert len(touched) > 200
    h = 64.0 / GRID   # 64 grid steps: the perturbation is exact in fixed point

    def loss(t):
        o = orc.render_diff(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, t)
        i = _exact_image(o["sums"], o["count"])
        return i, o["count"]

    picks = list(rng.choice(touched, 12, replace=False)) + list(rng.choice(np.setdiff1d(np.arange(npts), touched), 3))
    for p in picks:
        k = int(rng.integers(3))
        tp, tm = theta.copy(), theta.copy()
        tp[p, k] += np.float32(h)
        tm[p, k] -= np.float32(h)
        ip, cp = loss(tp)
        im, cm = loss(tm)
        assert np.array_equal(cp, fwd["count"]) and np.array_equal(cm, fwd["count"])
        cov = fwd["count"] > 0
        d = (ip - im)[cov]
        s = (ip + im - 2.0 * target)[cov]
        fd = float((d * s).sum()) / (2.0 * h)
        g = grad[p, k]
        assert abs(fd - g) <= 1e-5 * max(abs(g), 1e-3), (p, k, fd, g)


def test_backward_mask_selects_points_and_views_accumulate():
    packed, cloud, cam = _scene(1500, 0x53)
    cfg = scr.ComputeRasteriserConfiguration()
    W, H = 64, 48
    npts = len(cloud.xyzLow)
    theta = _grid_params(npts, 5)
    fwd = orc.render_diff(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, theta)
    dLdI = np.random.default_rng(6).normal(size=(H, W, 3)).astype(np.float32)
    full = np.zeros((npts, 4))
    orc.diff_backward(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, fwd["depth"], fwd["count"], dLdI, full)
    mask = np.zeros(npts, np.uint8)
    mask[::7] = 1
    part = np.zeros((npts, 4))
    orc.diff_backward(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, fwd["depth"], fwd["count"], dLdI, part,
                      mask=mask)
    assert np.array_equal(part[mask == 1], full[mask == 1]) and not part[mask == 0].any()
    twice = full.copy()
    orc.diff_backward(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, fwd["depth"], fwd["count"], dLdI, twice)
    assert np.allclose(twice, 2 * full, rtol=0, atol=1e-15)
    # per pixel, the gradient mass is dLdI (each fragment carries 1/count of it)
    assert np.isclose(full[:, :3].sum(), float(dLdI[fwd["count"] > 0].astype(np.float64).sum()), rtol=1e-9)
