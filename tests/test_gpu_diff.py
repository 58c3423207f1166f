"""Differentiable HQS on the GPU through the C ABI: the forward's integer
planes and float image bit for bit against the oracle, the backward against
the oracle's FP64 gradient, and the torch autograd.Function harness against
the FP64 gradient and central finite differences (h = 1e-3) of a re-rendered
loss, plus a short Adam fit."""
import numpy as np
import pytest

from _pkg import orc, scr

pytestmark = pytest.mark.gpu

W, H = 160, 120


@pytest.fixture(scope="module")
def diffscene(ctx):
    n = 90_000
    pos, col = scr.gen_scene(ctx, n, on_device=False, seed=0x5C0005, halfExtent=200.0, numBoxes=40,
                             quantizeColors=0)
    cloud = scr.ComputeRasteriserPointCloud(ctx, 1, 4096)
    src = scr.GPUPacker(ctx, keyBits=63).pack(cloud, pos, col, sourceIndices=True)
    ppb, nb = cloud.pointsPerBatch, cloud.batchCount
    batches = cloud.readPool(scr.POOL_BATCHES)
    # ground-truth float colours in pool order (batch b of the pack sits in slot b)
    gt = np.zeros((nb * ppb, 4), np.float32)
    k = 0
    for s in range(nb):
        m = int(batches["numPoints"][s])
        first = int(batches["firstPoint"][s])
        gt[first:first + m, :3] = col[src[k:k + m], :3]
        k += m
    assert k == n
    pool_rgb = cloud.readPool(scr.POOL_COLORS)
    live = np.zeros(nb * ppb, bool)
    for s in range(nb):
        live[int(batches["firstPoint"][s]):int(batches["firstPoint"][s]) + int(batches["numPoints"][s])] = True
    trunc = (np.clip(gt[:, :3], 0, 1) * np.float32(255)).astype(np.uint32)
    assert np.array_equal((pool_rgb[live] & 0xFFFFFF), (trunc[live] * [1, 256, 65536]).sum(1).astype(np.uint32))
    lo, hi = pos[:, :3].min(0), pos[:, :3].max(0)
    cams, _ = scr.sphere_views((lo + hi) / 2, 1.3 * float(np.linalg.norm(hi - lo)), 4, 0x5C0005,
                               resolutions=((W, H),))
    ocloud = orc.Cloud(batches, scr.identity_files(1), cloud.readPool(scr.POOL_XYZ_LOW),
                       cloud.readPool(scr.POOL_XYZ_MED), cloud.readPool(scr.POOL_XYZ_HIGH), pool_rgb,
                       cloud.readPool(scr.POOL_LEVELS))
    return cloud, ocloud, gt, cams


CFGS = [
    dict(minimumPointSize=1.0, maximumPointSize=1.0, enableCLOD=False),
    dict(minimumPointSize=1.0, maximumPointSize=3.0, pointSizeScale=300.0, enableCLOD=True),
    dict(minimumPointSize=2.0, maximumPointSize=2.0, enableCLOD=True, enableLODDither=False),
]


def _dev(ctx, a):
    return scr.DeviceBuffer.from_array(ctx, np.ascontiguousarray(a))


@pytest.mark.parametrize("kw", CFGS)
def test_diff_forward_bit_exact(ctx, diffscene, kw):
    cloud, ocloud, gt, cams = diffscene
    cfg = scr.ComputeRasteriserConfiguration(**kw)
    r = scr.DifferentiableHQS(ctx, cloud, [(W, H)] * len(cams), cfg)
    params = gt.copy()
    params[::5, :3] += np.float32(0.37)      # off-grid values, some outside [0, 1]
    dparams = _dev(ctx, params)
    img = scr.DeviceBuffer(ctx, W * H * 12)
    for v, cam in enumerate(cams):
        r.forward(v, cam, dparams.ptr, img.ptr)
        got = img.to_array(np.float32, W * H * 3).reshape(H, W, 3)
        ref = orc.render_diff(ocloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, params, threads=4)
        assert ref["count"].sum() > 1000
        assert np.array_equal(r.read(v, scr.PLANE_DEPTH), ref["depth"])
        assert np.array_equal(r.read(v, scr.PLANE_COUNT), ref["count"])
        assert np.array_equal(r.sums(v), ref["sums"])
        assert np.array_equal(got.view(np.uint32), ref["image"].view(np.uint32))
    r.close()


def test_diff_backward_matches_fp64_reference(ctx, diffscene):
    cloud, ocloud, gt, cams = diffscene
    cfg = scr.ComputeRasteriserConfiguration(**CFGS[1])
    r = scr.DifferentiableHQS(ctx, cloud, [(W, H)] * len(cams), cfg)
    npool = r.poolPoints
    params = np.full((npool, 4), 0.5, np.float32)
    dparams = _dev(ctx, params)
    img = scr.DeviceBuffer(ctx, W * H * 12)
    grad = _dev(ctx, np.zeros((npool, 4), np.float64))
    ref = np.zeros((npool, 4), np.float64)
    rng = np.random.default_rng(7)
    for v, cam in enumerate(cams):
        r.forward(v, cam, dparams.ptr, img.ptr)
        g = rng.normal(size=(H, W, 3)).astype(np.float32)
        dg = _dev(ctx, g)
        r.backward(v, cam, dg.ptr, grad.ptr)
        o = orc.render_diff(ocloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, params, threads=4)
        orc.diff_backward(ocloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, o["depth"], o["count"], g, ref)
    got = grad.to_array(np.float64, npool * 4).reshape(npool, 4)
    nz = ref != 0
    assert nz.sum() > 3000
    assert np.array_equal(got == 0, ~nz)
    rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    assert rel.max() <= 1e-12, rel.max()
    r.close()


@pytest.mark.parametrize("kw", CFGS)
def test_batched_views_equal_per_view_calls(ctx, diffscene, kw):
    """One launch per phase for all views gives the per-view images, planes
    and FP64 gradients bit for bit."""
    cloud, ocloud, gt, cams = diffscene
    cfg = scr.ComputeRasteriserConfiguration(**kw)
    sizes = [(W, H)] * len(cams)
    one = scr.DifferentiableHQS(ctx, cloud, sizes, cfg)
    many = scr.DifferentiableHQS(ctx, cloud, sizes, cfg, batched=True)
    npool = one.poolPoints
    params = gt.copy()
    params[::3, :3] *= np.float32(0.8)
    dparams = _dev(ctx, params)
    nv = len(cams)
    imgs = [scr.DeviceBuffer(ctx, W * H * 12) for _ in range(2 * nv)]
    one.forward_all(cams, dparams.ptr, [b.ptr for b in imgs[:nv]])
    many.forward_all(cams, dparams.ptr, [b.ptr for b in imgs[nv:]])
    for v in range(nv):
        a = imgs[v].to_array(np.uint32, W * H * 3)
        b = imgs[nv + v].to_array(np.uint32, W * H * 3)
        assert np.array_equal(a, b), v
        assert np.array_equal(one.sums(v), many.sums(v))
        for pl in (scr.PLANE_DEPTH, scr.PLANE_COUNT):
            assert np.array_equal(one.read(v, pl), many.read(v, pl))
    rng = np.random.default_rng(9)
    dls = [_dev(ctx, rng.normal(size=(H, W, 3)).astype(np.float32)) for _ in range(nv)]
    ga = _dev(ctx, np.zeros((npool, 4), np.float64))
    gb = _dev(ctx, np.zeros((npool, 4), np.float64))
    one.backward_all(cams, [d.ptr for d in dls], ga.ptr)
    many.backward_all(cams, [d.ptr for d in dls], gb.ptr)
    a = ga.to_array(np.float64, npool * 4)
    b = gb.to_array(np.float64, npool * 4)
    assert np.count_nonzero(a) > 3000
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    one.close()
    many.close()


def _torch_setup(ctx, diffscene, cfgkw):
    from _pkg import PKG_DIR
    import importlib.util
    spec = importlib.util.spec_from_file_location("scr_torch_diff", PKG_DIR / "torch_diff.py")
    td = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(td)
    cloud, ocloud, gt, cams = diffscene
    cfg = scr.ComputeRasteriserConfiguration(**cfgkw)
    r = scr.DifferentiableHQS(ctx, cloud, [(W, H)] * len(cams), cfg, batched=True)
    return td, r, cfg


def test_torch_function_gradient_vs_fp64_and_finite_differences(ctx, diffscene):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch has no GPU")
    cloud, ocloud, gt, cams = diffscene
    td, r, cfg = _torch_setup(ctx, diffscene, CFGS[1])
    dev = torch.device("cuda:0")
    npool = r.poolPoints
    render = td.HQSRender(r, cams)
    with torch.no_grad():
        target = render(torch.from_numpy(gt).to(dev))
        target = target + 0.02 * torch.randn(target.shape, generator=torch.Generator(device=dev).manual_seed(1),
                                             device=dev)
    theta = torch.full((npool, 4), 0.5, dtype=torch.float32, device=dev, requires_grad=True)
    seen = []
    images = render(theta)
    images.register_hook(lambda g: seen.append(g.detach().clone()))
    loss = ((images - target) ** 2).mean()
    loss.backward()
    g32 = theta.grad.detach().cpu().numpy()
    dLdI = seen[0].cpu().numpy()
    ref = np.zeros((npool, 4), np.float64)
    th = theta.detach().cpu().numpy()
    for v, cam in enumerate(cams):
        o = orc.render_diff(ocloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, th, threads=4)
        orc.diff_backward(ocloud, cfg, cam.viewMatrix, cam.projectionMatrix, W, H, o["depth"], o["count"], dLdI[v],
                          ref)
    nz = ref[:, :3] != 0
    assert nz.sum() > 3000
    rel = np.abs(g32[:, :3][nz] - ref[:, :3][nz]) / np.abs(ref[:, :3][nz])
    assert rel.max() <= 1e-5, rel.max()
    # central differences of the re-rendered loss, h = 1e-3, in FP64
    rng = np.random.default_rng(11)
    cand = np.nonzero(np.abs(ref[:, :3]).max(1) > 0.1 * np.abs(ref[:, :3]).max())[0]
    h = 1e-3
    tgt = target.double()
    with torch.no_grad():
        for p in rng.choice(cand, 12, replace=False):
            k = int(np.argmax(np.abs(ref[p, :3])))
            tp = theta.detach().clone()
            tm = theta.detach().clone()
            tp[p, k] += h
            tm[p, k] -= h
            ip, im = render(tp).double(), render(tm).double()
            dl = (((ip - im) * (ip + im - 2 * tgt)).sum() / ip.numel()).item()
            step = (orc.lib().orc_diff_fixed(float(tp[p, k])) - orc.lib().orc_diff_fixed(float(tm[p, k]))) / 65536.0
            fd = dl / step
            assert abs(fd - ref[p, k]) <= 2e-2 * abs(ref[p, k]), (p, k, fd, ref[p, k])
    r.close()


def test_torch_adam_fit_reduces_loss(ctx, diffscene):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch has no GPU")
    cloud, ocloud, gt, cams = diffscene
    td, r, cfg = _torch_setup(ctx, diffscene, CFGS[0])
    dev = torch.device("cuda:0")
    render = td.HQSRender(r, cams)
    with torch.no_grad():
        target = render(torch.from_numpy(gt).to(dev))
    theta = torch.full((r.poolPoints, 4), 0.5, dtype=torch.float32, device=dev, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=1e-2)
    losses = []
    for _ in range(25):
        opt.zero_grad()
        loss = ((render(theta) - target) ** 2).mean()
        loss.backward()
        opt.step()
        losses.append(loss.item())
    assert losses[-1] < 0.5 * losses[0], losses
    r.close()


def test_diff_rejects_nearest_and_tint_modes(ctx, diffscene):
    cloud, ocloud, gt, cams = diffscene
    for kw in (dict(mode=scr.MODE_NEAREST_POINT), dict(applyTint=True), dict(holeFillIterations=2),
               dict(motionBlur=0.5), dict(applyDisplacement=True)):
        r = scr.DifferentiableHQS(ctx, cloud, [(W, H)], scr.ComputeRasteriserConfiguration(**kw))
        p = scr.DeviceBuffer(ctx, r.poolPoints * 16)
        img = scr.DeviceBuffer(ctx, W * H * 12)
        with pytest.raises(scr.ScrError) as e:
            r.forward(0, cams[0], p.ptr, img.ptr)
        assert e.value.code == -8
        r.close()


def test_config5_full_size_gradient_checks(ctx):
    """5M points of the 2 km scene with float colours, 32 views of 512x512:
    forward planes equal the oracle's; the torch-harness gradient (FP32)
    against the oracle's FP64 gradient and against central differences
    (h = 1e-3, FP64 loss differences of re-rendered images) on 1,000 seeded
    random points that the views see."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch has no GPU")
    from _pkg import PKG_DIR
    import importlib.util
    spec = importlib.util.spec_from_file_location("scr_torch_diff", PKG_DIR / "torch_diff.py")
    td = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(td)
    n, size, nv = 5_000_000, 512, 32
    pos, col = scr.gen_scene(ctx, n, on_device=False, seed=0x5C0005, halfExtent=1000.0, numBoxes=2000,
                             quantizeColors=0)
    cloud = scr.ComputeRasteriserPointCloud(ctx, 1, 10240)
    src = scr.GPUPacker(ctx, keyBits=63).pack(cloud, pos, col, sourceIndices=True)
    batches = cloud.readPool(scr.POOL_BATCHES)
    npool = cloud.batchCount * cloud.pointsPerBatch
    gt = np.zeros((npool, 4), np.float32)
    k = 0
    for s in range(cloud.batchCount):
        m, first = int(batches["numPoints"][s]), int(batches["firstPoint"][s])
        gt[first:first + m, :3] = col[src[k:k + m], :3]
        k += m
    lo, hi = pos[:, :3].min(0), pos[:, :3].max(0)
    del pos, col, src
    cams, _ = scr.sphere_views((lo + hi) / 2, 1.2 * float(np.linalg.norm(hi - lo)), nv, 0x5C0005,
                               resolutions=((size, size),))
    cfg = scr.ComputeRasteriserConfiguration(minimumPointSize=1.0, maximumPointSize=3.0, pointSizeScale=1500.0)
    r = scr.DifferentiableHQS(ctx, cloud, [(size, size)] * nv, cfg, batched=True)
    dev = torch.device("cuda:0")
    render = td.HQSRender(r, cams)
    with torch.no_grad():
        clean = render(torch.from_numpy(gt).to(dev))
        target = clean + 0.02 * torch.randn(clean.shape, generator=torch.Generator(device=dev).manual_seed(0x5C0005),
                                            device=dev)
    theta = torch.full((npool, 4), 0.5, dtype=torch.float32, device=dev, requires_grad=True)
    seen = []
    images = render(theta)
    images.register_hook(lambda g: seen.append(g.detach().clone()))
    ((images - target) ** 2).mean().backward()
    g32 = theta.grad.detach().cpu().numpy().astype(np.float64)
    dLdI = seen[0].cpu().numpy()
    ocloud = orc.Cloud(batches, scr.identity_files(1), cloud.readPool(scr.POOL_XYZ_LOW),
                       cloud.readPool(scr.POOL_XYZ_MED), cloud.readPool(scr.POOL_XYZ_HIGH),
                       cloud.readPool(scr.POOL_COLORS), cloud.readPool(scr.POOL_LEVELS))
    touched = np.nonzero(np.abs(g32[:, :3]).max(1) > 0)[0]
    rng = np.random.default_rng(0x5C0005)
    picks = np.sort(rng.choice(touched, 1000, replace=False))
    mask = np.zeros(npool, np.uint8)
    mask[picks] = 1
    th = theta.detach().cpu().numpy()
    ref = np.zeros((npool, 4), np.float64)
    for v, cam in enumerate(cams):
        o = orc.render_diff(ocloud, cfg, cam.viewMatrix, cam.projectionMatrix, size, size, th, threads=16)
        assert np.array_equal(r.read(v, scr.PLANE_COUNT), o["count"])
        assert np.array_equal(r.sums(v), o["sums"])
        orc.diff_backward(ocloud, cfg, cam.viewMatrix, cam.projectionMatrix, size, size, o["depth"], o["count"],
                          dLdI[v], ref, mask=mask)
    gp, rp = g32[picks, :3], ref[picks, :3]
    nz = rp != 0
    rel64 = float((np.abs(gp[nz] - rp[nz]) / np.abs(rp[nz])).max())
    assert rel64 <= 1e-5, rel64
    h = 1e-3
    tgt = target.double()
    f = orc.lib().orc_diff_fixed
    worst = 0.0
    gmax = float(np.abs(rp).max())
    tp = theta.detach().clone()
    with torch.no_grad():
        for p in picks:
            kk = int(np.argmax(np.abs(ref[p, :3])))
            base = float(th[p, kk])
            tp[p, kk] = base + h
            ip = render(tp).double()
            tp[p, kk] = base - h
            im = render(tp).double()
            tp[p, kk] = base
            dl = (((ip - im) * (ip + im - 2 * tgt)).sum() / ip.numel()).item()
            step = (f(np.float32(base + h)) - f(np.float32(base - h))) / 65536.0
            fd = dl / step
            err = abs(fd - ref[p, kk])
            assert err <= 1e-2 * abs(ref[p, kk]) + 1e-4 * gmax, (int(p), kk, fd, ref[p, kk])
            if abs(ref[p, kk]) >= 1e-2 * gmax:
                worst = max(worst, err / abs(ref[p, kk]))
    print(f"config-5 gradient checks: 1000 points, max rel err vs FP64 {rel64:.3e}, "
          f"central differences (h=1e-3) max rel err {worst:.3e}")
    r.close()
    cloud.close()
