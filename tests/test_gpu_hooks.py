"""User kernels over a cloud's pool (TintPass.swift:215-338 preamble helpers,
here include/scr_device.h): a HIP tint kernel built against the header
writes tints that must equal the same function evaluated on the CPU from the
oracle's decode, bit for bit, and a render with applyTint must equal the
oracle's render with those tints."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from _pkg import orc, scr
from fixtures import jittered_cloud
from helpers import assert_planes_equal, gpu_cloud, oracle_cloud, render_gpu_hqs

pytestmark = pytest.mark.gpu

HOOK = Path(__file__).resolve().parent / "hooks" / "libuser_hooks.so"


def _hook():
    if not HOOK.exists():
        raise RuntimeError(f"{HOOK} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(str(HOOK))
    lib.tint_hook_run.restype = C.c_int
    lib.tint_hook_run.argtypes = [C.c_void_p, C.c_float, C.c_float]
    lib.displace_hook_run.restype = C.c_int
    lib.displace_hook_run.argtypes = [C.c_void_p]
    return lib


def _cpu_tints(o: orc.Cloud, ppb: int, lo: float, hi: float) -> np.ndarray:
    """tint_by_height (tests/hooks/tint_hook.hip) on the CPU, zero where no thread writes."""
    f32 = np.float32
    k = f32(1.0) / f32(255.0)
    out = np.zeros((o.xyzLow.shape[0], 4), np.float32)
    for s, b in enumerate(o.batches):
        if b["state"] == 0:
            continue
        for local in range(int(b["numPoints"])):
            i = int(b["firstPoint"]) + local
            level = int(o.levels[i]) & 7
            p = orc.decode_point(b, (o.xyzLow[i], o.xyzMed[i], o.xyzHigh[i]), min(level, 2), fused=True)
            c = int(o.colors[i])
            cr, cg, cb = (f32(c & 0xFF) * k, f32((c >> 8) & 0xFF) * k, f32((c >> 16) & 0xFF) * k)
            if f32(lo) < p[1] < f32(hi):
                out[i] = (0, 0, 0, -1)
            elif p[1] >= f32(hi):
                out[i] = (f32(1) - cr, f32(1) - cg, f32(1) - cb, 0.75)
            else:
                out[i] = (p[0] * f32(0.1), 0.5, f32(local) * f32(1.0 / 1024.0), 0.5)
    return out


@pytest.mark.parametrize("clod", [False, True])
def test_user_tint_kernel_matches_cpu_and_renders_like_the_oracle(ctx, clod):
    pos, col = jittered_cloud(6000, 0x7137)
    packed = orc.pack(pos, col, pointsPerBatch=500, lodLevels=4)
    g, o = gpu_cloud(ctx, packed), oracle_cloud(packed)
    view = g.deviceView(withTints=True)
    assert view.tints and view.slots * view.pointsPerBatch == o.xyzLow.shape[0]
    lo, hi = float(np.percentile(pos[:, 1], 30)), float(np.percentile(pos[:, 1], 45))
    assert _hook().tint_hook_run(C.addressof(view), lo, hi) == 0
    got = np.zeros((o.xyzLow.shape[0], 4), np.float32)
    ctx.check(ctx.lib.scr_memcpy_d2h(ctx.handle, got.ctypes.data, view.tints, got.nbytes))
    want = _cpu_tints(o, view.pointsPerBatch, lo, hi)
    assert (want[:, 3] < 0).sum() > 100 and (want[:, 3] == 0.75).sum() > 100
    bad = np.argwhere(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (bad[:4], got[bad[:4, 0]], want[bad[:4, 0]])
    cam = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    cfg = scr.ComputeRasteriserConfiguration(applyTint=True, enableCLOD=clod, minimumPointSize=1,
                                             maximumPointSize=2, pointSizeScale=20.0)
    o.tints = want
    gpu = render_gpu_hqs(ctx, [g], cfg, cam, 320, 240)
    ref = orc.render_hqs([o], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
    assert ref["count"].sum() > 1000
    assert_planes_equal(gpu, ref)


def test_device_view_without_tints_and_after_set_tints(ctx):
    pos, col = jittered_cloud(1000, 0x55)
    packed = orc.pack(pos, col, pointsPerBatch=250)
    g = gpu_cloud(ctx, packed)
    v = g.deviceView()
    assert not v.tints and v.batches and v.xyzLow and v.colors and v.levels and v.stream
    n = v.slots * v.pointsPerBatch
    t = np.full((n, 4), 0.25, np.float32)
    g.setTints(t)
    v2 = g.deviceView(withTints=True)
    back = np.zeros_like(t)
    ctx.check(ctx.lib.scr_memcpy_d2h(ctx.handle, back.ctypes.data, v2.tints, back.nbytes))
    assert np.array_equal(back, t)


def _cpu_displacements(o: orc.Cloud) -> np.ndarray:
    """displace_wave (tests/hooks/displace_hook.hip) on the CPU, zero where no thread writes."""
    f32 = np.float32
    out = np.zeros((o.xyzLow.shape[0], 4), np.float32)
    for s, b in enumerate(o.batches):
        if b["state"] == 0:
            continue
        for local in range(int(b["numPoints"])):
            i = int(b["firstPoint"]) + local
            level = int(o.levels[i]) & 7
            p = orc.decode_point(b, (o.xyzLow[i], o.xyzMed[i], o.xyzHigh[i]), min(level, 2), fused=True)
            if local % 17 == 0:
                out[i] = (0.0, np.nan, 0.0, 0.0)
            else:
                out[i] = (0.0, f32(local) * f32(1.0 / 4096.0), p[0] * f32(0.25), 0.0)
    return out


def test_user_displacement_kernel_matches_cpu_and_renders_like_the_oracle(ctx):
    """DisplacementPass (DisplacementPass.swift:120-172): a HIP displacement
    kernel built against scr_device.h writes the cloud's displacement buffer;
    a render with applyDisplacement draws the displaced points, NaN ones
    removed, as the oracle does."""
    pos, col = jittered_cloud(6000, 0xD15)
    packed = orc.pack(pos, col, pointsPerBatch=500, lodLevels=4)
    g, o = gpu_cloud(ctx, packed), oracle_cloud(packed)
    view = g.deviceView(withDisplacements=True)
    assert view.displacements and not view.tints
    assert _hook().displace_hook_run(C.addressof(view)) == 0
    got = np.zeros((o.xyzLow.shape[0], 4), np.float32)
    ctx.check(ctx.lib.scr_memcpy_d2h(ctx.handle, got.ctypes.data, view.displacements, got.nbytes))
    want = _cpu_displacements(o)
    assert np.isnan(want[:, 1]).sum() > 100
    bad = np.argwhere(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (bad[:4], got[bad[:4, 0]], want[bad[:4, 0]])
    cam = scr.Camera.perspective((11.0, 7.0, 13.0), (0, 0, 0), fov=55, aspect=4 / 3)
    o.disp = want
    for kw in (dict(), dict(minimumPointSize=1, maximumPointSize=3, pointSizeScale=20.0)):
        cfg = scr.ComputeRasteriserConfiguration(applyDisplacement=True, **kw)
        gpu = render_gpu_hqs(ctx, [g], cfg, cam, 320, 240)
        ref = orc.render_hqs([o], cfg, cam.viewMatrix, cam.projectionMatrix, 320, 240)
        still = orc.render_hqs([o], scr.ComputeRasteriserConfiguration(**kw), cam.viewMatrix, cam.projectionMatrix,
                               320, 240)
        assert ref["count"].sum() > 1000 and (ref["count"] != still["count"]).sum() > 500
        assert_planes_equal(gpu, ref)
