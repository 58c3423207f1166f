#!/usr/bin/env python3
"""Headline benchmark: HQS frames of a 100M-point cloud at 1920x1080 on one MI355X.

Workload (one "step" = one full HQS frame: clear, cull, depth, colour,
resolve): the 100M-point LiDAR-like scene (8 km x 8 km value-noise terrain +
32,000 boxes, seed 0x5C0003), generated on the GPU and packed on the GPU with
63-bit Morton keys into 10,240-point batches, rendered at 1920x1080 with CLOD
off (every point drawn), 1-pixel points, from an overview pose that keeps the
whole scene in view.  Inputs stay resident in HBM; the timed region is K
frames bracketed by device synchronisation.

value = processed points (all points of all visible batches) per second.
roofline: the depth pass (algorithmic bytes = per processed point 1 B level
+ 4/8/12 B position by batch precision, + 80 B per visible batch) against the
MI355X HBM peak; traffic comes from the rocprofv3 FETCH_SIZE/WRITE_SIZE runs
summarised in profiles/ when present.
extra.atomic_roofline: the depth and colour passes' global atomics (counts per
launch from the rocprofv3 TCC_ATOMIC summary in profiles/) per second against
the device's measured atomic rate for the same access pattern.
cpu_baseline: the CPU oracle (scalar C transcription of the same passes,
multithreaded over batches, built -O3 -march=native for the host) on the
same frame, median of 5, with per-pass times and the host's CPU model.
"""
from __future__ import annotations

import argparse
import contextlib
import ctypes as C
import gc
import importlib.util
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
PKG_DIR = REPO / "satin-computerasteriser_amd"
HBM_PEAK_GBS = 8000.0
# the secondary measurements (bench.py --sections)
SECTIONS = ("clod", "nearest", "pack", "blend", "holefill", "ply", "flythrough", "diff", "multiview")


def load_pkg():
    name = "satin_computerasteriser_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


@contextlib.contextmanager
def quiet_host():
    """Timed regions run with Python's cyclic collector paused (collected just
    before): a collection inside a region stalls the host thread that feeds
    the GPU and shows up as several milliseconds of device idle (measured)."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def overview_camera(scr, width, height):
    return scr.overview_camera(width, height)


FLY_SCENE = dict(seed=0x5C0003, halfExtent=4000.0, numBoxes=32000)


def blend_bench(scr, ctx, ras, cloud, cam, processed, steps, width, height):
    """Weighted-blend modes on the headline frame: translucent defocus (tints
    all zero: every point fully in focus, so every fragment blends) and motion
    blur between two poses rendered alternately.  The blend colour pass makes
    4 sums per fragment (1-px points, no sweep: a fragment per processed
    point), merged in LDS tiles before the planes; the sum rate is set beside
    the device's global atomic rate (above 1: the tiles absorb the traffic)."""
    cloud.deviceView(withTints=True)
    out = {}
    ras.configuration = scr.ComputeRasteriserConfiguration(enableCLOD=False, applyTint=True,
                                                           tintAlphaIsCoverage=True)
    out["defocus_ms_per_frame"] = round(timed_frames(ctx, ras, cam, steps, 1), 4)
    color_ms = pass_profile(ctx, ras, cam, steps)["color"]
    peak = ctx.measure_atomic_rate(add=True) / 1e9
    achieved = 4.0 * processed / (color_ms * 1e-3) / 1e9
    out["defocus_color_ms"] = round(color_ms, 4)
    out["defocus_color_Gpts_per_s"] = round(processed / (color_ms * 1e-3) / 1e9, 2)
    out["defocus_sums"] = {"sums_per_s_G": round(achieved, 2), "global_atomic_rate_G": round(peak, 2),
                           "ratio": round(achieved / peak, 4), "sums": 4 * processed}
    cloud.setTints(None)
    # the overview pose with the eye 40 m to the side and 25 m up
    cam2 = scr.Camera(scr.look_at((40.0, -7200.0, 5225.0), (0.0, 300.0, 0.0), (0.0, 0.0, 1.0)),
                      scr.perspective_reverse_z(55.0, width / height, 10.0, 30000.0))
    ras.configuration = scr.ComputeRasteriserConfiguration(enableCLOD=False, motionBlur=0.5, motionBlurSamples=4,
                                                           motionBlurMaxSpread=8.0)
    for k in range(2):
        ras.render(cam if k % 2 == 0 else cam2)
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        ras.render(cam if k % 2 == 0 else cam2)
    ctx.sync()
    out["motion_blur_ms_per_frame"] = round((time.perf_counter() - t0) * 1e3 / steps, 4)
    return out


def hole_fill_bench(scr, ctx, ras, cam, width, height, passes=20):
    """Hole-fill passes over the 1080p output: 4 B read + 4 B written per
    pixel and pass (the neighbour reads hit the cache), against HBM peak."""
    ras.render(cam)
    ras.frame.holeFill(2)
    ctx.sync()
    ctx.profile(True)
    for _ in range(passes):
        ras.frame.holeFill(2)
    prof = ctx.collect_profile()
    ctx.profile(False)
    tot, n = prof["hole_fill"]
    ms = tot / n / 2.0
    gbs = 8.0 * width * height / (ms * 1e-3) / 1e9
    return {"ms_per_pass": round(ms, 5), "achieved_GBs": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}


def ply_bench(scr, ctx, n_binary=10_000_000, n_ascii=1_000_000):
    """PLY input: host parse rate of a binary little-endian file (x y z float,
    red green blue uchar) and of an ASCII one, and the binary file's parse +
    GPU pack."""
    rng = np.random.default_rng(0x91E)
    rec = np.zeros(n_binary, [("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("r", "u1"), ("g", "u1"), ("b", "u1")])
    for k in "xyz":
        rec[k] = rng.uniform(-100, 100, n_binary).astype(np.float32)
    for k in "rgb":
        rec[k] = rng.integers(0, 256, n_binary, dtype=np.uint8)
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
            "property float z\nproperty uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n")
    binary = (head % n_binary).encode() + rec.tobytes()
    t0 = time.perf_counter()
    pos, col = scr.PLYPointCloudLoader.parse(binary)
    t_bin = time.perf_counter() - t0
    sub = rec[:n_ascii]
    body = "\n".join(f"{a:.6g} {b:.6g} {c:.6g} {r} {g} {bl}" for a, b, c, r, g, bl in
                     zip(sub["x"], sub["y"], sub["z"], sub["r"], sub["g"], sub["b"]))
    ascii_ = ((head.replace("binary_little_endian", "ascii") % n_ascii) + body).encode()
    t0 = time.perf_counter()
    scr.PLYPointCloudLoader.parse(ascii_)
    t_asc = time.perf_counter() - t0
    cloud = scr.ComputeRasteriserPointCloud(ctx, 1, scr.DEFAULT_POINTS_PER_BATCH)
    ctx.sync()
    t0 = time.perf_counter()
    scr.PLYPointCloudLoader.packInto(scr.GPUPacker(ctx, keyBits=63), cloud, binary)
    ctx.sync()
    t_pack = time.perf_counter() - t0
    cloud.close()
    del pos, col
    return {"binary_points": n_binary, "binary_parse_ms": round(t_bin * 1e3, 1),
            "binary_MB_per_s": round(len(binary) / t_bin / 1e6, 1),
            "ascii_points": n_ascii, "ascii_parse_ms": round(t_asc * 1e3, 1),
            "ascii_MB_per_s": round(len(ascii_) / t_asc / 1e6, 1),
            "binary_parse_and_gpu_pack_ms": round(t_pack * 1e3, 1),
            "binary_threads": min(16, os.cpu_count() or 1, max(1, n_binary // 262144)),
            "ascii_threads": min(16, os.cpu_count() or 1, max(1, n_ascii // 32768))}


def flythrough_bench(scr, ctx, cloud, frames=120, width=2560, height=1440):
    """The 120-frame CLOD fly-through (Catmull-Rom spline through 8 seeded
    control points 50-400 m above the terrain) at 2560x1440, CLOD and dither on."""
    cams = scr.flythrough_cameras(frames, pathSeed=0x5C0003, width=width, height=height, **FLY_SCENE)
    ras = scr.ComputeRasteriser(ctx, scr.ComputeRasteriserConfiguration(enableCLOD=True, enableLODDither=True))
    ras.addPointCloud(cloud)
    ras.resize(width, height)
    processed, per_frame = [], []
    for cam in cams:   # untimed pass: visible work and per-frame times
        t0 = time.perf_counter()
        ras.render(cam)
        ctx.sync()
        per_frame.append((time.perf_counter() - t0) * 1e3)
        processed.append(int(cloud.readVisible()["activePoints"].astype(np.int64).sum()))
    with quiet_host():
        t0 = time.perf_counter()
        for cam in cams:
            ras.render(cam)
        ctx.sync()
        total_ms = (time.perf_counter() - t0) * 1e3
    ras.frame.close()
    return {"frames": frames, "width": width, "height": height, "ms_per_frame_mean": round(total_ms / frames, 4),
            "ms_per_frame_synced_min_max": [round(min(per_frame), 4), round(max(per_frame), 4)],
            "processed_points_per_frame_mean": int(np.mean(processed)),
            "processed_points_per_frame_min_max": [min(processed), max(processed)],
            "Gpts_per_s": round(sum(processed) / (total_ms * 1e-3) / 1e9, 3)}


def algorithmic_bytes(vis, per_point_extra=0):
    """Depth-pass bytes: per processed point 1 B level + 4 B per position word
    (3/2/1 words for precision level 0/1/>=2), + 80 B per visible batch."""
    lv = vis["level"].astype(np.int64)
    words = np.where(lv == 0, 3, np.where(lv == 1, 2, 1))
    pts = vis["activePoints"].astype(np.int64)
    return int((pts * (1 + 4 * words + per_point_extra)).sum() + 80 * len(vis)), int(pts.sum())


def timed_frames(ctx, ras, cam, steps, warmup):
    for _ in range(warmup):
        ras.render(cam)
    ctx.sync()
    with quiet_host():
        t0 = time.perf_counter()
        for _ in range(steps):
            ras.render(cam)
        ctx.sync()
        return (time.perf_counter() - t0) * 1e3 / steps


def pass_profile(ctx, ras, cam, steps):
    ctx.profile(True)
    for _ in range(steps):
        ras.render(cam)
    prof = ctx.collect_profile()
    ctx.profile(False)
    return {k: v[0] / v[1] for k, v in prof.items()}  # average ms per launch


STALE_PROFILES: dict = {}


def profile_text(name: str, library_version: str):
    """A committed rocprof summary, only when it was measured on this build:
    each summary names the library version string (with its source hash) it
    was taken on; a summary of other sources is reported as stale, not used."""
    f = REPO / "profiles" / name
    if not f.exists():
        STALE_PROFILES[name] = "missing"
        return None
    text = f.read_text()
    m = re.search(r"source ([0-9a-f]{12})", text)
    want = re.search(r"source ([0-9a-f]{12})", library_version)
    if not m or not want or m.group(1) != want.group(1):
        STALE_PROFILES[name] = (f"measured on source {m.group(1) if m else 'unknown'}, "
                                f"this build is {want.group(1) if want else 'unknown'}")
        return None
    return text


def read_traffic(kernel_tag: str, library_version: str):
    """HBM bytes per launch of the named kernel from the committed rocprof summary."""
    text = profile_text("traffic_summary.md", library_version)
    if text is None:
        return None
    for line in text.splitlines():
        m = re.match(rf"\|\s*{re.escape(kernel_tag)}(<[^|]*>)?\s*\|\s*([0-9.eE+]+)\s*\|", line)
        if m:
            return float(m.group(2))
    return None


def read_counter(kernel_tag: str, counter: str, library_version: str):
    """Per-dispatch value of a rocprof counter for the named kernel from the
    committed counter summary (the table whose header names the counter)."""
    text = profile_text("hqs_100m_counters.md", library_version)
    if text is None:
        return None
    col = None
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if counter in cells:
            col = cells.index(counter)
            continue
        if col is not None and cells and re.fullmatch(rf"{re.escape(kernel_tag)}(<[^>]*>)?", cells[0]):
            try:
                return float(cells[col])
            except (ValueError, IndexError):
                return None
    return None


def atomic_roofline(scr, ctx, kernel_tag: str, ms: float, add: bool):
    """Global atomics per second of a pass (the TCC_ATOMIC count per launch
    from the committed counter summary over the pass time measured here)
    against the device's measured rate for the flushes' pattern."""
    count = read_counter(kernel_tag, "TCC_ATOMIC_sum", scr.version())
    peak = ctx.measure_atomic_rate(add=add) / 1e9
    if count is None or not (ms > 0):
        return {"unit": "Gatomic/s", "peak": round(peak, 2), "achieved": None, "frac": None, "atomics": count}
    achieved = count / (ms * 1e-3) / 1e9
    return {"bound": "atomics", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "Gatomic/s",
            "frac": round(achieved / peak, 4) if peak > 0 else None, "atomics": count,
            "op": "atomicAdd" if add else "atomicMax"}


def node_like_points(n, seed=0x7F4A7C15):
    """COPC-node-like points as the reference's pack benchmark builds them
    (GPUPackBenchTests.swift:37-61): 24 blobs at random centres of the unit
    box, each point a centre plus a uniform jitter of +-0.06 per axis, clamped
    to the box and centred on the origin; random RGB, alpha 1.  numpy's
    generator stands in for the reference's LCG (same distribution)."""
    rng = np.random.default_rng(seed)
    centres = rng.random((24, 3), dtype=np.float32)
    jitter = (rng.random((n, 3), dtype=np.float32) - np.float32(0.5)) * np.float32(0.12)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.clip(centres[np.arange(n) % 24] + jitter, 0, 1) - np.float32(0.5)
    col = np.ones((n, 4), np.float32)
    col[:, :3] = rng.random((n, 3), dtype=np.float32)
    return pos, col


def median_ms(fn, reps):
    out = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        out.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(out))


def pack_bench(scr, ctx, node=140_000, nodes_per_call=10, nodes_per_frame=8, big=13_600_000):
    """The packer on the reference's published workloads (GPUPackBenchTests.
    swift:80-204, tasks/todo.md:189, 226-228), inputs resident in HBM, each
    timing a host wall clock around the call and a device synchronisation
    (the reference's encode + commit + wait): one 140k-point node wholesale
    (median of 30), ten nodes in one call (median of 15, per node), eight
    nodes per frame through the slot-targeted chunk pack (addRawChunksGPU,
    median of 20, pool cleared in between), and a 13.6M-point scene with
    30- and 63-bit keys (median of 5); Gkeys/s = points / time."""
    ppb = scr.DEFAULT_POINTS_PER_BATCH
    out = {}
    pos, col = node_like_points(node * nodes_per_call)
    dpos, dcol = scr.DeviceBuffer.from_array(ctx, pos), scr.DeviceBuffer.from_array(ctx, col)
    nb = -(-node // ppb)
    for bits in (30, 63):
        packer = scr.GPUPacker(ctx, keyBits=bits)
        one = scr.ComputeRasteriserPointCloud(ctx, nb + 4, ppb)
        many = scr.ComputeRasteriserPointCloud(ctx, -(-node * nodes_per_call // ppb) + 8, ppb)

        def pack(cloud, n):
            packer.packDevice(cloud, dpos, dcol, n)
            ctx.sync()

        pack(one, node)
        per_node = median_ms(lambda: pack(one, node), 30)
        pack(many, node * nodes_per_call)
        batched = median_ms(lambda: pack(many, node * nodes_per_call), 15)
        frame_cloud = scr.ComputeRasteriserPointCloud(ctx, nb * nodes_per_frame + 4, ppb)
        o = packer._opts(True)
        first = C.c_int32(-1)

        def frame():
            for _ in range(nodes_per_frame):
                ctx.check(ctx.lib.scr_pack_chunk(frame_cloud.handle, dpos.ptr, dcol.ptr, node, C.byref(o),
                                                 C.byref(first)))
            ctx.sync()
            frame_cloud.clearAllBatches()

        frame()
        per_frame = median_ms(frame, 20)
        out[f"{bits}bit"] = {
            "node_points": node, "per_node_ms": round(per_node, 3),
            f"x{nodes_per_call}_one_call_ms": round(batched, 3),
            "amortised_per_node_ms": round(batched / nodes_per_call, 3),
            "per_call_overhead_ms": round(per_node - batched / nodes_per_call, 3),
            f"chunk_pack_{nodes_per_frame}_nodes_per_frame_ms": round(per_frame, 3),
            "node_Gkeys_per_s": round(node / (per_node * 1e-3) / 1e9, 3),
            "x10_Gkeys_per_s": round(node * nodes_per_call / (batched * 1e-3) / 1e9, 3)}
        one.close(); many.close(); frame_cloud.close()
    dpos.free(); dcol.free()
    del pos, col
    gpos, gcol = scr.gen_scene(ctx, big, on_device=True, seed=0x5C0002, halfExtent=1000.0, numBoxes=2000)
    cloud = scr.ComputeRasteriserPointCloud(ctx, -(-big // ppb) + 4, ppb)
    for bits in (30, 63):
        packer = scr.GPUPacker(ctx, keyBits=bits)

        def pack_big():
            packer.packDevice(cloud, gpos, gcol, big)
            ctx.sync()

        pack_big()
        ms = median_ms(pack_big, 5)
        ctx.profile(True)
        pack_big()
        prof = {k: round(v[0], 3) for k, v in ctx.collect_profile().items()}
        ctx.profile(False)
        out[f"scene_{big // 100_000 / 10:g}M_{bits}bit"] = {"points": big, "ms": round(ms, 3),
                                                            "Gkeys_per_s": round(big / (ms * 1e-3) / 1e9, 3),
                                                            "pass_ms": prof}
    cloud.close()
    gpos.free(); gcol.free()
    return out


def cloud_bounds(scr, cloud):
    return scr.cloud_bounds(cloud)


def multiview_bench(scr, ctx, steps, warmup, n_views=64):
    """64 cameras over 4 resident clouds (1M/5M/20M/50M points of the 2 km
    scene, seeds 0x5C0003..6); view v renders cloud v mod 4 from an eye
    uniform on the sphere of 1.5x the cloud's diagonal, at a resolution drawn
    from {256x256, 512x384, 1024x768, 1920x1080}; screen-space points 1-3 px.
    Batched (one launch per phase for all views, graph replay, autotuned
    shapes) against one scr_render per view on the same GPU."""
    import ctypes as C
    clouds, cams, sizes, which, cfg = scr.multiview_workload(ctx, n_views)
    pixels = sum(w * h for w, h in sizes)

    # one launch sequence per view: scr_render with prepared arguments
    frames = [scr.Frame(ctx, w, h) for w, h in sizes]
    ccfg = cfg.to_c()
    ccams = [cam.to_c() for cam in cams]
    handles = [(C.c_void_p * 1)(clouds[k].handle.value) for k in which]
    lib = ctx.lib

    def per_view_pass():
        for v in range(n_views):
            ctx.check(lib.scr_render(ctx.handle, C.byref(ccfg), C.byref(ccams[v]), handles[v], 1, frames[v].handle))

    processed = 0
    for v in range(n_views):   # visible work of every view (untimed)
        ctx.check(lib.scr_render(ctx.handle, C.byref(ccfg), C.byref(ccams[v]), handles[v], 1, frames[v].handle))
        processed += int(clouds[which[v]].readVisible()["activePoints"].astype(np.int64).sum())
    for _ in range(warmup):
        per_view_pass()
    ctx.sync()
    with quiet_host():
        t0 = time.perf_counter()
        for _ in range(steps):
            per_view_pass()
        ctx.sync()
        ms_per_view_path = (time.perf_counter() - t0) * 1e3 / steps
    want = [f.read(scr.PLANE_RGBA8) for f in frames]
    want_count = [f.read(scr.PLANE_COUNT) for f in frames]
    for f in frames:
        f.close()

    mv = scr.MultiViewRasteriser(ctx, clouds, sizes, which, cfg)
    t0 = time.perf_counter()
    tuning = mv.autotune(cams, reps=3)
    tune_s = time.perf_counter() - t0
    # untimed warm-up: the first render captures the graph; then a run of
    # back-to-back renders, since the first such run of a process (renders
    # that wait for the staging slot of the one two before) is ~7 ms slower
    # once (measured)
    mv.render(cams)
    ctx.sync()
    for _ in range(max(warmup, 6)):
        mv.render(cams)
    ctx.sync()
    with quiet_host():
        t0 = time.perf_counter()
        for _ in range(steps):
            mv.render(cams)
        ctx.sync()
        ms_batched = (time.perf_counter() - t0) * 1e3 / steps
    exact = all(np.array_equal(mv.frames[v].read(scr.PLANE_RGBA8), want[v]) and
                np.array_equal(mv.frames[v].read(scr.PLANE_COUNT), want_count[v]) for v in range(n_views))
    ctx.profile(True)
    mv.render(cams, useGraph=False)
    prof = {k: round(v[0] / v[1], 4) for k, v in ctx.collect_profile().items()}
    ctx.profile(False)
    mv.close()
    for c in clouds:
        c.close()
    return {
        "views": n_views, "pixels_per_launch": pixels, "processed_points_per_launch": processed,
        "views_per_s_batched": round(n_views / (ms_batched * 1e-3), 1),
        "views_per_s_per_view_launches": round(n_views / (ms_per_view_path * 1e-3), 1),
        "speedup": round(ms_per_view_path / ms_batched, 3),
        "ms_batched": round(ms_batched, 4), "ms_per_view_launches": round(ms_per_view_path, 4),
        "Gpts_per_s_batched": round(processed / (ms_batched * 1e-3) / 1e9, 3),
        "bit_exact_vs_per_view": bool(exact), "tuning": tuning, "autotune_s": round(tune_s, 2),
        "pass_ms_eager": prof,
    }


def diff_bench(scr, ctx, steps, fit_steps=200, n_points=5_000_000, n_views=32, size=512):
    """Differentiable HQS on 5M points of the 2 km scene (seed 0x5C0005,
    float ground-truth colours), 32 views of 512x512 on a sphere: forward +
    backward kernel throughput, then an FP32 Adam fit (lr 1e-2) of per-point
    colours from 0.5 to noisy (sigma 0.02) renders of the ground truth through
    the torch autograd.Function."""
    pos, col = scr.gen_scene(ctx, n_points, on_device=False, seed=0x5C0005, halfExtent=1000.0, numBoxes=2000,
                             quantizeColors=0)
    cloud = scr.ComputeRasteriserPointCloud(ctx, 1, scr.DEFAULT_POINTS_PER_BATCH)
    src = scr.GPUPacker(ctx, keyBits=63).pack(cloud, pos, col, sourceIndices=True)
    batches = cloud.readPool(scr.POOL_BATCHES)
    npool = cloud.batchCount * cloud.pointsPerBatch
    gt = np.zeros((npool, 4), np.float32)
    k = 0
    for s_ in range(cloud.batchCount):
        m, first = int(batches["numPoints"][s_]), int(batches["firstPoint"][s_])
        gt[first:first + m, :3] = col[src[k:k + m], :3]
        k += m
    ctr, diag = cloud_bounds(scr, cloud)
    cams, _ = scr.sphere_views(ctr, 1.2 * diag, n_views, 0x5C0005, resolutions=((size, size),))
    cfg = scr.ComputeRasteriserConfiguration(minimumPointSize=1.0, maximumPointSize=3.0, pointSizeScale=1500.0)
    r = scr.DifferentiableHQS(ctx, cloud, [(size, size)] * n_views, cfg, batched=True)
    r1 = scr.DifferentiableHQS(ctx, cloud, [(size, size)] * n_views, cfg)
    del pos, col, src

    # kernel throughput: forward + backward of every view, device-resident buffers,
    # batched (one launch per phase for all views) and one view at a time
    params = scr.DeviceBuffer.from_array(ctx, np.full((npool, 4), 0.5, np.float32))
    imgs = [scr.DeviceBuffer(ctx, size * size * 12) for _ in range(n_views)]
    dl = scr.DeviceBuffer.from_array(ctx, np.random.default_rng(5).normal(size=size * size * 3).astype(np.float32))
    grad = scr.DeviceBuffer(ctx, npool * 32)

    def fwd_bwd(rr):
        rr.forward_all(cams, params.ptr, [b.ptr for b in imgs])
        rr.backward_all(cams, [dl.ptr] * n_views, grad.ptr)

    processed = 0
    for v, cam in enumerate(cams):
        r1.forward(v, cam, params.ptr, imgs[v].ptr)
        processed += int(cloud.readVisible()["activePoints"].astype(np.int64).sum())
    out = {"points": n_points, "views": n_views, "size": size, "processed_points_per_pass": processed}
    for name, rr in (("batched", r), ("per_view", r1)):
        fwd_bwd(rr)
        ctx.sync()
        with quiet_host():
            t0 = time.perf_counter()
            for _ in range(steps):
                fwd_bwd(rr)
            ctx.sync()
            ms = (time.perf_counter() - t0) * 1e3 / steps
        ctx.profile(True)
        fwd_bwd(rr)
        prof = {k_: round(v_[0] / v_[1], 4) for k_, v_ in ctx.collect_profile().items()}
        ctx.profile(False)
        out[name] = {"ms_fwd_bwd_all_views": round(ms, 4), "Gpts_per_s_fwd_bwd": round(processed / (ms * 1e-3) / 1e9, 3),
                     "pass_ms": prof}
    out["Gpts_per_s_fwd_bwd"] = out["batched"]["Gpts_per_s_fwd_bwd"]
    params.free(); dl.free(); grad.free()
    for b in imgs:
        b.free()
    r1.close()

    try:
        import torch
        if not torch.cuda.is_available():
            raise ImportError("torch sees no GPU")
        import importlib.util
        spec = importlib.util.spec_from_file_location("scr_torch_diff", PKG_DIR / "torch_diff.py")
        td = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(td)
    except ImportError as e:
        out["fit"] = f"skipped: {e}"
        r.close(); cloud.close()
        return out
    dev = torch.device("cuda:0")
    render = td.HQSRender(r, cams)
    with torch.no_grad():
        gt_t = torch.from_numpy(gt).to(dev)
        clean = render(gt_t)
        gen = torch.Generator(device=dev).manual_seed(0x5C0005)
        target = clean + 0.02 * torch.randn(clean.shape, generator=gen, device=dev)
    theta = torch.full((npool, 4), 0.5, dtype=torch.float32, device=dev, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    losses = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(fit_steps):
        opt.zero_grad()
        loss = ((render(theta) - target) ** 2).mean()
        loss.backward()
        opt.step()
        losses.append(loss.item())
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    with torch.no_grad():
        seen = (clean.sum(dim=-1) > 0)
        rmse_img = float(((render(theta) - clean) ** 2).mean().sqrt())
    out["fit"] = {"steps": fit_steps, "lr": 1e-2, "loss_first": losses[0], "loss_last": losses[-1],
                  "loss_at": {str(i): losses[i] for i in (0, 10, 50, 100, 150) if i < len(losses)},
                  "noise_floor_sigma2": 0.02 ** 2, "image_rmse_vs_clean": rmse_img,
                  "ms_per_step": round(fit_s * 1e3 / fit_steps, 3), "covered_pixel_frac": float(seen.float().mean())}
    out["gradient_checks"] = gradient_checks(torch, r, render, target, npool, n_views)
    r.close()
    cloud.close()
    return out


def gradient_checks(torch, r, render, target, npool, n_views, n_fd=200, h=1e-3):
    """Accuracy of the gradient the fit uses (params 0.5, the fit's loss):
    the FP32 gradient torch receives against the library's FP64 accumulation
    of the same dL/dimage, over every point the views see; and against
    central differences (h = 1e-3, FP64 loss differences of re-rendered
    images, divided by the step the 2^-16 fixed-point colour actually takes)
    on n_fd seeded points, counting components at least 1% of the largest."""
    dev = torch.device("cuda:0")
    theta = torch.full((npool, 4), 0.5, dtype=torch.float32, device=dev, requires_grad=True)
    seen = []
    images = render(theta)
    images.register_hook(lambda g: seen.append(g.detach().clone().contiguous()))
    ((images - target) ** 2).mean().backward()
    g32 = theta.grad.detach().double()
    g64 = torch.zeros((npool, 4), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    r.backward_all(render.cameras, [seen[0][v].data_ptr() for v in range(n_views)], g64.data_ptr())
    r.ctx.sync()
    nz = g64[:, :3] != 0
    rel32 = float(((g32[:, :3] - g64[:, :3]).abs()[nz] / g64[:, :3].abs()[nz]).max())
    g = g64.cpu().numpy()
    touched = np.nonzero(np.abs(g[:, :3]).max(1) > 0)[0]
    picks = np.sort(np.random.default_rng(0x5C0005).choice(touched, min(n_fd, len(touched)), replace=False))
    gmax = float(np.abs(g[:, :3]).max())
    tgt = target.double()
    tp = theta.detach().clone()

    def fixed(v):  # the library's colour quantisation: rint(float32(v) * 2^16)
        return float(np.rint(np.float32(v) * np.float32(65536.0)))

    worst, counted = 0.0, 0
    with torch.no_grad():
        for p in picks:
            k = int(np.argmax(np.abs(g[p, :3])))
            if abs(g[p, k]) < 1e-2 * gmax:
                continue
            base = float(tp[p, k])
            tp[p, k] = base + h
            ip = render(tp).double()
            tp[p, k] = base - h
            im = render(tp).double()
            tp[p, k] = base
            dl = (((ip - im) * (ip + im - 2 * tgt)).sum() / ip.numel()).item()
            fd = dl / ((fixed(base + h) - fixed(base - h)) / 65536.0)
            worst = max(worst, abs(fd - g[p, k]) / abs(g[p, k]))
            counted += 1
    return {"fp32_vs_fp64_max_rel": rel32, "fp32_vs_fp64_points": int(len(touched)),
            "central_diff_max_rel": worst, "central_diff_points": counted, "h": h}


def host_cpu() -> dict:
    """The host's CPU model, logical CPU count and the CPUs this process may use."""
    model = "unknown"
    with contextlib.suppress(OSError):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "usable_cpus": usable}


def cpu_baseline(scr, cloud, cfg, cam, width, height, vis, runs=5):
    """The oracle (scalar C restatement of the same passes, threads over
    batches, private planes merged) on the same frame: built at -O3
    -march=native for this host (oracle/Makefile `native`), timed as the
    median of `runs` whole frames; per-pass medians from the oracle's own
    phase clock; plus a one-thread run on every 16th visible batch for the
    per-core rate."""
    import subprocess
    import tempfile
    sys.path.insert(0, str(REPO))
    from oracle import oracle as orc

    host = host_cpu()
    # the GPU host's CPU share is 16 threads per GPU, whatever it has in total
    threads = max(1, min(16, host["usable_cpus"]))
    build = "-O3 -march=native -ffp-contract=off"
    tmp = Path(tempfile.mkdtemp(prefix="scr_oracle_native_"))
    so = tmp / "libscr_oracle_native.so"
    rc = subprocess.run(["make", "-s", "-C", str(REPO / "oracle"), "native", f"OUT={so}"], capture_output=True)
    prev = orc.use_library(so) if rc.returncode == 0 and so.exists() else None
    if rc.returncode != 0:
        build = "-O2 (oracle/Makefile default; the -O3 -march=native build failed)"
    try:
        low = cloud.readPool(scr.POOL_XYZ_LOW)
        med = cloud.readPool(scr.POOL_XYZ_MED)
        high = cloud.readPool(scr.POOL_XYZ_HIGH)
        colors = cloud.readPool(scr.POOL_COLORS)
        levels = cloud.readPool(scr.POOL_LEVELS)
        batches = cloud.readPool(scr.POOL_BATCHES)
        vis_idx = np.sort(vis["batchIndex"].astype(np.int64))

        def frame(sel, nthreads):
            c = orc.Cloud(batches[sel], scr.identity_files(1), low, med, high, colors, levels)
            t0 = time.perf_counter()
            orc.render_hqs([c], cfg, cam.viewMatrix, cam.projectionMatrix, width, height, threads=nthreads)
            return time.perf_counter() - t0, orc.last_hqs_pass_seconds()

        npts = int(batches["numPoints"][vis_idx].astype(np.int64).sum())
        times, passes = [], []
        for _ in range(runs):
            dt, ps = frame(vis_idx, threads)
            times.append(dt)
            passes.append(ps)
        med_s = float(np.median(times))
        pass_med = {k: round(float(np.median([p[k] for p in passes])) * 1e3, 2) for k in passes[0]}
        sub = vis_idx[::16]
        n1 = int(batches["numPoints"][sub].astype(np.int64).sum())
        dt1, _ = frame(sub, 1)
        del low, med, high, colors, levels, batches
        # the CPU pack (PackedPointCloudFixtures.pack, one thread) on the
        # packer benchmark's inputs: one 140k-point node (median of 5) and the
        # 13.6M-point scene (one run), 30-bit keys as the reference sorts
        npos, ncol = node_like_points(140_000)
        node_ms = median_ms(lambda: orc.pack(npos[:, :3], ncol, pointsPerBatch=scr.DEFAULT_POINTS_PER_BATCH), 5)
        spos, scol = scr.gen_scene(cloud.ctx, 13_600_000, on_device=False, seed=0x5C0002, halfExtent=1000.0,
                                   numBoxes=2000)
        t0 = time.perf_counter()
        orc.pack(spos[:, :3], scol, pointsPerBatch=scr.DEFAULT_POINTS_PER_BATCH)
        scene_ms = (time.perf_counter() - t0) * 1e3
        del spos, scol
    finally:
        orc.use_library(prev)
        with contextlib.suppress(OSError):
            so.unlink()
            tmp.rmdir()
    return {"value": npts / med_s / 1e9, "unit": "Gpts/s", "cores": threads, "kind": "port",
            "sample": f"the whole frame: {npts} points of {len(vis_idx)} visible batches, HQS {width}x{height}, "
                      f"{threads} threads, median of {runs} frames ({min(times):.2f}-{max(times):.2f} s each)",
            "build": build, "pass_ms_median": pass_med, "frame_ms_runs": [round(t * 1e3, 1) for t in times],
            "one_thread_Gpts_s": round(n1 / dt1 / 1e9, 5), "one_thread_sample_points": n1,
            "pack_30bit_one_thread": {"node_140k_ms_median_of_5": round(node_ms, 2),
                                      "scene_13.6M_ms": round(scene_ms, 1),
                                      "scene_13.6M_Gkeys_per_s": round(13.6e6 / (scene_ms * 1e-3) / 1e9, 4)},
            **host}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--points", type=int, default=100_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary measurements")
    ap.add_argument("--no-multiview", action="store_true", help="skip the batched multi-view measurement")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable-mode measurement and fit")
    ap.add_argument("--sections", default="all",
                    help="comma-separated secondary measurements to run (" + ",".join(SECTIONS) + "), or all")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (the known-size copy kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    sections = set(SECTIONS) if args.sections == "all" else {x for x in args.sections.split(",") if x}
    unknown = sections - set(SECTIONS)
    if unknown:
        ap.error(f"unknown sections {sorted(unknown)}; choose from {SECTIONS}")
    if args.no_extra:
        sections = set()
    if args.no_multiview:
        sections.discard("multiview")
    if args.no_diff:
        sections.discard("diff")
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the library has no "
              "multi-GPU path", file=sys.stderr)
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        return 2

    if "diff" in sections:
        # torch (for the fit) must own the process's HIP runtime: import it
        # before the library creates its context
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    scr = load_pkg()
    ctx = scr.Context(0)
    W, H, N = args.width, args.height, args.points

    log(f"generating {N} points on the GPU")
    pos, col = scr.gen_scene(ctx, N, on_device=True, seed=0x5C0003, halfExtent=4000.0, numBoxes=32000)
    cloud = scr.ComputeRasteriserPointCloud(ctx, 1, scr.DEFAULT_POINTS_PER_BATCH)
    pack_100m = {}
    # the 63-bit pack last: it is the headline workload's cloud
    for bits in ((30, 63) if "pack" in sections else (63,)):
        packer = scr.GPUPacker(ctx, keyBits=bits)
        ctx.sync()
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            packer.packDevice(cloud, pos, col, N)
            ctx.sync()
            runs.append((time.perf_counter() - t0) * 1e3)
        ctx.profile(True)
        packer.packDevice(cloud, pos, col, N)
        prof = {k: round(v[0], 3) for k, v in ctx.collect_profile().items()}
        ctx.profile(False)
        pack_100m[f"{bits}bit"] = {"ms_median_of_3": round(float(np.median(runs)), 2),
                                   "ms_first_call": round(runs[0], 2),
                                   "Gkeys_per_s": round(N / (float(np.median(runs)) * 1e-3) / 1e9, 3),
                                   "pass_ms": prof}
    pack_ms = pack_100m["63bit"]["ms_median_of_3"]
    log(f"packed in {pack_ms:.1f} ms ({cloud.residentBatchCount} batches): {pack_100m}")
    pos.free()
    col.free()

    if args.copy_calibration_mb > 0:
        nbytes = args.copy_calibration_mb << 20
        a, b = scr.DeviceBuffer(ctx, nbytes), scr.DeviceBuffer(ctx, nbytes)
        ctx.check(ctx.lib.scr_memcpy_d2d(ctx.handle, b.ptr, a.ptr, nbytes))
        ctx.sync()
        a.free()
        b.free()
        log(f"copy calibration: {args.copy_calibration_mb} MiB device-to-device")

    cfg = scr.ComputeRasteriserConfiguration(enableCLOD=False)
    cam = overview_camera(scr, W, H)
    ras = scr.ComputeRasteriser(ctx, cfg)
    ras.addPointCloud(cloud)
    ras.resize(W, H)

    log("timing HQS frames")
    ms = timed_frames(ctx, ras, cam, args.steps, args.warmup)
    vis = cloud.readVisible()
    depth_bytes, processed = algorithmic_bytes(vis)
    prof = pass_profile(ctx, ras, cam, max(3, args.steps // 2))
    depth_ms = prof.get("depth", float("nan"))
    achieved = depth_bytes / (depth_ms * 1e-3) / 1e9
    traffic = read_traffic("k_depth", scr.version())
    value = processed / (ms * 1e-3) / 1e9
    log(f"HQS {ms:.3f} ms/frame, {value:.2f} Gpts/s; passes {prof}")

    extra = {"pack_ms_63bit": round(pack_ms, 2), "pack_100M": pack_100m,
             "pass_ms": {k: round(v, 4) for k, v in prof.items()},
             "processed_points": processed, "visible_batches": int(len(vis)),
             "depth_algorithmic_bytes": depth_bytes}
    cb_bytes, _ = algorithmic_bytes(vis, per_point_extra=4)
    extra["color_roofline_frac"] = round(cb_bytes / (prof.get("color", float("nan")) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if sections:
        extra["atomic_roofline"] = {"k_depth": atomic_roofline(scr, ctx, "k_depth", depth_ms, add=False),
                                    "k_color": atomic_roofline(scr, ctx, "k_color", prof.get("color", float("nan")),
                                                               add=True)}
    if "clod" in sections:
        # CLOD on at the same pose
        ras.configuration = scr.ComputeRasteriserConfiguration(enableCLOD=True)
        ms_c = timed_frames(ctx, ras, cam, args.steps, 1)
        vis_c = cloud.readVisible()
        _, proc_c = algorithmic_bytes(vis_c)
        extra["hqs_clod_on"] = {"ms_per_frame": round(ms_c, 4), "processed_points": proc_c,
                                "Gpts_per_s_of_cloud": round(N / (ms_c * 1e-3) / 1e9, 3)}
    if "nearest" in sections:
        # nearest-point mode, CLOD off
        ras.configuration = scr.ComputeRasteriserConfiguration(mode=scr.MODE_NEAREST_POINT, enableCLOD=False)
        ms_n = timed_frames(ctx, ras, cam, args.steps, 1)
        extra["nearest"] = {"ms_per_frame": round(ms_n, 4), "Gpts_per_s": round(processed / (ms_n * 1e-3) / 1e9, 3)}
    if "pack" in sections:
        log("packer: 140k-point nodes, 10 per call, 8 per frame, 13.6M scene (30/63-bit)")
        extra["pack"] = pack_bench(scr, ctx)
        log(f"pack {extra['pack']}")
    if "blend" in sections:
        log("weighted-blend modes (defocus, motion blur)")
        extra["blend_modes"] = blend_bench(scr, ctx, ras, cloud, cam, processed, max(3, args.steps // 10), W, H)
        log(f"blend {extra['blend_modes']}")
    if "holefill" in sections:
        ras.configuration = cfg
        extra["hole_fill_1080p"] = hole_fill_bench(scr, ctx, ras, cam, W, H)
        log(f"hole fill {extra['hole_fill_1080p']}")
    if "ply" in sections:
        extra["ply"] = ply_bench(scr, ctx)
        log(f"ply {extra['ply']}")
    ras.configuration = cfg
    if "flythrough" in sections:
        log("CLOD fly-through, 120 frames at 2560x1440")
        extra["flythrough_120_1440p"] = flythrough_bench(scr, ctx, cloud)
    if "diff" in sections:
        log("differentiable HQS (5M points, 32 views 512x512) and Adam fit")
        extra["diff_hqs"] = diff_bench(scr, ctx, max(2, args.steps // 5))
        log(f"diff {extra['diff_hqs']}")
    if "multiview" in sections:
        log("batched multi-view (64 views, 4 clouds)")
        extra["multiview_64"] = multiview_bench(scr, ctx, max(3, args.steps // 4), 1)
        log(f"multi-view {extra['multiview_64']}")

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle) on a bounded sample")
        ras.render(cam)
        vis = cloud.readVisible()
        cpu = cpu_baseline(scr, cloud, cfg, cam, W, H, vis)

    line = {
        "metric": "Gpts/s HQS 1080p 100M pts",
        "value": round(value, 4),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u32",
        "data": "synthetic",
        "config": {"workload": f"HQS frame, {N} pt LiDAR-like cloud (8 km square, 32k boxes, seed 0x5C0003), "
                               f"GPU-packed 63-bit Morton, {W}x{H}, CLOD off, 1 px points, overview pose",
                   "points": N, "width": W, "height": H, "points_per_batch": 10240, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_depth", "depth_ms": round(depth_ms, 4)},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    line["extra"]["library"] = scr.version()
    if STALE_PROFILES:
        line["extra"]["profiles_not_used"] = STALE_PROFILES
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
