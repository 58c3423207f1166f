"""Pins for the oracle's arithmetic against closed forms, published
algorithms and brute force (so that a dropped term, wrong shift or
transposed operand in the oracle fails here)."""
import math
from fractions import Fraction

import numpy as np
import pytest

from _pkg import orc, scr
from fixtures import jittered_cloud, lattice


def test_log2_matches_numpy_within_one_ulp_and_is_exact_on_powers_of_two():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(0.04, 1e6, 4000), rng.uniform(0.9, 1.1, 2000),
                         [1.0, 2.0, 0.5, 1024.0, 0.04, 4.0, 25.0]]).astype(np.float32)
    for x in xs:
        got = np.float32(orc.log2f(float(x)))
        ref = np.float32(np.log2(np.float64(x)))
        assert abs(int(got.view(np.int32)) - int(ref.view(np.int32))) <= 1, (x, got, ref)
    for e in range(-20, 30):
        assert orc.log2f(2.0 ** e) == float(e)
    assert math.isnan(orc.log2f(float("nan")))
    assert orc.log2f(float("inf")) == float("inf")


def _lowbias32_inverse(x):
    # published inverse of lowbias32 (Wellons, hash-prospector)
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x43021123) & 0xFFFFFFFF
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1D69E2A5) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def test_hash_unit_is_lowbias32_scaled():
    rng = np.random.default_rng(2)
    for v in list(rng.integers(0, 2 ** 32, 2000)) + [0, 1, 0xFFFFFFFF, 12345]:
        u = orc.hash_unit(int(v))
        assert 0.0 <= u < 1.0 or u == 1.0
        # u = float32(h) / 2^32; recover h for values that are exact in binary32
        h = int(np.float64(u) * 2.0 ** 32)
        if h < 2 ** 24:
            assert _lowbias32_inverse(h) == int(v)
    # statistical sanity
    us = np.array([orc.hash_unit(i) for i in range(20000)])
    assert abs(us.mean() - 0.5) < 0.01


def test_hash_unit_small_outputs_invert_exactly():
    # find inputs whose hash is small (exactly representable) by inverting small hashes
    for h in (0, 1, 7, 123456, 2 ** 23 + 5):
        x = _lowbias32_inverse(h)
        assert orc.hash_unit(x) == np.float32(h) / np.float32(2.0 ** 32)


def test_depth_to_uint_closed_forms():
    assert orc.depth_to_uint(0.5) == 2 ** 31
    assert orc.depth_to_uint(0.25) == 2 ** 30
    assert orc.depth_to_uint(1.0) == 0xFFFFFFFF  # 2^32 saturates
    assert orc.depth_to_uint(0.0) == 0
    assert orc.depth_to_uint(-1.0) == 0
    assert orc.depth_to_uint(float("nan")) == 0
    z = np.float32(0.7)
    assert orc.depth_to_uint(float(z)) == int(np.float64(z) * 2 ** 32)


def _cvt_u32(y: np.float32) -> int:
    # float -> u32 under contract C4: truncate, saturate, NaN -> 0
    if y != y or y <= 0:
        return 0
    if y >= np.float32(2.0 ** 32):
        return 0xFFFFFFFF
    return int(y)


def test_depth_to_uint_needs_no_clamp_before_a_saturating_conversion():
    # the device form converts z * 2^32 without clamping z to [0, 1] first
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2 ** 32, size=20000, dtype=np.uint64).astype(np.uint32)
    zs = list(bits.view(np.float32))
    zs += [np.float32(v) for v in (0.0, -0.0, 1.0, -1.0, 1.0000001, 0.99999994, 2.0 ** -149, -2.0 ** -149,
                                   2.0 ** -126, 2.0 ** 96, 2.0 ** 127, float("inf"), float("-inf"), float("nan"))]
    zs += list(rng.random(5000, dtype=np.float32))
    two32 = np.float32(2.0 ** 32)
    with np.errstate(over="ignore", invalid="ignore"):
        for z in zs:
            assert orc.depth_to_uint(float(z)) == _cvt_u32(np.float32(z) * two32), z


def _interleave(x, y, z, bits):
    k = 0
    for i in range(bits):
        k |= ((x >> i) & 1) << (3 * i + 2)
        k |= ((y >> i) & 1) << (3 * i + 1)
        k |= ((z >> i) & 1) << (3 * i)
    return k


def test_morton_matches_bit_interleave():
    rng = np.random.default_rng(3)
    for _ in range(500):
        x, y, z = (int(v) for v in rng.integers(0, 1024, 3))
        assert orc.morton30(x, y, z) == _interleave(x, y, z, 10)
        X, Y, Z = (int(v) for v in rng.integers(0, 2 ** 21, 3))
        assert orc.morton63(X, Y, Z) == _interleave(X, Y, Z, 21)


def test_mat_mul_exact_on_small_integers_and_order():
    rng = np.random.default_rng(4)
    a = rng.integers(-8, 8, (4, 4)).astype(np.float32)
    b = rng.integers(-8, 8, (4, 4)).astype(np.float32)
    # storage is m[column, row]: the math product A·B is (b.T @ a.T).T in storage
    got = orc.mat_mul(a, b)
    assert np.array_equal(got.T, a.T @ b.T)


def test_pack_order_is_stable_sort_by_key():
    pos, col = jittered_cloud(3000, 0x5EED)
    pos[100:200] = pos[5]  # equal keys: ties must keep input order
    for bits in (30, 63):
        packed = orc.pack(pos, col, pointsPerBatch=60000, lodLevels=1, keyBits=bits)
        keys = orc.morton_keys(pos, bits)
        assert np.array_equal(packed.sourceIndices, np.argsort(keys, kind="stable").astype(np.uint32))


def test_lod_levels_first_claimer_brute_force():
    pos, _ = jittered_cloud(1500, 0xC0FFEE)
    bmin, bmax = pos.min(0), pos.max(0)
    lev = orc.lod_levels(pos, bmin, bmax, lodLevels=4, coarse=4)
    ext = np.maximum(bmax - bmin, np.float32(1e-6))
    longest = ext.max()
    base = np.float32(longest / np.float32(4))
    expect = np.full(len(pos), 3, np.uint8)
    for level in range(3):
        voxel = np.float32(base * np.float32(0.5 ** level))
        inv = np.float32(np.float32(1) / np.maximum(voxel, np.float32(1e-6)))
        seen = set()
        for i in range(len(pos)):
            if expect[i] != 3:
                continue
            cell = tuple(int(np.floor(np.float32((pos[i, a] - bmin[a]) * inv))) for a in range(3))
            if cell not in seen:
                seen.add(cell)
                expect[i] = level
    assert np.array_equal(lev, expect)


def test_decode_identity_camera_single_point_pixel():
    # one point at (0, 0, 0.5) with identity camera lands in the centre pixel
    # with depth 2^31 (reverse-Z uint encoding, Common.metal:10-12)
    packed = orc.pack([(0.0, 0.0, 0.5)], [(1, 1, 1, 1)], pointsPerBatch=1)
    cloud = orc.Cloud.from_packed(packed)
    cfg = scr.ComputeRasteriserConfiguration(enableFrustumCulling=False, enableCLOD=False)
    eye = np.eye(4, dtype=np.float32)
    out = orc.render_hqs([cloud], cfg, eye, eye, 8, 8)
    assert out["count"][8 - 1 - 4, 4] == 1
    assert out["depth"][3, 4] == 2 ** 31
    assert out["count"].sum() == 1
    assert out["rgba8"][3, 4] == 0xFFFFFFFF


def test_hqs_invariant_under_batch_slot_permutation_and_threads():
    pos, col = jittered_cloud(4000, 0x1234)
    packed = orc.pack(pos, col, pointsPerBatch=256)
    cloud = orc.Cloud.from_packed(packed)
    cam = scr.Camera.perspective((12.0, 9.0, 14.0), (0, 0, 0), fov=50, aspect=1.5)
    cfg = scr.ComputeRasteriserConfiguration(minimumPointSize=1, maximumPointSize=3, pointSizeScale=25.0)
    ref = orc.render_hqs([cloud], cfg, cam.viewMatrix, cam.projectionMatrix, 96, 64)
    # permute slots (the batch table order) — results must not change
    perm = np.random.default_rng(5).permutation(len(cloud.batches))
    c2 = orc.Cloud(cloud.batches[perm], cloud.files, cloud.xyzLow, cloud.xyzMed, cloud.xyzHigh, cloud.colors,
                   cloud.levels)
    out = orc.render_hqs([c2], cfg, cam.viewMatrix, cam.projectionMatrix, 96, 64, threads=4)
    for k in ref:
        assert np.array_equal(ref[k], out[k]), k
    assert ref["count"].sum() > 100


def test_nearest_two_pass_matches_packed_max_rule():
    # nearest index = min point index among max-depth fragments
    pos, col = jittered_cloud(3000, 0x77)
    packed = orc.pack(pos, col, pointsPerBatch=300)
    cloud = orc.Cloud.from_packed(packed)
    cam = scr.Camera.perspective((10.0, 4.0, 12.0), (0, 0, 0), fov=45, aspect=1.0)
    cfg = scr.ComputeRasteriserConfiguration(mode=1, enableCLOD=False, minimumPointSize=1, maximumPointSize=4,
                                             pointSizeScale=30.0)
    out = orc.render_nearest(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, 64, 64)
    d, idx = out["depth"], out["index"]
    covered = d > 0
    assert covered.sum() > 50
    # every covered pixel's winner maps to a real point with that pixel's colour
    rgba = out["rgba8"]
    assert np.all((rgba[covered] & 0x00FFFFFF) == (cloud.colors[idx[covered]] & 0x00FFFFFF))
    assert np.all((rgba[covered] >> 24) == 255)


def _round_f32(v: Fraction) -> np.float32:
    """The binary32 nearest to the exact rational v (ties to even; normal range)."""
    if v == 0:
        return np.float32(0.0)
    sign, a = (-1, -v) if v < 0 else (1, v)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if a < Fraction(2) ** e:
        e -= 1
    m = a / Fraction(2) ** (e - 23)          # in [2^23, 2^24)
    n = m.numerator // m.denominator
    rem = m - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return np.float32(sign * float(Fraction(n) * Fraction(2) ** (e - 23)))


def _fma(a, b, c) -> np.float32:
    return _round_f32(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c)))


def test_point_transform_is_fused_in_source_order():
    # fma(c2, z, fma(c1, y, c0 x)) + c3 per row, each fma rounded once (exact
    # rationals); the unfused chain differs often enough that it would fail
    rng = np.random.default_rng(0x5EED)
    differ = 0
    for _ in range(1000):
        m = (rng.uniform(0.5, 2.0, 16) * rng.choice([-1.0, 1.0], 16)).astype(np.float32)
        p = (rng.uniform(0.5, 2.0, 3) * rng.choice([-1.0, 1.0], 3)).astype(np.float32)
        got = orc.transform_point(m, p)
        for i in range(4):
            a = np.float32(m[i] * p[0])
            a = _fma(m[4 + i], p[1], a)
            a = _fma(m[8 + i], p[2], a)
            a = np.float32(a + m[12 + i])
            assert got[i] == a, (m, p, i)
            u = np.float32(np.float32(np.float32(m[i] * p[0]) + np.float32(m[4 + i] * p[1])) +
                           np.float32(m[8 + i] * p[2]))
            differ += int(np.float32(u + m[12 + i]) != a)
    assert differ > 200


def test_decode_point_is_fused_and_the_swift_reference_is_not():
    # decodePoint of the passes: fma(float(q), size * 2^-bits, min); the
    # NearestPointCPUReference restatement rounds the product first
    rng = np.random.default_rng(0xDEC0)
    b = np.zeros(1, dtype=orc.RASTER_BATCH_DTYPE)
    differ = 0
    for _ in range(1000):
        mn = rng.uniform(-3000.0, 3000.0, 3).astype(np.float32)
        size = rng.uniform(1.0, 500.0, 3).astype(np.float32)
        b["minX"], b["minY"], b["minZ"] = mn
        b["maxX"], b["maxY"], b["maxZ"] = mn + size
        level = int(rng.integers(0, 3))
        words = [int(w) for w in rng.integers(0, 1 << 30, 3)]
        got = orc.decode_point(b[0], words, level)
        ref = orc.decode_point(b[0], words, level, fused=False)
        bits = (30, 20, 10)[level]
        for a in range(3):
            lo, me, hi = ((w >> (10 * a)) & 1023 for w in words)
            q = ((lo << 20) | (me << 10) | hi) if level == 0 else (((lo << 10) | me) if level == 1 else lo)
            ext = np.float32(b["max" + "XYZ"[a]][0] - b["min" + "XYZ"[a]][0])
            s = np.float32(max(ext, np.float32(1e-6)) * np.float32(2.0 ** -bits))
            assert got[a] == _fma(np.float32(q), s, mn[a]), (a, q, s, mn[a])
            assert ref[a] == np.float32(np.float32(np.float32(q) * s) + mn[a])
            differ += int(got[a] != ref[a])
    assert differ > 100


def _hole_fill_loops(img, iterations):
    """holeFillUpdate restated with scalar binary32 loops (tiny images only)."""
    f32 = np.float32
    k = f32(1) / f32(255)
    h, w = img.shape
    cur = img.copy()
    for _ in range(iterations):
        nxt = cur.copy()
        for y in range(h):
            for x in range(w):
                c = int(cur[y, x])
                if (c >> 24) >= 128:
                    continue
                s = [f32(0), f32(0), f32(0)]
                n = f32(0)
                for oy in (-1, 0, 1):
                    for ox in (-1, 0, 1):
                        if (ox, oy) == (0, 0) or not (0 <= x + ox < w and 0 <= y + oy < h):
                            continue
                        q = int(cur[y + oy, x + ox])
                        if (q >> 24) < 128:
                            continue
                        for ch in range(3):
                            s[ch] = f32(s[ch] + f32((q >> (8 * ch)) & 0xFF) * k)
                        n = f32(n + f32(1))
                if n > 0:
                    r = f32(f32(1) / n)
                    b = [int(np.rint(f32(min(max(f32(s[ch] * r), f32(0)), f32(1)) * f32(255)))) for ch in range(3)]
                    nxt[y, x] = b[0] | (b[1] << 8) | (b[2] << 16) | 0xFF000000
        cur = nxt
    return cur


def test_hole_fill_closed_forms_and_loops():
    # an empty texel ringed by one colour takes that colour exactly; a texel
    # with one valid neighbour copies it; one with none stays; valid texels
    # (alpha byte >= 128, i.e. alpha >= 0.5) never change
    img = np.zeros((5, 5), np.uint32)
    img[1:4, 1:4] = 0xFF336699
    img[2, 2] = 0x00000000
    out = orc.hole_fill(img, 1)
    assert out[2, 2] == 0xFF336699 and out[1, 1] == 0xFF336699
    img = np.zeros((3, 4), np.uint32)
    img[0, 0] = 0x80FF0010      # alpha 128: valid
    img[2, 3] = 0x7F00FF00      # alpha 127: empty
    out = orc.hole_fill(img, 1)
    assert out[0, 0] == 0x80FF0010 and out[1, 1] == 0xFFFF0010 and out[2, 3] == 0x7F00FF00
    assert out[2, 0] == 0 and out[0, 3] == 0
    rng = np.random.default_rng(3)
    for shape in ((1, 1), (1, 7), (6, 5), (9, 9)):
        img = rng.integers(0, 2 ** 32, shape, dtype=np.uint64).astype(np.uint32)
        img[rng.random(shape) < 0.6] &= 0x00FFFFFF          # mostly holes
        for it in (1, 2, 3):
            assert np.array_equal(orc.hole_fill(img, it), _hole_fill_loops(img, it)), (shape, it)


def test_expf_is_exp_rounded_once():
    # contract C8: e^x from double + - * / steps, rounded once to binary32;
    # against exp in float64 rounded to binary32 (the correctly rounded value
    # except for inputs within 1e-16 of a rounding boundary, none drawn here)
    rng = np.random.default_rng(8)
    xs = np.concatenate([rng.uniform(-104, 88, 20000), rng.uniform(-1, 1, 5000), rng.uniform(0, 16, 5000),
                         -np.arange(0, 70, 0.125)]).astype(np.float32)
    bad = 0
    for x in xs:
        got = np.float32(orc.expf(float(x)))
        want = np.float32(np.exp(np.float64(x)))
        bad += int(got.view(np.uint32) != want.view(np.uint32))
    assert bad == 0
    assert orc.expf(0.0) == 1.0 and orc.expf(float("-inf")) == 0.0 and orc.expf(float("inf")) == float("inf")
    assert math.isnan(orc.expf(float("nan"))) and orc.expf(-120.0) == 0.0 and orc.expf(89.0) == float("inf")
    assert orc.expf(1.0) == np.float32(math.e)
