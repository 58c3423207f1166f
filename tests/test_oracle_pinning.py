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


def pin_batch_projection():
    """The draw passes' projection of an undisplaced point: decodePoint
    (Common.metal:73-117) and the file transform folded per batch into
    A[c][r] = T[c][r] * step_c (rounded once), B = T (min, 1) (transform_point,
    pinned above), clip_r = fma(A[2][r], qz, fma(A[1][r], qy, fma(A[0][r], qx,
    B_r))) with each fma rounded once (exact rationals); and, where every
    product and sum is exact (small integer T, power-of-two steps, lattice
    minimum), it equals T (min + q * step, 1) exactly."""
    rng = np.random.default_rng(0xAFF1)
    for _ in range(600):
        T = (rng.uniform(0.5, 2.0, 16) * rng.choice([-1.0, 1.0], 16)).astype(np.float32)
        mn = rng.uniform(-3000.0, 3000.0, 3).astype(np.float32)
        bits = int(rng.choice([10, 20, 30]))
        step = (rng.uniform(1.0, 500.0, 3).astype(np.float32) * np.float32(2.0 ** -bits)).astype(np.float32)
        qf = np.float32(rng.integers(0, 1 << bits, 3))
        got = orc.project_q(T, mn, step, qf)
        B = orc.transform_point(T, mn)
        for r in range(4):
            a = B[r]
            for c in range(3):
                a = _fma(np.float32(T[c * 4 + r] * step[c]), qf[c], a)
            assert got[r] == a, (T, mn, step, qf, r)
    for _ in range(300):
        T = rng.integers(-8, 9, 16).astype(np.float32)
        mn = (rng.integers(-64, 65, 3) / 4.0).astype(np.float32)
        step = np.float32(2.0) ** rng.integers(-10, -2, 3).astype(np.float32)
        qf = np.float32(rng.integers(0, 1024, 3))
        p = mn.astype(np.float64) + qf.astype(np.float64) * step.astype(np.float64)
        want = T.astype(np.float64).reshape(4, 4).T @ np.append(p, 1.0)
        assert np.array_equal(orc.project_q(T, mn, step, qf).astype(np.float64), want), (T, mn, step, qf)


def test_batch_projection_is_the_fused_affine_map():
    pin_batch_projection()


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


# ---------------------------------------------------------------- cull geometry
# The pins below are module-level functions so that the mutation test
# (test_oracle_mutants.py) can run them against oracles with planted mistakes.

def _look_at(eye, target, up=(0.0, 1.0, 0.0)):
    return scr.look_at(eye, target, up)


# View-space boxes for a camera at the origin looking down -Z.  Perspective:
# fov 90 deg, aspect 2, near 1, far 100, so the frustum is |x| <= -2z,
# |y| <= -z, -100 <= z <= -1 (an x/y mix-up fails the 12..15 cases).
PERSPECTIVE_BOXES = [
    ((-0.5, -0.5, -10.5), (0.5, 0.5, -9.5), True),
    ((-30.0, -0.5, -10.5), (-25.0, 0.5, -9.5), False),   # left of the frustum
    ((-22.0, -0.5, -10.5), (-18.0, 0.5, -9.5), True),    # across the left plane
    ((25.0, -0.5, -10.5), (30.0, 0.5, -9.5), False),     # right
    ((18.0, -0.5, -10.5), (22.0, 0.5, -9.5), True),
    ((-0.5, -15.0, -10.5), (0.5, -12.0, -9.5), False),   # below
    ((-0.5, -11.0, -10.5), (0.5, -9.0, -9.5), True),
    ((-0.5, 12.0, -10.5), (0.5, 15.0, -9.5), False),     # above
    ((-0.5, 9.0, -10.5), (0.5, 11.0, -9.5), True),
    ((12.0, -0.5, -10.5), (15.0, 0.5, -9.5), True),      # inside only because the frustum is twice as wide
    ((-0.1, -0.1, -0.8), (0.1, 0.1, -0.2), False),       # before the near plane
    ((-0.1, -0.1, -1.5), (0.1, 0.1, -0.5), True),
    ((-1.0, -1.0, -150.0), (1.0, 1.0, -120.0), False),   # past the far plane
    ((-1.0, -1.0, -110.0), (1.0, 1.0, -90.0), True),
    ((-1.0, -1.0, 5.0), (1.0, 1.0, 10.0), False),        # behind the camera
]
# Orthographic, 10 high, aspect 1, near 1, far 50: |x|, |y| <= 5, -50 <= z <= -1.
ORTHO_BOXES = [
    ((-1.0, -1.0, -20.0), (1.0, 1.0, -10.0), True),
    ((-9.0, -1.0, -20.0), (-6.0, 1.0, -10.0), False),
    ((-6.0, -1.0, -20.0), (-4.0, 1.0, -10.0), True),
    ((6.0, -1.0, -20.0), (9.0, 1.0, -10.0), False),
    ((-1.0, 6.0, -20.0), (1.0, 9.0, -10.0), False),
    ((-1.0, -9.0, -20.0), (1.0, -6.0, -10.0), False),
    ((-1.0, -1.0, -1.5), (1.0, 1.0, -1.2), True),        # just past the near plane
    ((-1.0, -1.0, -0.9), (1.0, 1.0, -0.1), False),
    ((-1.0, -1.0, -80.0), (1.0, 1.0, -60.0), False),     # past the far plane
    ((-1.0, -1.0, -55.0), (1.0, 1.0, -45.0), True),
]


def pin_frustum():
    """intersectsFrustum (Common.metal:119-153): boxes inside, outside and
    across each of the six planes, perspective and orthographic, also with
    the camera moved (the planes' translation column)."""
    P = scr.perspective_reverse_z(90.0, 2.0, 1.0, 100.0)
    O = scr.orthographic_reverse_z(10.0, 1.0, 1.0, 50.0)
    for proj, boxes in ((P, PERSPECTIVE_BOXES), (O, ORTHO_BOXES)):
        for shift in (0.0, 100.0):
            V = _look_at((shift, 0.0, 0.0), (shift, 0.0, -1.0))
            m = orc.mat_mul(proj, V)
            for mn, mx, want in boxes:
                lo = (mn[0] + shift, mn[1], mn[2])
                hi = (mx[0] + shift, mx[1], mx[2])
                assert orc.intersects_frustum(m, lo, hi) == want, (proj[3, 3], shift, mn, mx)


def test_frustum_planes():
    pin_frustum()


def pin_pixel_size_and_precision():
    """pixelSizeOnScreen (Common.metal:155-172): the box centre and the
    centre moved by the box diagonal R along view x share clip w, so the
    distance is W/2 * P00 * R / w (perspective w = view depth, orthographic
    w = 1), for any centre; precisionLevel (:174-181) cuts at 100, 200, 500
    and 10,000 px."""
    rng = np.random.default_rng(11)
    for fov, W, H in ((90.0, 1000, 1000), (50.0, 1920, 1080), (20.0, 640, 360)):
        P = scr.perspective_reverse_z(fov, W / H, 0.1, 1e5)
        for _ in range(40):
            d = float(rng.uniform(2.0, 400.0))
            c = np.array([rng.uniform(-0.3, 0.3) * d, rng.uniform(-0.3, 0.3) * d, -d])
            half = rng.uniform(0.05, 3.0, 3)
            mn, mx = (c - half).astype(np.float32), (c + half).astype(np.float32)
            R = float(np.linalg.norm(mx.astype(np.float64) - mn.astype(np.float64)))
            cz = (float(mn[2]) + float(mx[2])) / 2.0
            want = W / 2.0 * float(P[0, 0]) * R / -cz
            got = orc.pixel_size_on_screen(mn, mx, np.eye(4, dtype=np.float32), P, W, H)
            assert abs(got - want) <= 1e-4 * want + 1e-3, (fov, W, H, got, want)
    O = scr.orthographic_reverse_z(8.0, 2.0, 0.5, 100.0)
    mn, mx = np.float32([-1.0, -2.0, -10.0]), np.float32([2.0, 2.0, -8.0])
    got = orc.pixel_size_on_screen(mn, mx, np.eye(4, dtype=np.float32), O, 800, 400)
    assert abs(got - 400.0 * float(O[0, 0]) * float(np.linalg.norm(mx - mn))) < 1e-3
    # precision level around each threshold
    for px, level in ((0.0, 4), (99.99, 4), (100.0, 3), (199.99, 3), (200.0, 2), (499.9, 2), (500.0, 1),
                      (9999.0, 1), (10000.0, 0), (1e9, 0)):
        assert orc.precision_level(px) == level, (px, level)
    # the same through pixelSizeOnScreen: an on-axis box of diagonal R at depth d
    P = scr.perspective_reverse_z(90.0, 1.0, 0.1, 1e5)      # P00 = 1: px = 500 R / d at W = 1000
    for want_px, level in ((99.0, 4), (101.0, 3), (199.0, 3), (201.0, 2), (499.0, 2), (501.0, 1),
                           (9990.0, 1), (10010.0, 0)):
        R = want_px * 4.0 / 500.0
        mn, mx = np.float32([-R / 2, 0.0, -4.0]), np.float32([R / 2, 0.0, -4.0])
        px = orc.pixel_size_on_screen(mn, mx, np.eye(4, dtype=np.float32), P, 1000, 1000)
        assert abs(px - want_px) < 0.05 and orc.precision_level(px) == level, (want_px, px)


def test_pixel_size_and_precision_level():
    pin_pixel_size_and_precision()


def pin_lod_threshold():
    """lodThresholdFromPixelSize (Common.metal:185-187): log2(max(px, 1) / 25)
    + 3 + bias, so 25 * 2^k px gives exactly k + 3 + bias; and the cull's use
    of it (Cull/Shaders.metal:55-74): CLOD off gives 99 and every point,
    CLOD on gives max(0, threshold) and the prefix cum[clamp(floor(t + 0.5),
    0, 7)] of the batch's level counts."""
    for k in range(-4, 12):
        for bias in (-2, 0, 3):
            assert orc.lod_threshold(25.0 * 2.0 ** k, bias) == float(k + 3 + bias), (k, bias)
    assert orc.lod_threshold(0.25, 0) == orc.lod_threshold(1.0, 0)       # max(px, 1)
    assert abs(orc.lod_threshold(1.0, 0) - (math.log2(1 / 25) + 3)) < 1e-6
    # one batch of 100 points with level counts cum = 1, 5, 5, 40, 60, 75, 90, 100
    cum = [1, 5, 5, 40, 60, 75, 90, 100]
    levels = np.repeat(np.arange(8, dtype=np.uint8), np.diff([0] + cum))
    pos = np.zeros((100, 3), np.float32)
    pos[:, 0] = np.linspace(-0.1, 0.1, 100, dtype=np.float32)
    packed = orc.pack(pos, np.ones((100, 4), np.float32), pointsPerBatch=100, lodLevels=1)
    cloud = orc.Cloud.from_packed(packed)
    cloud.levels[:100] = levels
    b = cloud.batches
    b["lodCum01"], b["lodCum23"] = cum[0] | (cum[1] << 16), cum[2] | (cum[3] << 16)
    b["lodCum45"], b["lodCum67"] = cum[4] | (cum[5] << 16), cum[6] | (cum[7] << 16)
    P = scr.perspective_reverse_z(90.0, 1.0, 0.1, 1e4)
    box_r = float(np.linalg.norm([b["maxX"][0] - b["minX"][0], b["maxY"][0] - b["minY"][0],
                                  b["maxZ"][0] - b["minZ"][0]]))
    for k, bias in ((2, 0), (0, 0), (1, 1), (4, 2), (-3, -1)):
        # a camera at distance d along +z sees the batch at 25 * 2^k px (W = 1000, P00 = 1)
        d = 500.0 * box_r / (25.0 * 2.0 ** k)
        V = _look_at((0.0, 0.0, d), (0.0, 0.0, 0.0))
        for clod in (True, False):
            cfg = scr.ComputeRasteriserConfiguration(enableCLOD=clod, lodBias=bias, enableFrustumCulling=False)
            vb = orc.cull(cloud, cfg, V, P, 1000, 1000)
            assert len(vb) == 1
            if not clod:
                assert vb["lodThreshold"][0] == 99.0 and vb["activePoints"][0] == 100
                continue
            t = max(0.0, k + 3 + bias)
            assert abs(vb["lodThreshold"][0] - t) < 1e-3, (k, bias, vb["lodThreshold"][0])
            lmax = min(max(math.floor(t + 0.5), 0), 7)
            assert vb["activePoints"][0] == cum[lmax], (k, bias, lmax)


def test_lod_threshold_and_survivor_prefix():
    pin_lod_threshold()


def pin_footprint_radius():
    """pointFootprintRadius (Common.metal:18-61) against hand-computed radii:
    size = scale / |view position| (screen space) or 2 * scale * focal /
    view depth (world space, focal = H / 2 * P11), or scale / 2 * scale *
    focal under an orthographic projection; non-finite sizes take the
    minimum; clamped to [lo, hi]; radius = ceil(max(size - 1, 0) / 2) capped
    at 16."""
    I = np.eye(4, dtype=np.float32)
    P = scr.perspective_reverse_z(90.0, 1.0, 0.1, 1000.0)            # P11 = 1
    O = scr.orthographic_reverse_z(4.0, 1.0, 0.1, 1000.0)            # P11 = 0.5
    S, Wd = scr.POINT_SIZE_SCREEN_SPACE, scr.POINT_SIZE_WORLD_SPACE
    cases = [
        # (mode, min, max, scale, point, proj, H, radius)
        (S, 1, 9, 10.0, (0, 0, -2), P, 100, 2),        # size 5
        (S, 1, 9, 10.0, (0, 0, -4), P, 100, 1),        # size 2.5: ceil(0.75)
        (S, 1, 9, 10.0, (0, 0, -10), P, 100, 0),       # size 1
        (S, 1, 9, 10.0, (0, 0, -1), P, 100, 4),        # size 10 -> clamped to 9
        (S, 1, 9, 16.0, (4, 0, -3), P, 100, 2),        # |v| = 5: size 3.2 (view depth 3 would give 5.3)
        (S, 3, 9, float("nan"), (0, 0, -2), P, 100, 1),  # non-finite: the minimum, 3
        (S, 1, 100, 1e6, (0, 0, -2), P, 100, 16),      # size 100 -> 50, capped at 16
        (Wd, 1, 16, 0.14, (4, 0, -3), P, 100, 2),      # 2 * 0.14 * 50 / 3 = 4.67 (|v| = 5 would give 2.8)
        (Wd, 1, 16, 0.07, (0, 0, -2), P, 100, 2),      # 2 * 0.07 * 50 / 2 = 3.5: ceil(1.25)
        (S, 1, 16, 7.0, (3, 1, -9), O, 100, 3),        # orthographic screen space: size = scale
        (Wd, 1, 16, 0.09, (3, 1, -9), O, 100, 2),      # orthographic world space: 2 * 0.09 * 25 = 4.5
    ]
    for mode, lo, hi, scale, pt, proj, H, want in cases:
        cfg = scr.ComputeRasteriserConfiguration(pointSizeMode=mode, minimumPointSize=lo, maximumPointSize=hi,
                                                 pointSizeScale=scale)
        got = orc.footprint_radius(np.float32(pt), cfg, I, proj, 2 * H, H)
        assert got == want, (mode, lo, hi, scale, pt, got, want)
    # the file's world matrix moves the point first: (0, 0, 0) placed at z = -2
    cfg = scr.ComputeRasteriserConfiguration(minimumPointSize=1, maximumPointSize=9, pointSizeScale=10.0)
    world = np.eye(4, dtype=np.float32)
    world[3, 2] = -2.0
    assert orc.footprint_radius(np.float32([0, 0, 0]), cfg, I, P, 200, 100, world=world) == 2


def test_footprint_radius_closed_forms():
    pin_footprint_radius()


# ---------------------------------------------------------------- nearest
def _lattice_cloud(n=24, dup=40):
    """A single-batch lattice whose decoded positions are exact (box size 2
    and 1, 10-bit decode) plus `dup` exact duplicates of lattice points."""
    g = np.linspace(-1.0, 1.0, n, dtype=np.float32)
    X, Y = np.meshgrid(g, g, indexing="ij")
    z = (np.arange(n * n) % 7 + 1).astype(np.float32) / 8.0
    pos = np.stack([X.ravel(), Y.ravel(), z], 1)
    pos = np.concatenate([pos, [[-1.0, -1.0, 0.0], [1.0, 1.0, 1.0]], pos[::(n * n) // dup][:dup]]).astype(np.float32)
    col = np.random.default_rng(3).random((len(pos), 4)).astype(np.float32)
    packed = orc.pack(pos, col, pointsPerBatch=len(pos), lodLevels=1)
    return orc.Cloud.from_packed(packed)


def pin_nearest_matches_cpu_reference():
    """The render passes' nearest mode (NearestDepth/NearestIndex) equals the
    restated NearestPointCPUReference (NearestPointCPUReference.swift:16-81)
    where their arithmetic coincides (exact lattice decode, identity camera):
    same depths, same winning indices, duplicates resolved to the lowest
    index; plus a hand-built tie at one pixel."""
    cloud = _lattice_cloud()
    I = np.eye(4, dtype=np.float32)
    cfg = scr.ComputeRasteriserConfiguration(mode=scr.MODE_NEAREST_POINT, enableCLOD=False,
                                             enableFrustumCulling=False)
    for W, H in ((64, 64), (48, 30)):
        out = orc.render_nearest(cloud, cfg, I, I, W, H)
        ref = orc.nearest_cpu_reference(cloud, W, H, I, I)
        assert (out["depth"] > 0).sum() > 100
        assert np.array_equal(out["depth"].ravel(), ref["depths"])
        assert np.array_equal(out["index"].ravel(), ref["indices"])
    # three points on one pixel: two at the same (front) depth, one behind
    pos = np.float32([[0.3, 0.3, 0.5], [0.3, 0.3, 0.75], [0.3, 0.3, 0.75], [-0.5, -0.5, 0.25]])
    col = np.float32([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1], [1, 1, 1, 1]])
    packed = orc.pack(pos, col, pointsPerBatch=4, lodLevels=1)
    c = orc.Cloud.from_packed(packed)
    out = orc.render_nearest(c, cfg, I, I, 8, 8)
    front = [i for i in range(4) if packed.sourceIndices[i] in (1, 2)]
    iy = 8 - 1 - int(np.float32(0.65) * 8)
    ix = int(np.float32(0.65) * 8)
    assert out["index"][iy, ix] == min(front), (out["index"][iy, ix], front)
    # z = 0.75 is the box's maximum: it decodes to min + 1023 steps of 10 bits
    assert out["depth"][iy, ix] == orc.depth_to_uint(float(np.float32(0.25 + 1023 * (0.5 / 1024))))


def test_nearest_equals_cpu_reference_and_lowest_index_wins():
    pin_nearest_matches_cpu_reference()


# ---------------------------------------------------------------- weighted blend
def _one_point(x=0.0, y=0.0, z=0.5, rgb=(1.0, 1.0, 1.0)):
    packed = orc.pack([(x, y, z)], [(*rgb, 1.0)], pointsPerBatch=1, lodLevels=1)
    return orc.Cloud.from_packed(packed)


def pin_blend_sums_and_coverage():
    """Weighted-blend accumulation (ColorPass/Shaders.metal:146-198) and its
    resolve (Resolve/Shaders.metal:27-45) for a single point with an
    identity camera: coverage a = 1 - tint alpha, sums uint(c / 255 * a / N *
    4096 + 0.5) on each of the N sweep taps; colour = sum / coverage sum,
    alpha = 1 - e^(-coverage sum / 4096).  Values chosen so the exact
    rationals sit far from rounding boundaries."""
    I = np.eye(4, dtype=np.float32)
    W = H = 64
    cfg = scr.ComputeRasteriserConfiguration(applyTint=True, tintAlphaIsCoverage=True, enableCLOD=False,
                                             enableFrustumCulling=False)
    for tint_a, rgb, want_c, want_a in ((0.5, (1.0, 0.2, 0.0), (2048, 410, 0), 2048),
                                        (0.75, (1.0, 1.0, 1.0), (1024, 1024, 1024), 1024),
                                        (0.0, (0.6, 0.4, 1.0), (2458, 1638, 4096), 4096)):
        c = _one_point(rgb=rgb)
        c.tints = np.float32([[0.0, 0.0, 0.0, tint_a]])
        out = orc.render_hqs([c], cfg, I, I, W, H)
        iy, ix = 31, 32
        assert (out["red"][iy, ix], out["green"][iy, ix], out["blue"][iy, ix]) == want_c
        assert out["count"][iy, ix] == want_a and out["count"].sum() == want_a
        rgba = int(out["rgba8"][iy, ix])
        alpha = 1.0 - math.exp(-want_a / 4096.0)
        assert rgba >> 24 == int(np.rint(alpha * 255.0)), (rgba >> 24, alpha)
        for ch in range(3):
            assert (rgba >> (8 * ch)) & 0xFF == int(np.rint(want_c[ch] / want_a * 255.0))
    # motion blur: the previous view-projection shifts the point 8 px left and
    # 4 rows down, so the sweep (scaled by 0.5) is (4, 2): 5 taps, step (1, 0.5)
    c = _one_point()
    prev = np.eye(4, dtype=np.float32)
    prev[3, 0], prev[3, 1] = -0.25, 0.125
    cfg = scr.ComputeRasteriserConfiguration(motionBlur=0.5, motionBlurSamples=8, motionBlurMaxSpread=64.0,
                                             enableCLOD=False, enableFrustumCulling=False)
    out = orc.render_hqs([c], cfg, I, I, W, H, prev_view_proj=prev)
    taps = [(31, 32), (30, 31), (30, 30), (29, 29), (29, 28)]     # (row, column): centre - round(s * step)
    got = {(int(r), int(q)) for r, q in zip(*np.nonzero(out["count"]))}
    assert got == set(taps), sorted(got)
    for r, q in taps:
        assert out["count"][r, q] == 819 and out["red"][r, q] == 819    # uint(4096 / 5 + 0.5)
    # a sweep longer than the sample cap: 8 taps
    cfg.motionBlurSamples = 3
    out = orc.render_hqs([c], cfg, I, I, W, H, prev_view_proj=prev)
    assert (out["count"] > 0).sum() == 3 and out["count"][31, 32] == 1365      # uint(4096 / 3 + 0.5)


def test_blend_sums_sweep_and_coverage_alpha():
    pin_blend_sums_and_coverage()


def test_nearest_independent_of_threads():
    pos, col = jittered_cloud(20000, 0x4E4)
    packed = orc.pack(pos, col, pointsPerBatch=500)
    cloud = orc.Cloud.from_packed(packed)
    cam = scr.Camera.perspective((10.0, 4.0, 12.0), (0, 0, 0), fov=45, aspect=1.5)
    cfg = scr.ComputeRasteriserConfiguration(mode=1, minimumPointSize=1, maximumPointSize=4, pointSizeScale=30.0)
    ref = orc.render_nearest(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, 120, 80)
    assert (ref["depth"] > 0).sum() > 1000
    for t in (2, 5, 16):
        out = orc.render_nearest(cloud, cfg, cam.viewMatrix, cam.projectionMatrix, 120, 80, threads=t)
        for k in ref:
            assert np.array_equal(ref[k], out[k]), (t, k)


def test_gpu_lod_gate_closed_forms():
    """GPUPacker.init (GPUPacker.swift:35, 86-99): LOD runs when
    (coarse * 2^(L - 2) + 2)^3 <= 268,435,456, else every level stays L - 1."""
    for levels, coarse, on in ((4, 64, True), (5, 64, True), (6, 64, False), (8, 64, False), (1, 64, False),
                               (6, 40, True), (6, 41, False), (2, 643, True), (2, 644, False), (3, 321, True),
                               (3, 322, False)):
        assert orc.gpu_lod_enabled(levels, coarse) == on, (levels, coarse)
        dim = coarse * 2 ** (levels - 2) + 2
        assert on == (levels > 1 and dim ** 3 <= 268_435_456)
    pos, col = jittered_cloud(2000, 0x10D)
    off = orc.pack(pos, col, pointsPerBatch=300, lodLevels=8, coarseVoxelDivisions=64, gpuRules=True)
    assert np.all(off.levels == 7)
    cum = off.batches["lodCum67"] >> 16
    assert np.array_equal(cum, off.batches["numPoints"])
    assert np.all(off.batches["lodCum01"] == 0) and np.all(off.batches["lodCum45"] == 0)
    # below the cap the gate changes nothing
    a = orc.pack(pos, col, pointsPerBatch=300, lodLevels=6, coarseVoxelDivisions=4, gpuRules=True)
    b = orc.pack(pos, col, pointsPerBatch=300, lodLevels=6, coarseVoxelDivisions=4)
    assert np.array_equal(a.levels, b.levels) and a.batches.tobytes() == b.batches.tobytes()
    assert len(set(a.levels.tolist())) > 2


def pin_displacement():
    """Displacement (DepthPass/Shaders.metal:69-74, ColorPass/Shaders.metal:
    72-78, 150-153) for one point under an identity camera, 64 x 64: drawn at
    position + displacement; a NaN component removes it (also from the
    weighted blend, which has no depth test); a cloud without a buffer draws
    undisplaced; nearest mode ignores it; the motion-blur sweep runs from the
    current pixel back to that of position + previous displacement."""
    I = np.eye(4, dtype=np.float32)
    W = H = 64
    base = dict(enableCLOD=False, enableFrustumCulling=False)
    c = _one_point()
    c.disp = np.float32([[0.5, 0.0, 0.0, 0.0]])
    out = orc.render_hqs([c], scr.ComputeRasteriserConfiguration(applyDisplacement=True, **base), I, I, W, H)
    assert out["count"][31, 48] == 1 and out["count"].sum() == 1               # x: int(0.75 * 64)
    out = orc.render_hqs([c], scr.ComputeRasteriserConfiguration(**base), I, I, W, H)
    assert out["count"][31, 32] == 1 and out["count"].sum() == 1                # applyDisplacement off
    nearest = orc.render_nearest(c, scr.ComputeRasteriserConfiguration(mode=1, applyDisplacement=True, **base),
                                 I, I, W, H)
    assert nearest["index"][31, 32] == 0 and (nearest["depth"] > 0).sum() == 1
    bare = _one_point()
    out = orc.render_hqs([bare], scr.ComputeRasteriserConfiguration(applyDisplacement=True, **base), I, I, W, H)
    assert out["count"][31, 32] == 1 and out["count"].sum() == 1                # no buffer: zeros
    for nan_at in range(3):
        c.disp = np.zeros((1, 4), np.float32)
        c.disp[0, nan_at] = np.nan
        for extra in (dict(), dict(applyTint=True, tintAlphaIsCoverage=True), dict(motionBlur=0.5)):
            cfg = scr.ComputeRasteriserConfiguration(applyDisplacement=True, **base, **extra)
            out = orc.render_hqs([c], cfg, I, I, W, H)
            assert out["count"].sum() == 0 and out["depth"].sum() == 0, (nan_at, extra)
    # sweep: current displacement (0.25, 0, 0) puts the point at column 40,
    # previous (0, 0, 0) at 32; blur 0.5 gives 4 px: taps 40, 39, 37, 36
    c.disp = np.float32([[0.25, 0.0, 0.0, 0.0]])
    c.prevDisp = np.zeros((1, 4), np.float32)
    cfg = scr.ComputeRasteriserConfiguration(applyDisplacement=True, motionBlur=0.5, motionBlurSamples=8, **base)
    out = orc.render_hqs([c], cfg, I, I, W, H)
    got = {(int(r), int(q)) for r, q in zip(*np.nonzero(out["count"]))}
    assert got == {(31, 40), (31, 39), (31, 37), (31, 36)}, sorted(got)
    assert all(out["count"][31, q] == 1024 for q in (40, 39, 37, 36))
    # without a previous copy the sweep starts from the current displacement: no velocity
    c.prevDisp = None
    out = orc.render_hqs([c], cfg, I, I, W, H)
    assert out["count"][31, 40] == 4096 and out["count"].sum() == 4096


def test_displacement_closed_forms():
    pin_displacement()


def pin_composite():
    """ComputeRasteriser.draw's composite (Post.metal:23-50, ComputeRasteriser.
    swift:672-683) against exact arithmetic and closed forms."""
    # source-over on every source byte and alpha against a spread of target
    # bytes: the byte written is the exact rational (s*a + t*(255 - a)) / 255
    # rounded to nearest, wherever that value is not within 1e-3 of a tie;
    # alpha is a + ta*(255 - a) / 255 the same way
    s_, a_, t_ = np.meshgrid(np.arange(256), np.arange(256), np.arange(0, 256, 5), indexing="ij")
    s_, a_, t_ = (x.ravel().astype(np.int64) for x in (s_, a_, t_))
    src = (s_ | (s_ << 8) | (s_ << 16) | (a_ << 24)).astype(np.uint32)
    dst = (t_ | ((255 - t_) << 8) | (t_ << 16) | (t_ << 24)).astype(np.uint32)
    one = np.ones(src.shape, np.float32)
    out, _ = orc.composite(src, one, dst, one * 0, mode=orc.COMPOSITE_OVER)
    for shift, dv in ((0, t_), (8, 255 - t_), (16, t_)):
        num = s_ * a_ + dv * (255 - a_)                 # value * 255
        exact = num / 255.0
        clear = np.abs(exact - np.floor(exact) - 0.5) > 1e-3
        want = np.floor(exact + 0.5).astype(np.int64)
        got = (out >> shift) & 0xFF
        assert np.array_equal(got[clear], want[clear]), shift
    exact = a_ + t_ * (255 - a_) / 255.0
    clear = np.abs(exact - np.floor(exact) - 0.5) > 1e-3
    assert np.array_equal(((out >> 24) & 0xFF)[clear], np.floor(exact + 0.5).astype(np.int64)[clear])
    # opaque sources replace, transparent ones leave the target
    assert np.array_equal(out[a_ == 255], src[a_ == 255])
    assert np.array_equal(out[a_ == 0], dst[a_ == 0])
    # depth mode: no cloud (depth <= 0, -0 included) never draws; the compare
    # decides ties; the depth write follows the compare
    d = np.float32([0.0, -0.0, -1.0, 0.5, 0.5, 0.5, 0.25])
    td = np.float32([0.0, 0.0, 0.0, 0.25, 0.5, 0.75, 0.25])
    c = np.full(7, 0xFF102030, np.uint32)
    t = np.full(7, 0x80405060, np.uint32)
    drawn = {orc.COMPARE_ALWAYS: [0, 0, 0, 1, 1, 1, 1], orc.COMPARE_GREATER: [0, 0, 0, 1, 0, 0, 0],
             orc.COMPARE_GREATER_EQUAL: [0, 0, 0, 1, 1, 0, 1], orc.COMPARE_LESS: [0, 0, 0, 0, 0, 1, 0],
             orc.COMPARE_LESS_EQUAL: [0, 0, 0, 0, 1, 1, 1]}
    for cmp, want in drawn.items():
        for write in (False, True):
            o, od = orc.composite(c, d, t, td, depthCompare=cmp, writeDepth=write, blend=orc.BLEND_REPLACE)
            w = np.array(want, bool)
            assert np.array_equal(o == c, w), (cmp, write)
            assert np.array_equal(od, np.where(w, d, td) if write else td), (cmp, write)
    # the colour-only mode draws every texel, whatever the depths
    o, _ = orc.composite(c, d, t, None, mode=orc.COMPOSITE_OVER, blend=orc.BLEND_REPLACE)
    assert np.array_equal(o, c)
    # a depth-tested replace with depth writes is a z-buffer: two layers with
    # distinct depths give the same image in either order (greater = nearer)
    rng = np.random.default_rng(5)
    ca, cb, tt = (rng.integers(0, 2**32, 4096, dtype=np.uint64).astype(np.uint32) for _ in range(3))
    da, db = (rng.choice(np.float32([0.0, 0.1, 0.2, 0.3, 0.4]), 4096) for _ in range(2))
    db = np.where(db == da, np.float32(0.45), db).astype(np.float32)
    z0 = np.zeros(4096, np.float32)
    kw = dict(blend=orc.BLEND_REPLACE, writeDepth=True)
    ab = orc.composite(cb, db, *orc.composite(ca, da, tt, z0, **kw), **kw)
    ba = orc.composite(ca, da, *orc.composite(cb, db, tt, z0, **kw), **kw)
    assert np.array_equal(ab[0], ba[0]) and np.array_equal(ab[1], ba[1])
    assert np.array_equal(ab[1], np.maximum(da, db))


def test_composite_closed_forms():
    pin_composite()
