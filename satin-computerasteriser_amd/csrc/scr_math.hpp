// scr_math.hpp — device arithmetic of the render passes (gfx950).
//
// Implements the shared arithmetic contract of the rasteriser:
//   * binary32 RNE, source operation order of the Metal passes; no implicit
//     contraction (every .hip file is built with -ffp-contract=off), and
//     explicit fmaf exactly where the per-point front end multiplies then
//     adds: decodePoint's q * s + min, the point transform (mul_point), the
//     batch projection of quantised coordinates (project_q: decode and
//     transform folded into one per-batch affine map) and the viewport's
//     ndc * 0.5 + 0.5;
//   * a / b in the Metal passes is a * (1.0f / b) with the correctly rounded
//     reciprocal (Metal fast-math reciprocal division made exact);
//   * sqrt correctly rounded; log2 = scr_log2f (double-precision, + - * / only);
//   * float→int truncates and saturates, NaN → 0;
//   * RGBA8 bytes = rint(saturate(x) * 255).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/scr.h"

namespace scr {

constexpr float kTwo32 = 4294967296.0f;
constexpr float kInvTwo32 = 2.3283064365386963e-10f;  // 2^-32, exact
constexpr float kInv255 = 1.0f / 255.0f;

__host__ __device__ inline float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__host__ __device__ inline float rcp(float x) { return 1.0f / x; }

// The same correctly rounded 1 / x for 2^-126 <= x < 2^126, in three
// instructions: the hardware reciprocal (within 1 ulp) refined by one
// Newton step, e = 1 - x * r exact through the fused multiply-add (the
// whole range is compared with 1.0f / x, bit for bit, on the device by
// scr_check_fast_reciprocal).  Callers prove the range first.
__device__ inline float rcp_fast(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    const float e = fmaf(-x, r, 1.0f);
    return fmaf(e, r, r);
}

// e^x for binary32 x (the coverage resolve's 1 - e^-sum): evaluated in
// double with + - * / only and rounded to binary32 once, so the device and
// a CPU evaluation of the same steps agree bit for bit.  x = k ln2 + r with
// k = rint(x / ln2) and r = (x - k ln2_hi) - k ln2_lo (the fdlibm split: k
// ln2_hi is exact for |k| < 2^11), |r| <= 0.35; e^r by the Horner form
// 1 + r/1 (1 + r/2 (1 + ... (1 + r/13))), whose truncation error (< 1e-17)
// is below the double rounding; then the exact scaling by 2^k.
__host__ __device__ inline float scr_expf(float xf) {
    if (!(xf == xf)) return xf;
    if (xf > 100.0f) return __builtin_inff();
    if (xf < -110.0f) return 0.0f;
    const double x = (double)xf;
    const double k = rint(x * 1.4426950408889634);
    const double r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    double e = 1.0;
    for (int n = 13; n >= 1; n--) e = 1.0 + (r / (double)n) * e;
    return (float)ldexp(e, (int)k);
}

__host__ __device__ inline uint32_t f2u(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= kTwo32) return 0xffffffffu;
    return (uint32_t)x;
}

__host__ __device__ inline int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)x;
}

// Device forms of f2i / f2u: the gfx950 conversions truncate, saturate and
// map NaN to 0 — exactly the contract above — in one instruction.
__device__ inline int32_t f2i_hw(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
__device__ inline uint32_t f2u_hw(float x) {
    uint32_t r;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Common.metal:10-12 — float(CR_MAX_DEPTH - 1u) is 2^32 in binary32.
__host__ __device__ inline uint32_t depth_to_uint(float z) { return f2u(saturate(z) * kTwo32); }
// Common.metal:14-16
__host__ __device__ inline float uint_to_depth(uint32_t d) { return (float)d * kInvTwo32; }
// depth_to_uint on the device: scaling by 2^32 is exact, and the saturating
// conversion already sends z < 0, -0 and NaN to 0 and z > 1 to 2^32 - 1, so
// the clamp of depth_to_uint changes no result and is left out.
__device__ inline uint32_t depth_to_uint_hw(float z) { return f2u_hw(z * kTwo32); }

// Common.metal:190-197 (lowbias32)
__host__ __device__ inline float hash_unit(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return (float)x * kInvTwo32;
}

// log2 of the cull pass (Common.metal:185-187): x = 2^e m, m in
// [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s) with s = (m-1)/(m+1), series to
// s^21; evaluated in double with + - * / only, rounded once to float.
__host__ __device__ inline float scr_log2f(float x) {
    if (x != x) return x;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_inff();
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    uint32_t expo = (u >> 23) & 0xffu;
    if (expo == 0xffu) return x;
    int e;
    if (expo == 0u) {
        const uint32_t w = __builtin_bit_cast(uint32_t, x * 16777216.0f);
        expo = (w >> 23) & 0xffu;
        e = (int)expo - 127 - 24;
        u = w;
    } else {
        e = (int)expo - 127;
    }
    double m = (double)__builtin_bit_cast(float, (u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    const double lnm = 2.0 * s * p;
    return (float)((double)e + lnm * 1.4426950408889634);
}

struct f4 {
    float x, y, z, w;
};

// Metal float4x4 * float4: ((c0 x + c1 y) + c2 z) + c3 w, column-major m[c*4+r].
__host__ __device__ inline f4 mul(const float* m, f4 v) {
    f4 r;
    r.x = m[0] * v.x;
    r.y = m[1] * v.x;
    r.z = m[2] * v.x;
    r.w = m[3] * v.x;
    r.x = r.x + m[4] * v.y;
    r.y = r.y + m[5] * v.y;
    r.z = r.z + m[6] * v.y;
    r.w = r.w + m[7] * v.y;
    r.x = r.x + m[8] * v.z;
    r.y = r.y + m[9] * v.z;
    r.z = r.z + m[10] * v.z;
    r.w = r.w + m[11] * v.z;
    r.x = r.x + m[12] * v.w;
    r.y = r.y + m[13] * v.w;
    r.z = r.z + m[14] * v.w;
    r.w = r.w + m[15] * v.w;
    return r;
}

// A point (w = 1) through m, the per-point transform of the draw passes:
// the column sums in source order with each product-then-add fused,
// fma(c2, z, fma(c1, y, c0 * x)) + c3 (c3 * 1 is c3 exactly).
__host__ __device__ inline f4 mul_point(const float* m, float x, float y, float z) {
    f4 r;
    r.x = fmaf(m[8], z, fmaf(m[4], y, m[0] * x)) + m[12];
    r.y = fmaf(m[9], z, fmaf(m[5], y, m[1] * x)) + m[13];
    r.z = fmaf(m[10], z, fmaf(m[6], y, m[2] * x)) + m[14];
    r.w = fmaf(m[11], z, fmaf(m[7], y, m[3] * x)) + m[15];
    return r;
}

// The draw passes' projection of an undisplaced point: decodePoint (q * s +
// min) followed by the file transform T, folded per batch into one affine map
// of the quantised coordinates,
//     A[c][r] = T[c][r] * s_c        (one rounding per entry)
//     B       = mul_point(T, min)    (the transformed batch minimum)
//     clip_r  = fma(A[2][r], qz, fma(A[1][r], qy, fma(A[0][r], qx, B_r)))
// with q converted to float as decodePoint converts it.  Three fused
// multiply-adds per clip coordinate instead of the decode's three and the
// transform's four.  (Metal's fast math leaves the rounding of decode and
// transform unspecified; this fixes one order.)
struct BatchAffine {
    float a[12];  // column-major like T: a[c * 4 + r], c = x, y, z
    float b[4];
};
__host__ __device__ inline BatchAffine batch_affine(const float* T, float sx, float sy, float sz, float mnx, float mny,
                                                    float mnz) {
    BatchAffine m;
    const f4 o = mul_point(T, mnx, mny, mnz);
    m.b[0] = o.x;
    m.b[1] = o.y;
    m.b[2] = o.z;
    m.b[3] = o.w;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        m.a[r] = T[r] * sx;
        m.a[4 + r] = T[4 + r] * sy;
        m.a[8 + r] = T[8 + r] * sz;
    }
    return m;
}
__host__ __device__ inline f4 project_q(const BatchAffine& m, float qx, float qy, float qz) {
    f4 r;
    r.x = fmaf(m.a[8], qz, fmaf(m.a[4], qy, fmaf(m.a[0], qx, m.b[0])));
    r.y = fmaf(m.a[9], qz, fmaf(m.a[5], qy, fmaf(m.a[1], qx, m.b[1])));
    r.z = fmaf(m.a[10], qz, fmaf(m.a[6], qy, fmaf(m.a[2], qx, m.b[2])));
    r.w = fmaf(m.a[11], qz, fmaf(m.a[7], qy, fmaf(m.a[3], qx, m.b[3])));
    return r;
}

__host__ __device__ inline void mat_mul(const float* a, const float* b, float* out) {
    float r[16];
    for (int c = 0; c < 4; c++) {
        const f4 col = mul(a, f4{b[c * 4 + 0], b[c * 4 + 1], b[c * 4 + 2], b[c * 4 + 3]});
        r[c * 4 + 0] = col.x;
        r[c * 4 + 1] = col.y;
        r[c * 4 + 2] = col.z;
        r[c * 4 + 3] = col.w;
    }
    for (int i = 0; i < 16; i++) out[i] = r[i];
}

__host__ __device__ inline float len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }

// Camera / point-size constants of one render, shared by every point pass.
struct PassParams {
    float view[16];
    float proj[16];
    int width, height;
    int pointSizeMode;
    float minPointSize, maxPointSize, pointSizeScale;
    int lodDither;
    int enableFrustumCulling;
    int lodBias;
    int enableCLOD;
    float depthKeep;        // 1 - depthTolerance
    int colorizeChunks;
    int colorizeOverdraw;
    int applyTint;
    int applyDisplacement;  // points are drawn at position + displacement (depth, colour, blend passes)
    int constRadius;        // >= 0: footprint radius is this constant (lo == hi); -1: per point
    int maxRadius;          // upper bound of the footprint radius (radius of hi)
    int ortho;              // proj[3][3] == 1
    float focal;            // float(height) * 0.5 * proj[1][1]
    float lo, hi;           // point-size clamp
    float bg[4];
};

// Radius of a point of size ps (the tail of pointFootprintRadius,
// Common.metal:52-60): non-finite sizes take lo, then clamp to [lo, hi];
// radius = clamp(int(ceil(max(ps - 1, 0) * 0.5)), 0, 16).  Non-decreasing in ps.
__host__ __device__ inline int radius_of_size(float ps, float lo, float hi) {
    if (!__builtin_isfinite(ps)) ps = lo;
    ps = fminf(fmaxf(ps, lo), hi);
    int r = f2i(ceilf(fmaxf(ps - 1.0f, 0.0f) * 0.5f));
    r = r < 0 ? 0 : r;
    return r > 16 ? 16 : r;
}

// pointFootprintRadius — Common.metal:18-61 (world/view transform only when
// the size depends on the point).
__host__ __device__ inline int footprint_radius(const PassParams& p, const float* world, float x, float y, float z) {
    if (p.constRadius >= 0) return p.constRadius;
    const f4 wp = mul_point(world, x, y, z);
    const f4 v = mul_point(p.view, wp.x, wp.y, wp.z);
    float ps;
    if (p.pointSizeMode == 1) {
        if (p.ortho) {
            ps = (2.0f * p.pointSizeScale) * p.focal;
        } else {
            const float viewZ = fmaxf(-v.z, 0.000001f);
            ps = ((2.0f * p.pointSizeScale) * p.focal) * rcp(viewZ);
        }
    } else {
        if (p.ortho) {
            ps = p.pointSizeScale;
        } else {
            const float d = fmaxf(len3(v.x, v.y, v.z), 0.000001f);
            ps = p.pointSizeScale * rcp(d);
        }
    }
    return radius_of_size(ps, p.lo, p.hi);
}

// Footprint radius shared by every point of a batch, or -1.
//
// In the perspective modes the size is k * rcp(d), d = max(|v|, 1e-6)
// (screen space) or max(-v.z, 1e-6) (world space), v = view * world * point.
// Bounding v over the batch box (the box's corners through view * world,
// widened by 1e-4 of the magnitudes involved, which dwarfs the float error of
// the per-point products, and d by a further 1e-4 relative) brackets the d
// every point computes; rcp, a product with k >= 0 and the radius are all
// non-decreasing or non-increasing in d, so equal radii at the two ends mean
// that radius for every point of the batch.
__host__ __device__ inline int batch_const_radius(const PassParams& p, const float* world, const scr_raster_batch& b) {
    if (p.constRadius >= 0) return p.constRadius;
    if (p.ortho) return -1;
    const float k = p.pointSizeMode == 1 ? (2.0f * p.pointSizeScale) * p.focal : p.pointSizeScale;
    if (!(k >= 0.0f) || !__builtin_isfinite(k)) return -1;
    float m[16];
    mat_mul(p.view, world, m);
    const float bmin[3] = {b.minX, b.minY, b.minZ}, bmax[3] = {b.maxX, b.maxY, b.maxZ};
    float lo[3], hi[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float t = m[12 + i];
        float l = t, h = t, mag = fabsf(t);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float a = m[j * 4 + i] * bmin[j], c = m[j * 4 + i] * bmax[j];
            l += fminf(a, c);
            h += fmaxf(a, c);
            mag += fmaxf(fabsf(a), fabsf(c));
        }
        const float margin = 1e-4f * mag;
        lo[i] = l - margin;
        hi[i] = h + margin;
    }
    float dmin, dmax;
    if (p.pointSizeMode == 1) {
        dmin = -hi[2];
        dmax = -lo[2];
    } else {
        float n2 = 0.0f, f2 = 0.0f;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float near = fmaxf(0.0f, fmaxf(lo[i], -hi[i]));
            const float far = fmaxf(fabsf(lo[i]), fabsf(hi[i]));
            n2 += near * near;
            f2 += far * far;
        }
        dmin = sqrtf(n2);
        dmax = sqrtf(f2);
    }
    dmin = fmaxf(dmin * (1.0f - 1e-4f), 0.000001f);
    dmax = fmaxf(dmax * (1.0f + 1e-4f), 0.000001f);
    if (!__builtin_isfinite(dmin) || !__builtin_isfinite(dmax)) return -1;
    const int rNear = radius_of_size(k * rcp(dmin), p.lo, p.hi);
    const int rFar = radius_of_size(k * rcp(dmax), p.lo, p.hi);
    return rNear == rFar ? rNear : -1;
}

// Host-side radius for lo == hi (size clamps to lo for every point).
__host__ inline int const_radius_for(float lo) { return radius_of_size(lo, lo, lo); }

// ---------------------------------------------------------------- batch bounds
//
// What the cull pass proves about every point of a visible batch from the
// eight corners of its box, for the draw passes to rely on.  A point's
// quantised coordinates (q converted to float: q_f <= 2^bits) stand for the
// position P = min + q_f * s, s = size * 2^-bits exactly, which lies in the
// box [min, min + size] (evaluated exactly, in double).  The transform is
// affine, so the exact clip coordinates of P lie between the corner values,
// and since w > 0 on the whole box (all corners), so do x / w, y / w and
// z / w.  The float clip coordinates the passes compute (project_q: one
// rounding per A entry, four in B = mul_point(T, min), three in the fma
// chain) are within 7 * 2^-24 A_r of the exact ones, with
// A_r = sum_c |T[c][r]| (|min_c| + size_c) + |T[3][r]|, which bounds every
// partial sum (for displaced points nothing is proven); so with the
// allowance e = 2^-20 A the float ratio n = x_f * rcp(w_f) satisfies
//     |n - x / w| <= (e_x + N e_w) / (w_min - e_w) = D,  plus |n| 2^-22.9
// for the two roundings of rcp and the product (N = the largest |x / w| of
// the corners).  The viewport (fma(n, 0.5, 0.5) * W) adds at most 2^-23 W.
// Corner values are evaluated in double: a product of a float and a corner
// coordinate (a sum of two floats) is exact to 2^-52 relative, and the sums'
// rounding (a few 2^-52 A) is far inside the allowance.
struct BatchBounds {
    int32_t x0, x1, y0, y1;  // kProven: every fragment centre's column ix and unflipped row iy lie in [x0, x1] x [y0, y1]
    float zmax;              // kProven: every fragment's ndc z is <= zmax
    uint32_t flags;          // kBoundsProven | kBoundsInside | kBoundsFastRcp
    int32_t rconst;          // footprint radius every point of the batch computes, or -1
    uint32_t pad;
};
constexpr uint32_t kBoundsProven = 1u;   // the rectangle and zmax hold
constexpr uint32_t kBoundsInside = 2u;   // no point fails the w, NDC or viewport tests (nearest's ndc z > 0 included)
constexpr uint32_t kBoundsFastRcp = 4u;  // every point's clip w lies in [2^-125, 2^125] (rcp_fast is exact)

// What the cull pass leaves for the draws of one visible batch: the bounds
// its box proves and everything the draws would otherwise fetch through the
// batch record and its file (a chain of dependent loads per batch): the
// first point, the file, the batch minimum, the decode step at the drawn
// precision and the batch projection.  Eight 16-byte words.
struct DrawRecord {
    BatchBounds bb;
    uint32_t first, fileIndex;
    float mn[3], step[3];
    BatchAffine aff;
};
static_assert(sizeof(DrawRecord) == 128, "eight 16-byte words");

// One corner through m in double: clip value per row (the magnitude bound A
// per row comes from the whole box, batch_magnitudes).
struct CornerTerms {
    double c[4], a[4];
};
__host__ __device__ inline CornerTerms corner_terms(const float* m, double x, double y, double z, const double* a) {
    CornerTerms t;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double tx = (double)m[i] * x, ty = (double)m[4 + i] * y, tz = (double)m[8 + i] * z;
        t.c[i] = ((tx + ty) + tz) + (double)m[12 + i];
        t.a[i] = a[i];
    }
    return t;
}

// Running extremes over the corners.
struct CornerExtremes {
    double wmin, wmax, nmin[3], nmax[3], amax[4];
    int bad;  // a corner behind the camera or not finite
};
__host__ __device__ inline CornerExtremes corner_extremes(const CornerTerms& t) {
    CornerExtremes e;
    const double w = t.c[3];
    e.bad = (w > 0.0 && w < 1e300) ? 0 : 1;
    e.wmin = e.wmax = w;
#pragma unroll
    for (int i = 0; i < 3; i++) e.nmin[i] = e.nmax[i] = e.bad ? 0.0 : t.c[i] / w;
#pragma unroll
    for (int i = 0; i < 4; i++) e.amax[i] = t.a[i];
    for (int i = 0; i < 4; i++) e.bad |= (t.c[i] == t.c[i] && fabs(t.c[i]) < 1e300 && t.a[i] < 1e300) ? 0 : 1;
    return e;
}
__host__ __device__ inline CornerExtremes combine(const CornerExtremes& a, const CornerExtremes& b) {
    CornerExtremes e;
    e.bad = a.bad | b.bad;
    e.wmin = fmin(a.wmin, b.wmin);
    e.wmax = fmax(a.wmax, b.wmax);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        e.nmin[i] = fmin(a.nmin[i], b.nmin[i]);
        e.nmax[i] = fmax(a.nmax[i], b.nmax[i]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) e.amax[i] = fmax(a.amax[i], b.amax[i]);
    return e;
}

// The box of the batch's positions, exactly (a sum of two floats in double):
// min and min + size, the corner q = 2^bits reaches (size * 2^-bits * 2^bits
// = size); and the magnitude bound A_r = sum_c |m[c][r]| (|min_c| + size_c) +
// |m[3][r]| of every partial sum the projection forms inside it.
__host__ __device__ inline void batch_box(const scr_raster_batch& b, double lo[3], double hi[3]) {
    const float mn[3] = {b.minX, b.minY, b.minZ};
    const float size[3] = {fmaxf(b.maxX - b.minX, 0.000001f), fmaxf(b.maxY - b.minY, 0.000001f),
                           fmaxf(b.maxZ - b.minZ, 0.000001f)};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        lo[c] = (double)mn[c];
        hi[c] = (double)mn[c] + (double)size[c];
    }
}
__host__ __device__ inline void batch_magnitudes(const float* m, const double lo[3], const double hi[3], double a[4]) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
        double s = fabs((double)m[12 + r]);
#pragma unroll
        for (int c = 0; c < 3; c++) s += fabs((double)m[c * 4 + r]) * (fabs(lo[c]) + (hi[c] - lo[c]));
        // the factor covers the rounding of this double sum; the constant
        // (2^-80, an allowance of 2^-100) covers products of A entries that
        // leave the normal float range
        a[r] = s * (1.0 + 0x1p-40) + 0x1p-80;
    }
}

// Bounds from the corners' extremes, for a width x height viewport.
__host__ __device__ inline BatchBounds finish_bounds(const CornerExtremes& e, int width, int height, int rconst) {
    BatchBounds bb;
    bb.x0 = bb.y0 = 0;
    bb.x1 = bb.y1 = -1;
    bb.zmax = __builtin_inff();
    bb.flags = 0u;
    bb.rconst = rconst;
    bb.pad = 0u;
    if (e.bad) return bb;
    const double ew = e.amax[3] * 0x1p-20;
    if (e.wmin - ew >= 0x1p-125 && e.wmax + ew <= 0x1p125) bb.flags |= kBoundsFastRcp;
    const double den = e.wmin - ew;
    if (!(den > 0.0) || width > (1 << 20) || height > (1 << 20)) return bb;
    double d[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double n = fmax(fabs(e.nmin[i]), fabs(e.nmax[i]));
        const double di = (e.amax[i] * 0x1p-20 + n * ew) / den;
        d[i] = di + (n + di) * 0x1p-22;
        if (!(d[i] < 1e30)) return bb;
    }
    // fragment columns: trunc(fl(fl(n / 2 + 1 / 2) * W)) of n in [-1, 1]
    const double W = (double)width, H = (double)height;
    const double sx = W * 0x1p-21 + 0x1p-8, sy = H * 0x1p-21 + 0x1p-8;
    const double lim = 0x1p30;
    const double lx = fmin(fmax((0.5 * (e.nmin[0] - d[0]) + 0.5) * W - sx, -lim), lim);
    const double hx = fmin(fmax((0.5 * (e.nmax[0] + d[0]) + 0.5) * W + sx, -lim), lim);
    const double ly = fmin(fmax((0.5 * (e.nmin[1] - d[1]) + 0.5) * H - sy, -lim), lim);
    const double hy = fmin(fmax((0.5 * (e.nmax[1] + d[1]) + 0.5) * H + sy, -lim), lim);
    bb.x0 = (int32_t)floor(lx);
    bb.x1 = (int32_t)floor(hx);
    bb.y0 = (int32_t)floor(ly);
    bb.y1 = (int32_t)floor(hy);
    const double z = e.nmax[2] + d[2];
    bb.zmax = (float)(z + fabs(z) * 0x1p-20 + 0x1p-60);  // rounded to a float above z
    bb.flags |= kBoundsProven;
    // every point passes: ndc x, y in [-1, 1 - 2^-20] (so the viewport
    // column stays below W), ndc z in (0, 1]
    const double top = 1.0 - 0x1p-20;
    if (e.nmin[0] - d[0] >= -1.0 && e.nmax[0] + d[0] <= top && e.nmin[1] - d[1] >= -1.0 && e.nmax[1] + d[1] <= top &&
        e.nmin[2] - d[2] > 0.0 && z <= 1.0)
        bb.flags |= kBoundsInside;
    return bb;
}

// Bounds of a batch under transform m (file.transform), all corners in one
// thread (the cull pass).
__host__ __device__ inline BatchBounds batch_bounds(const float* m, const scr_raster_batch& b, int width, int height,
                                                    int rconst) {
    double lo[3], hi[3], a[4];
    batch_box(b, lo, hi);
    batch_magnitudes(m, lo, hi, a);
    CornerExtremes e = corner_extremes(corner_terms(m, lo[0], lo[1], lo[2], a));
    for (int k = 1; k < 8; k++)
        e = combine(e, corner_extremes(corner_terms(m, (k & 1) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1],
                                                    (k & 4) ? hi[2] : lo[2], a)));
    return finish_bounds(e, width, height, rconst);
}

// insidePointFootprint — Common.metal:63-71
__host__ __device__ inline bool inside_footprint(int ox, int oy, int r) {
    if (r <= 0) return ox == 0 && oy == 0;
    const float fx = (float)ox, fy = (float)oy;
    const float rr = (float)r + 0.5f;
    return (fx * fx + fy * fy) <= rr * rr;
}

}  // namespace scr
