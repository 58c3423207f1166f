/*
 * scr_oracle.c — CPU oracle for the compute rasteriser (test infrastructure;
 * the arithmetic contract is stated in scr_oracle.h).
 *
 * Each function cites the reference file:line it restates (paths under
 * Sources/SatinComputeRasteriser/).  Build with -O2 -ffp-contract=off.
 */
#include "scr_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_TWO32 4294967296.0f
#define ORC_INV_TWO32 2.3283064365386963e-10f /* 2^-32, exact */
#define ORC_INV255 (1.0f / 255.0f)

/* Planted mistakes, compiled in only by the mutation test
 * (tests/test_oracle_mutants.py builds this file with -DORC_MUTANT=k and
 * requires a pin of tests/test_oracle_pinning.py to fail for every k).  The
 * default, 0, is the oracle itself. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* ------------------------------------------------------------------ C4 */
uint32_t orc_f2u(float x) {
    if (!(x > 0.0f)) return 0u;                 /* NaN, <= 0 */
    if (x >= ORC_TWO32) return 0xffffffffu;
    return (uint32_t)x;
}

int32_t orc_f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)x;
}

static float orc_saturate(float x) {
    /* clamp(x, 0, 1) as fmin(fmax(x, 0), 1): NaN -> 0 */
    float y = fmaxf(x, 0.0f);
    return fminf(y, 1.0f);
}

static float orc_rcp(float x) { return 1.0f / x; } /* C2: correctly rounded */

/* ------------------------------------------------------------------ C3
 * log2 for the cull pass (Common.metal:185-187 calls the Metal log2).
 * x = 2^e * m, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1),
 * atanh series to s^21; result = e + ln(m) * log2(e).  Only correctly rounded
 * double +,-,*,/ are used, so any IEEE machine gives the same bits. */
/* Contract C8: e^x for the coverage resolve, in double with + - * / and
 * exact operations only (rint, ldexp), rounded to binary32 once: k = rint(x /
 * ln2) taken as x * (1/ln2); r = (x - k ln2_hi) - k ln2_lo with fdlibm's
 * split of ln2; e^r as 1 + (r/1)(1 + (r/2)(1 + ... (1 + r/13))) evaluated
 * from n = 13 down; times 2^k. */
float orc_expf(float xf) {
    if (xf != xf) return xf;
    if (xf > 100.0f) return INFINITY;
    if (xf < -110.0f) return 0.0f;
    const double x = (double)xf;
    const double k = rint(x * 1.4426950408889634);
    const double r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    double e = 1.0;
    for (int n = 13; n >= 1; n--) e = 1.0 + (r / (double)n) * e;
    return (float)ldexp(e, (int)k);
}

float orc_log2f(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    if (x != x) return x;
    if (x < 0.0f) return (0.0f / 0.0f);
    if (x == 0.0f) return -INFINITY;
    uint32_t expo = (v.u >> 23) & 0xffu;
    if (expo == 0xffu) return x; /* +inf */
    int e;
    double m;
    if (expo == 0u) { /* subnormal: scale by 2^24 */
        union { float f; uint32_t u; } w;
        w.f = x * 16777216.0f;
        expo = (w.u >> 23) & 0xffu;
        e = (int)expo - 127 - 24;
        union { float f; uint32_t u; } mm;
        mm.u = (w.u & 0x007fffffu) | 0x3f800000u;
        m = (double)mm.f;
    } else {
        e = (int)expo - 127;
        union { float f; uint32_t u; } mm;
        mm.u = (v.u & 0x007fffffu) | 0x3f800000u; /* [1, 2) */
        m = (double)mm.f;
    }
    if (m > 1.4142135623730951) { m = m * 0.5; e = e + 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
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
    double lnm = 2.0 * s * p;
    double r = (double)e + lnm * 1.4426950408889634;
    return (float)r;
}

/* hashUnit — Common.metal:190-197 (lowbias32 finaliser) */
float orc_hash_unit(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return (float)x * ORC_INV_TWO32;
}

/* depthToUintReverseZ / uintToDepthReverseZ — Common.metal:10-16.
 * float(CR_MAX_DEPTH - 1u) rounds to 2^32; z == 1 saturates (C4). */
uint32_t orc_depth_to_uint(float ndcZ) { return orc_f2u(orc_saturate(ndcZ) * ORC_TWO32); }
float orc_uint_to_depth(uint32_t d) { return (float)d * ORC_INV_TWO32; }

/* ------------------------------------------------------------------ C1 */
void orc_mat4_vec(const float* m, const float* v, float* out) {
    float r[4];
    for (int i = 0; i < 4; i++) {
        float a = m[0 * 4 + i] * v[0];
        a = a + m[1 * 4 + i] * v[1];
        a = a + m[2 * 4 + i] * v[2];
        a = a + m[3 * 4 + i] * v[3];
        r[i] = a;
    }
    memcpy(out, r, sizeof r);
}

/* The per-point transform of the render passes (C1): a point (w = 1)
 * through m, column sums in source order with every product-then-add fused,
 * fma(c2, z, fma(c1, y, c0 x)) + c3. */
void orc_transform_point(const float* m, const float p[3], float out[4]) {
    float r[4];
    for (int i = 0; i < 4; i++) r[i] = fmaf(m[2 * 4 + i], p[2], fmaf(m[1 * 4 + i], p[1], m[0 * 4 + i] * p[0])) + m[3 * 4 + i];
    memcpy(out, r, sizeof r);
}

void orc_mat4_mul(const float* a, const float* b, float* out) {
    float r[16];
    for (int c = 0; c < 4; c++) orc_mat4_vec(a, b + c * 4, r + c * 4);
    memcpy(out, r, sizeof r);
}

static float orc_len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }

/* ------------------------------------------------------------------ Morton
 * mortonSpread10 — PackedPointCloudFixtures.swift:337-344; the 21-bit spread
 * is the same bit interleave widened to 63-bit keys. */
static uint32_t orc_spread10(uint32_t v) {
    uint32_t x = v & 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

static uint64_t orc_spread21(uint32_t v) {
    /* plain bit loop: bit i of v goes to bit 3i */
    uint64_t x = 0;
    for (int i = 0; i < 21; i++) x |= (uint64_t)((v >> i) & 1u) << (3 * i);
    return x;
}

uint32_t orc_morton30(uint32_t qx, uint32_t qy, uint32_t qz) {
    return (orc_spread10(qx) << 2) | (orc_spread10(qy) << 1) | orc_spread10(qz);
}

uint64_t orc_morton63(uint32_t qx, uint32_t qy, uint32_t qz) {
    return (orc_spread21(qx) << 2) | (orc_spread21(qy) << 1) | orc_spread21(qz);
}

/* ------------------------------------------------------------------ Common.metal */

/* pointFootprintRadius — Common.metal:18-61 */
int orc_footprint_radius(const float p[3], const scr_raster_file* file, const float* view,
                         const float* proj, int width, int height, const scr_config* cfg) {
    (void)width;
    const float mn = cfg->minimumPointSize, mx = cfg->maximumPointSize;
    const float lo = fmaxf(fminf(mn, mx), 1.0f);
    const float hi = fmaxf(fmaxf(mn, mx), lo);
    float wp[4], vp[4];
    orc_transform_point(file->world, p, wp);
    orc_transform_point(view, wp, vp);
    const int ortho = (proj[3 * 4 + 3] == 1.0f);
    float ps;
    if (cfg->pointSizeMode == 1) {
        const float focal = ((float)height * 0.5f) * proj[1 * 4 + 1];
        if (ortho) {
            ps = (2.0f * cfg->pointSizeScale) * focal;
        } else {
#if ORC_MUTANT == 10
            const float viewZ = fmaxf(orc_len3(vp[0], vp[1], vp[2]), 0.000001f);
#else
            const float viewZ = fmaxf(-vp[2], 0.000001f);
#endif
            ps = ((2.0f * cfg->pointSizeScale) * focal) * orc_rcp(viewZ);
        }
    } else {
        if (ortho) {
            ps = cfg->pointSizeScale;
        } else {
            const float d = fmaxf(orc_len3(vp[0], vp[1], vp[2]), 0.000001f);
            ps = cfg->pointSizeScale * orc_rcp(d);
        }
    }
    if (!isfinite(ps)) ps = lo;
    ps = fminf(fmaxf(ps, lo), hi);
#if ORC_MUTANT == 2
    int r = orc_f2i(floorf(fmaxf(ps - 1.0f, 0.0f) * 0.5f));
#else
    int r = orc_f2i(ceilf(fmaxf(ps - 1.0f, 0.0f) * 0.5f));
#endif
    if (r < 0) r = 0;
    if (r > 16) r = 16;
    return r;
}

/* insidePointFootprint — Common.metal:63-71 */
int orc_inside_footprint(int ox, int oy, int radius) {
    if (radius <= 0) return ox == 0 && oy == 0;
    const float x = (float)ox, y = (float)oy;
    const float r = (float)radius + 0.5f;
    return (x * x + y * y) <= r * r;
}

/* The quantised coordinates of point pi at the batch precision (Common.metal:
 * 73-117: 30, 20 or 10 bits from three, two or one 10-bit words) and the
 * decode step per axis, max(size, 1e-6) * 2^-bits (a power-of-two scaling:
 * exact). */
static void orc_point_q(uint32_t pi, const scr_raster_batch* b, const uint32_t* low, const uint32_t* med,
                        const uint32_t* high, int level, uint32_t q[3], float step[3]) {
    const float mn[3] = {b->minX, b->minY, b->minZ};
    const float mx[3] = {b->maxX, b->maxY, b->maxZ};
    float size[3];
    for (int i = 0; i < 3; i++) size[i] = fmaxf(mx[i] - mn[i], 0.000001f);
    float steps;
    const uint32_t lo = low[pi];
    if (level == 0) {
        const uint32_t me = med[pi], hi = high[pi];
        for (int a = 0; a < 3; a++) {
            const uint32_t s = 10u * (uint32_t)a;
            q[a] = (((lo >> s) & 1023u) << 20) | (((me >> s) & 1023u) << 10) | ((hi >> s) & 1023u);
        }
        steps = 1073741824.0f;
    } else if (level == 1) {
        const uint32_t me = med[pi];
        for (int a = 0; a < 3; a++) {
            const uint32_t s = 10u * (uint32_t)a;
            q[a] = (((lo >> s) & 1023u) << 10) | ((me >> s) & 1023u);
        }
        steps = 1048576.0f;
    } else {
        for (int a = 0; a < 3; a++) q[a] = (lo >> (10u * (uint32_t)a)) & 1023u;
        steps = 1024.0f;
    }
    const float invSteps = 1.0f / steps; /* power of two: exact */
    for (int a = 0; a < 3; a++) step[a] = size[a] * invSteps;
}

/* decodePoint — Common.metal:73-117 (fused: q * s + min as one fmaf, the
 * render passes' contract C1) and NearestPointCPUReference.swift:87-117
 * (fused = 0: Swift rounds the product and the sum separately). */
void orc_decode_point(uint32_t pi, const scr_raster_batch* b, const uint32_t* low,
                      const uint32_t* med, const uint32_t* high, int level, int fused, float out[3]) {
    const float mn[3] = {b->minX, b->minY, b->minZ};
    uint32_t q[3];
    float step[3];
    orc_point_q(pi, b, low, med, high, level, q, step);
    for (int a = 0; a < 3; a++)
        out[a] = fused ? fmaf((float)q[a], step[a], mn[a]) : (float)q[a] * step[a] + mn[a];
}

/* The render passes' projection of an undisplaced point (contract C1):
 * decodePoint (Common.metal:73-117) and the file transform (DepthPass/
 * Shaders.metal:75-78) folded into one affine map of the quantised
 * coordinates per batch, A[c][r] = T[c][r] * step_c (one rounding each),
 * B = T (min, 1) (orc_transform_point), clip_r = fma(A[2][r], qz,
 * fma(A[1][r], qy, fma(A[0][r], qx, B_r))). */
void orc_project_q(const float* T, const float mn[3], const float step[3], const float qf[3], float out[4]) {
    float B[4];
    orc_transform_point(T, mn, B);
#if ORC_MUTANT == 19
    for (int r = 0; r < 4; r++) B[r] -= T[12 + r];
#endif
    for (int r = 0; r < 4; r++) {
        float A[3];
        for (int c = 0; c < 3; c++) A[c] = T[c * 4 + r] * step[ORC_MUTANT == 20 && c == 1 ? 0 : c];
        out[r] = fmaf(A[2], qf[2], fmaf(A[1], qf[1], fmaf(A[0], qf[0], B[r])));
    }
}

/* makePlane / intersectsFrustum — Common.metal:119-153 */
static void orc_make_plane(float x, float y, float z, float w, float n[3], float* off) {
    const float len = fmaxf(orc_len3(x, y, z), 0.000001f);
    const float r = orc_rcp(len);
    n[0] = x * r;
    n[1] = y * r;
    n[2] = z * r;
    *off = w * r;
}

int orc_intersects_frustum(const float* m, const float mn[3], const float mx[3]) {
#define M(c, r) m[(c) * 4 + (r)]
    const float pl[6][4] = {
        {M(0, 3) - M(0, ORC_MUTANT == 1), M(1, 3) - M(1, ORC_MUTANT == 1), M(2, 3) - M(2, ORC_MUTANT == 1),
         M(3, 3) - M(3, ORC_MUTANT == 1)},
        {M(0, 3) + M(0, 0), M(1, 3) + M(1, 0), M(2, 3) + M(2, 0), M(3, 3) + M(3, 0)},
        {M(0, 3) + M(0, 1), M(1, 3) + M(1, 1), M(2, 3) + M(2, 1), M(3, 3) + M(3, 1)},
        {M(0, 3) - M(0, 1), M(1, 3) - M(1, 1), M(2, 3) - M(2, 1), M(3, 3) - M(3, 1)},
        {M(0, 3) - M(0, 2), M(1, 3) - M(1, 2), M(2, 3) - M(2, 2), M(3, 3) - M(3, 2)},
#if ORC_MUTANT == 11
        {M(0, 3) + M(0, 2), M(1, 3) + M(1, 2), M(2, 3) + M(2, 2), M(3, 3) + M(3, 2)},
#else
        {M(0, 2), M(1, 2), M(2, 2), M(3, 2)},
#endif
    };
#undef M
    for (int i = 0; i < 6; i++) {
        float n[3], off;
        orc_make_plane(pl[i][0], pl[i][1], pl[i][2], pl[i][3], n, &off);
        const float px = n[0] > 0.0f ? mx[0] : mn[0];
        const float py = n[1] > 0.0f ? mx[1] : mn[1];
        const float pz = n[2] > 0.0f ? mx[2] : mn[2];
        const float d = (n[0] * px + n[1] * py) + n[2] * pz;
        if (d + off < 0.0f) return 0;
    }
    return 1;
}

/* pixelSizeOnScreen — Common.metal:155-172 */
float orc_pixel_size_on_screen(const scr_raster_batch* b, const scr_raster_file* f,
                               const float* view, const float* proj, int width, int height) {
    const float mn[3] = {b->minX, b->minY, b->minZ};
    const float mx[3] = {b->maxX, b->maxY, b->maxZ};
    float c[4];
    for (int i = 0; i < 3; i++) c[i] = (mn[i] + mx[i]) * 0.5f;
    c[3] = 1.0f;
    const float radius = orc_len3(mn[0] - mx[0], mn[1] - mx[1], mn[2] - mx[2]);
    float vw[16], vc[4], ve[4], pc[4], pe[4];
    orc_mat4_mul(view, f->world, vw);
    orc_mat4_vec(vw, c, vc);
    ve[0] = vc[0] + radius;
    ve[1] = vc[1] + 0.0f;
    ve[2] = vc[2] + 0.0f;
    ve[3] = vc[3] + 0.0f;
    orc_mat4_vec(proj, vc, pc);
    orc_mat4_vec(proj, ve, pe);
    const float rc = orc_rcp(fmaxf(fabsf(pc[3]), 0.000001f));
    const float re = orc_rcp(fmaxf(fabsf(pe[3]), 0.000001f));
    pc[0] = pc[0] * rc;
    pc[1] = pc[1] * rc;
    pe[0] = pe[0] * re;
    pe[1] = pe[1] * re;
    const float sz[2] = {(float)(ORC_MUTANT == 4 ? height : width), (float)height};
    float sc[2], se[2];
    for (int i = 0; i < 2; i++) {
        sc[i] = (sz[i] * (pc[i] + 1.0f)) * 0.5f;
        se[i] = (sz[i] * (pe[i] + 1.0f)) * 0.5f;
    }
    const float dx = se[0] - sc[0], dy = se[1] - sc[1];
    return sqrtf(dx * dx + dy * dy);
}

/* precisionLevel — Common.metal:174-181 / Cull/Shaders.metal:47-53 */
int orc_precision_level(float px) {
    if (px < (ORC_MUTANT == 8 ? 200.0f : 100.0f)) return 4;
    if (px < (ORC_MUTANT == 8 ? 100.0f : 200.0f)) return 3;
    if (px < 500.0f) return 2;
    if (px < 10000.0f) return 1;
    return 0;
}

/* lodThresholdFromPixelSize — Common.metal:185-187 */
float orc_lod_threshold(float px, int lodBias) {
    return (orc_log2f(fmaxf(px, 1.0f) * (1.0f / 25.0f)) + (ORC_MUTANT == 5 ? 0.0f : 3.0f)) + (float)lodBias;
}

/* ------------------------------------------------------------------ host composition */

/* updateFiles — ComputeRasteriserPointCloud.swift:487-515 */
void orc_update_files(const scr_raster_file* in, int n, const float* vp, const float* model,
                      const float* frustum, const float* prevVP, scr_raster_file* out) {
    for (int i = 0; i < n; i++) {
        scr_raster_file f;
        float world[16];
        orc_mat4_mul(model, in[i].world, world);
        orc_mat4_mul(vp, world, f.transform);
        orc_mat4_mul(frustum ? frustum : vp, world, f.transformFrustum);
        memcpy(f.world, world, sizeof world);
        orc_mat4_mul(prevVP ? prevVP : vp, world, f.prevTransform);
        out[i] = f;
    }
}

/* ------------------------------------------------------------------ cull
 * cullUpdate — Cull/Shaders.metal:19-83.  The reference appends with an atomic
 * counter (order unspecified); the oracle lists survivors in slot order. */
static int orc_cull_one(const orc_cloud* c, int gid, const scr_config* cfg, const float* view,
                        const float* proj, int width, int height, scr_visible_batch* vb) {
    const scr_raster_batch* b = &c->batches[gid];
    if (b->state == 0) return 0;
    if ((int)b->fileIndex >= c->nFiles) return 0;
    const scr_raster_file* f = &c->files[b->fileIndex];
    if (cfg->enableFrustumCulling != 0) {
        const float mn[3] = {b->minX, b->minY, b->minZ};
        const float mx[3] = {b->maxX, b->maxY, b->maxZ};
        if (!orc_intersects_frustum(f->transformFrustum, mn, mx)) return 0;
    }
    const float px = orc_pixel_size_on_screen(b, f, view, proj, width, height);
    const int level = orc_precision_level(px);
    const float thr = (cfg->enableCLOD != 0) ? fmaxf(0.0f, orc_lod_threshold(px, cfg->lodBias)) : 99.0f;
    uint32_t active = b->numPoints;
    if (b->lodCum67 != 0u) {
        int lmax = orc_f2i(floorf(thr + 0.5f));
        if (lmax < 0) lmax = 0;
        if (lmax > 7) lmax = 7;
        const uint32_t words[4] = {b->lodCum01, b->lodCum23, b->lodCum45, b->lodCum67};
        const uint32_t w = words[lmax >> 1];
        active = (lmax & 1) ? (w >> 16) : (w & 0xffffu);
    }
    if (active == 0u) return 0;
    vb->batchIndex = (uint32_t)gid;
    vb->level = level;
    vb->lodThreshold = thr;
    vb->activePoints = active;
    return 1;
}

int orc_cull(const orc_cloud* c, const scr_config* cfg, const float* view, const float* proj,
             int width, int height, scr_visible_batch* visibleOut) {
    int n = 0;
    for (int gid = 0; gid < c->nSlots; gid++) {
        scr_visible_batch vb;
        if (orc_cull_one(c, gid, cfg, view, proj, width, height, &vb)) visibleOut[n++] = vb;
    }
    return n;
}

/* ------------------------------------------------------------------ per-point front end
 * Shared prologue of depthPassUpdate (DepthPass/Shaders.metal:48-97),
 * colorPassUpdate (ColorPass/Shaders.metal:58-93), nearestDepthUpdate /
 * nearestIndexUpdate (NearestDepth :37-72, NearestIndex :37-72).
 * Returns 1 when the point survives; fills ndcZ, pixel (x, y-flipped), radius. */
typedef struct orc_frag {
    float base[3];   /* decoded position */
    float point[3];  /* decoded position + displacement */
    float ndcZ;
    int px, py;
    uint32_t depth;
    int radius;
} orc_frag;

static int orc_front(const orc_cloud* c, const scr_raster_batch* b, const scr_raster_file* f,
                     const scr_visible_batch* vb, uint32_t pointIndex, const scr_config* cfg,
                     const float* view, const float* proj, int width, int height,
                     int nearestZ, int wantRadius, orc_frag* fr) {
    const float pointLevel = (float)((uint32_t)c->levels[pointIndex] & 7u);
    const float dither = (cfg->enableLODDither != 0) ? orc_hash_unit(pointIndex) : 0.5f;
    if (dither >= (vb->lodThreshold - pointLevel) + 0.5f) return 0;
    orc_decode_point(pointIndex, b, c->xyzLow, c->xyzMed, c->xyzHigh, vb->level, 1, fr->base);
    for (int a = 0; a < 3; a++) fr->point[a] = fr->base[a];
    /* displacement (DepthPass/Shaders.metal:69-74, ColorPass/Shaders.metal:
     * 72-78; the nearest passes do not displace): point += displacement, a
     * cloud without a buffer reads zeros, a NaN component drops the point */
    if (cfg->applyDisplacement != 0 && !nearestZ) {
        static const float zero4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const float* d = c->disp ? c->disp + (size_t)pointIndex * 4 : zero4;
        for (int a = 0; a < 3; a++) fr->point[a] = fr->point[a] + d[a];
        if (ORC_MUTANT != 12 && (isnan(fr->point[0]) || isnan(fr->point[1]) || isnan(fr->point[2]))) return 0;
    }
    float clip[4];
    if (cfg->applyDisplacement != 0 && !nearestZ) {
        orc_transform_point(f->transform, fr->point, clip);
    } else {
        const float mn[3] = {b->minX, b->minY, b->minZ};
        uint32_t q[3];
        float step[3];
        orc_point_q(pointIndex, b, c->xyzLow, c->xyzMed, c->xyzHigh, vb->level, q, step);
        const float qf[3] = {(float)q[0], (float)q[1], (float)q[2]};
        orc_project_q(f->transform, mn, step, qf, clip);
    }
    if (clip[3] <= 0.0f) return 0;
    const float rw = orc_rcp(clip[3]);
    const float nx = clip[0] * rw, ny = clip[1] * rw, nz = clip[2] * rw;
    if (nx < -1.0f || ny < -1.0f || nx > 1.0f || ny > 1.0f) return 0;
    if (nearestZ) {
        if (nz <= 0.0f || nz > 1.0f) return 0;
    } else {
        if (nz < 0.0f || nz > 1.0f) return 0;
    }
    int px = orc_f2i(fmaf(nx, 0.5f, 0.5f) * (float)width);
    int py = orc_f2i(fmaf(ny, 0.5f, 0.5f) * (float)height);
    if (px < 0 || px >= width || py < 0 || py >= height) return 0;
    py = height - 1 - py;
    fr->ndcZ = nz;
    fr->px = px;
    fr->py = py;
    fr->depth = orc_depth_to_uint(nz);
    fr->radius = wantRadius ? orc_footprint_radius(fr->point, f, view, proj, width, height, cfg) : 0;
    return 1;
}

/* ------------------------------------------------------------------ HQS */
typedef struct orc_vis_item {
    int cloud;
    scr_visible_batch vb;
} orc_vis_item;

typedef struct orc_job {
    const orc_cloud* clouds;
    const orc_vis_item* items;
    int first, last;
    const scr_config* cfg;
    const float* view;
    const float* proj;
    int width, height;
    uint32_t* depth;          /* private (depth phase) or shared final (colour phase) */
    uint32_t *r, *g, *b, *n;  /* private colour planes */
    int phase;
} orc_job;

/* depthPassUpdate — DepthPass/Shaders.metal:21-116 */
static void orc_depth_items(const orc_job* j) {
    for (int it = j->first; it < j->last; it++) {
        const orc_cloud* c = &j->clouds[j->items[it].cloud];
        const scr_visible_batch* vb = &j->items[it].vb;
        const scr_raster_batch* b = &c->batches[vb->batchIndex];
        const scr_raster_file* f = &c->files[b->fileIndex];
        for (uint32_t li = 0; li < vb->activePoints; li++) {
            const uint32_t pi = b->firstPoint + li;
            /* translucent defocus: the discard sentinel claims no depth
             * (DepthPass/Shaders.metal:60-67) */
            if (j->cfg->applyTint != 0 && j->cfg->tintAlphaIsCoverage != 0 && c->tints &&
                c->tints[(size_t)pi * 4 + 3] < 0.0f)
                continue;
            orc_frag fr;
            if (!orc_front(c, b, f, vb, pi, j->cfg, j->view, j->proj, j->width, j->height, 0, 1, &fr))
                continue;
            for (int oy = -fr.radius; oy <= fr.radius; oy++)
                for (int ox = -fr.radius; ox <= fr.radius; ox++) {
                    if (!orc_inside_footprint(ox, oy, fr.radius)) continue;
                    const int tx = fr.px + ox, ty = fr.py + oy;
                    if (tx < 0 || tx >= j->width || ty < 0 || ty >= j->height) continue;
                    uint32_t* d = &j->depth[(size_t)ty * j->width + tx];
                    if (fr.depth > *d) *d = fr.depth;
                }
        }
    }
}

/* colorPassUpdate, weighted-blend path (ColorPass/Shaders.metal:133-209):
 * no depth test; the fragment adds colour * a * 4096 and a * 4096 to every
 * footprint pixel of each motion-blur sample, the splat swept from its pixel
 * back toward its previous frame's pixel (file.prevTransform).  Integer pixel
 * differences wrap as the shader's int2 arithmetic does. */
static void orc_blend_fragment(const orc_job* j, const orc_cloud* c, uint32_t pi, const scr_raster_file* f,
                               const orc_frag* fr, uint32_t r, uint32_t g, uint32_t bl, float a) {
    if (a <= 0.0f) return;
    const int W = j->width, H = j->height;
    int samples = 1;
    float stepX = 0.0f, stepY = 0.0f;
    if (j->cfg->motionBlur > 0.0f) {
        /* the sweep starts from base + the previous displacement
         * (ColorPass/Shaders.metal:150-153; the current displacement when
         * there is no previous one, zeros without a buffer) */
        float prevPoint[3] = {fr->base[0], fr->base[1], fr->base[2]};
        if (ORC_MUTANT == 13)
            for (int a = 0; a < 3; a++) prevPoint[a] = fr->point[a];
        else if (j->cfg->applyDisplacement != 0) {
            static const float zero4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            const float* d = c->prevDisp ? c->prevDisp + (size_t)pi * 4 : (c->disp ? c->disp + (size_t)pi * 4 : zero4);
            for (int a = 0; a < 3; a++) prevPoint[a] = prevPoint[a] + d[a];
        }
        float prev[4];
        orc_transform_point(f->prevTransform, prevPoint, prev);
        if (prev[3] > 0.0f) {
            const float prw = orc_rcp(prev[3]);
            const int ppx = orc_f2i(fmaf(prev[0] * prw, 0.5f, 0.5f) * (float)W);
            const int ppy = H - 1 - orc_f2i(fmaf(prev[1] * prw, 0.5f, 0.5f) * (float)H);
            float vx = (float)(int32_t)((uint32_t)fr->px - (uint32_t)ppx) * j->cfg->motionBlur;
            float vy = (float)(int32_t)((uint32_t)fr->py - (uint32_t)ppy) * j->cfg->motionBlur;
            float len = sqrtf(vx * vx + vy * vy);
            if (len > j->cfg->motionBlurMaxSpread) {
                const float k = j->cfg->motionBlurMaxSpread * orc_rcp(len);
                vx = vx * k;
                vy = vy * k;
                len = j->cfg->motionBlurMaxSpread;
            }
            if (len > 0.75f) {
                int n = orc_f2i(ceilf(len));
                if (n < 1) n = 1;
                if (n > j->cfg->motionBlurSamples) n = j->cfg->motionBlurSamples;
                samples = n;
                const float d = orc_rcp((float)(n - 1 > 1 ? n - 1 : 1));
                stepX = vx * d;
                stepY = vy * d;
            }
        }
    }
    const float aN = ORC_MUTANT == 6 ? a : a * orc_rcp((float)samples);
    const uint32_t cr = orc_f2u((((float)r * ORC_INV255) * aN) * 4096.0f + 0.5f);
    const uint32_t cg = orc_f2u((((float)g * ORC_INV255) * aN) * 4096.0f + 0.5f);
    const uint32_t cb = orc_f2u((((float)bl * ORC_INV255) * aN) * 4096.0f + 0.5f);
    const uint32_t ca = orc_f2u(aN * 4096.0f + 0.5f);
    for (int s = 0; s < samples; s++) {
        const int cx = ORC_MUTANT == 9 ? (int32_t)((uint32_t)fr->px + (uint32_t)orc_f2i(roundf(stepX * (float)s)))
                                        : (int32_t)((uint32_t)fr->px - (uint32_t)orc_f2i(roundf(stepX * (float)s)));
        const int cy = (int32_t)((uint32_t)fr->py - (uint32_t)orc_f2i(roundf(stepY * (float)s)));
        for (int oy = -fr->radius; oy <= fr->radius; oy++)
            for (int ox = -fr->radius; ox <= fr->radius; ox++) {
                if (!orc_inside_footprint(ox, oy, fr->radius)) continue;
                const int tx = (int32_t)((uint32_t)cx + (uint32_t)ox), ty = (int32_t)((uint32_t)cy + (uint32_t)oy);
                if (tx < 0 || tx >= W || ty < 0 || ty >= H) continue;
                const size_t p = (size_t)ty * W + tx;
                j->r[p] += cr;
                j->g[p] += cg;
                j->b[p] += cb;
                j->n[p] += ca;
            }
    }
}

/* colorPassUpdate (opaque path) — ColorPass/Shaders.metal:30-219 */
static void orc_color_items(const orc_job* j) {
    const float keep = 1.0f - j->cfg->depthTolerance;
    for (int it = j->first; it < j->last; it++) {
        const orc_cloud* c = &j->clouds[j->items[it].cloud];
        const scr_visible_batch* vb = &j->items[it].vb;
        const scr_raster_batch* b = &c->batches[vb->batchIndex];
        const scr_raster_file* f = &c->files[b->fileIndex];
        for (uint32_t li = 0; li < vb->activePoints; li++) {
            const uint32_t pi = b->firstPoint + li;
            orc_frag fr;
            /* the colour pass computes the radius after the tint stage; it
             * depends only on the point, so computing it up front is the same */
            if (!orc_front(c, b, f, vb, pi, j->cfg, j->view, j->proj, j->width, j->height, 0, 1, &fr))
                continue;
            uint32_t color = c->colors[pi];
            if (j->cfg->colorizeChunks != 0) color = vb->batchIndex * 1234567u;
            else if (j->cfg->colorizeOverdraw != 0) color = 0x00010101u;
            uint32_t r = color & 0xffu, g = (color >> 8) & 0xffu, bl = (color >> 16) & 0xffu;
            const int coverageMode = j->cfg->applyTint != 0 && j->cfg->tintAlphaIsCoverage != 0;
            const int blend = coverageMode || j->cfg->motionBlur > 0.0f;
            float cover = 1.0f;
            if (j->cfg->applyTint != 0) {
                /* clouds without a tint buffer read a zeroed stand-in
                 * (ComputeRasteriser.swift:333-342) */
                static const float zero4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                const float* t = c->tints ? c->tints + (size_t)pi * 4 : zero4;
                if (t[3] < 0.0f) continue;
                if (coverageMode) {
                    cover = orc_saturate(1.0f - t[3]);
                    goto blended;
                }
                const float w = orc_saturate(t[3]);
                const float o[3] = {(float)r * ORC_INV255, (float)g * ORC_INV255, (float)bl * ORC_INV255};
                float m[3];
                for (int k = 0; k < 3; k++) m[k] = o[k] + (orc_saturate(t[k]) - o[k]) * w;
                r = orc_f2u(orc_saturate(m[0]) * 255.0f + 0.5f);
                g = orc_f2u(orc_saturate(m[1]) * 255.0f + 0.5f);
                bl = orc_f2u(orc_saturate(m[2]) * 255.0f + 0.5f);
            }
        blended:
            if (blend) {
                orc_blend_fragment(j, c, pi, f, &fr, r, g, bl, cover);
                continue;
            }
            for (int oy = -fr.radius; oy <= fr.radius; oy++)
                for (int ox = -fr.radius; ox <= fr.radius; ox++) {
                    if (!orc_inside_footprint(ox, oy, fr.radius)) continue;
                    const int tx = fr.px + ox, ty = fr.py + oy;
                    if (tx < 0 || tx >= j->width || ty < 0 || ty >= j->height) continue;
                    const size_t p = (size_t)ty * j->width + tx;
                    const uint32_t d = j->depth[p];
                    if (d == 0u) continue;
                    const float closest = orc_uint_to_depth(d);
                    if (!(fr.ndcZ >= closest * keep)) continue;
                    j->r[p] += r;
                    j->g[p] += g;
                    j->b[p] += bl;
                    j->n[p] += 1u;
                }
        }
    }
}

static void* orc_job_main(void* arg) {
    orc_job* j = (orc_job*)arg;
    if (j->phase == 0) orc_depth_items(j);
    else orc_color_items(j);
    return NULL;
}

static uint8_t orc_unorm8(float x) { return (uint8_t)orc_f2u(rintf(orc_saturate(x) * 255.0f)); }

static uint32_t orc_pack_rgba(float r, float g, float b, float a) {
    return (uint32_t)orc_unorm8(r) | ((uint32_t)orc_unorm8(g) << 8) | ((uint32_t)orc_unorm8(b) << 16) |
           ((uint32_t)orc_unorm8(a) << 24);
}

/* Wall-clock seconds of the phases of the process's last orc_render_hqs:
 * clear, cull, depth, colour, resolve (for the CPU baseline's per-pass split). */
static double orc_hqs_seconds[5];

static double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void orc_last_hqs_pass_seconds(double out[5]) { memcpy(out, orc_hqs_seconds, sizeof orc_hqs_seconds); }

int orc_render_hqs(const orc_cloud* clouds, int nClouds, const scr_config* cfg, const float* view,
                   const float* proj, int width, int height, orc_hqs_planes* out, int nThreads) {
    const size_t npx = (size_t)width * (size_t)height;
    double tick = orc_now();
    memset(orc_hqs_seconds, 0, sizeof orc_hqs_seconds);
#define ORC_LAP(k) do { const double t_ = orc_now(); orc_hqs_seconds[k] += t_ - tick; tick = t_; } while (0)
    /* clearUpdate — Clear/Shaders.metal:7-23 */
    memset(out->depth, 0, npx * 4);
    memset(out->red, 0, npx * 4);
    memset(out->green, 0, npx * 4);
    memset(out->blue, 0, npx * 4);
    memset(out->count, 0, npx * 4);
    ORC_LAP(0);

    int total = 0;
    for (int c = 0; c < nClouds; c++) total += clouds[c].nSlots;
    orc_vis_item* items = (orc_vis_item*)malloc(sizeof(orc_vis_item) * (size_t)(total > 0 ? total : 1));
    scr_visible_batch* tmp = (scr_visible_batch*)malloc(sizeof(scr_visible_batch) * (size_t)(total > 0 ? total : 1));
    if (!items || !tmp) { free(items); free(tmp); return -1; }
    int nItems = 0;
    for (int c = 0; c < nClouds; c++) {
        const int n = orc_cull(&clouds[c], cfg, view, proj, width, height, tmp);
        for (int i = 0; i < n; i++) { items[nItems].cloud = c; items[nItems].vb = tmp[i]; nItems++; }
    }
    free(tmp);
    if (nThreads < 1) nThreads = 1;
    if (nThreads > nItems) nThreads = nItems > 0 ? nItems : 1;
    ORC_LAP(1);

    /* depth phase (all clouds), then colour phase (all clouds):
     * ComputeRasteriser.swift:344-353 */
    uint32_t** priv = (uint32_t**)calloc((size_t)nThreads, sizeof(uint32_t*));
    orc_job* jobs = (orc_job*)calloc((size_t)nThreads, sizeof(orc_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nThreads, sizeof(pthread_t));
    int ok = priv && jobs && th;
    for (int phase = 0; ok && phase < 2; phase++) {
        for (int t = 0; t < nThreads; t++) {
            orc_job* j = &jobs[t];
            memset(j, 0, sizeof *j);
            j->clouds = clouds;
            j->items = items;
            j->first = (int)((long long)nItems * t / nThreads);
            j->last = (int)((long long)nItems * (t + 1) / nThreads);
            j->cfg = cfg;
            j->view = view;
            j->proj = proj;
            j->width = width;
            j->height = height;
            j->phase = phase;
            if (nThreads == 1) {
                j->depth = out->depth;
                j->r = out->red; j->g = out->green; j->b = out->blue; j->n = out->count;
            } else if (phase == 0) {
                priv[t] = (uint32_t*)calloc(npx, 4);
                if (!priv[t]) ok = 0;
                j->depth = priv[t];
            } else {
                priv[t] = (uint32_t*)calloc(npx * 4, 4);
                if (!priv[t]) ok = 0;
                j->depth = out->depth;
                j->r = priv[t]; j->g = priv[t] + npx; j->b = priv[t] + 2 * npx; j->n = priv[t] + 3 * npx;
            }
        }
        if (!ok) break;
        if (nThreads == 1) {
            orc_job_main(&jobs[0]);
            ORC_LAP(2 + phase);
            continue;
        }
        for (int t = 0; t < nThreads; t++) pthread_create(&th[t], NULL, orc_job_main, &jobs[t]);
        for (int t = 0; t < nThreads; t++) pthread_join(th[t], NULL);
        /* merge private planes: max (depth) / sum (colour) — both exact */
        for (int t = 0; t < nThreads; t++) {
            if (phase == 0) {
                for (size_t p = 0; p < npx; p++) if (priv[t][p] > out->depth[p]) out->depth[p] = priv[t][p];
            } else {
                for (size_t p = 0; p < npx; p++) {
                    out->red[p] += priv[t][p];
                    out->green[p] += priv[t][npx + p];
                    out->blue[p] += priv[t][2 * npx + p];
                    out->count[p] += priv[t][3 * npx + p];
                }
            }
            free(priv[t]);
            priv[t] = NULL;
        }
        ORC_LAP(2 + phase);
    }
    if (priv) for (int t = 0; t < nThreads; t++) free(priv[t]);
    free(priv);
    free(jobs);
    free(th);
    free(items);
    if (!ok) return -1;

    /* resolveUpdate, coverage path (Resolve/Shaders.metal:27-45) */
    if ((cfg->applyTint != 0 && cfg->tintAlphaIsCoverage != 0) || cfg->motionBlur > 0.0f) {
        for (size_t p = 0; p < npx; p++) {
            const uint32_t a = out->count[p];
            if (a == 0u) {
                if (out->rgba8)
                    out->rgba8[p] = orc_pack_rgba(cfg->backgroundColor[0], cfg->backgroundColor[1],
                                                  cfg->backgroundColor[2], 0.0f);
                if (out->depthF32) out->depthF32[p] = 0.0f;
                continue;
            }
            const float inv = orc_rcp((float)a);
            const float sumA = (float)a * (ORC_MUTANT == 7 ? 1.0f : 1.0f / 4096.0f);
            const float alpha = orc_saturate(1.0f - orc_expf(-sumA));
            if (out->rgba8)
                out->rgba8[p] = orc_pack_rgba((float)out->red[p] * inv, (float)out->green[p] * inv,
                                              (float)out->blue[p] * inv, alpha);
            if (out->depthF32) out->depthF32[p] = orc_uint_to_depth(out->depth[p]);
        }
        ORC_LAP(4);
        return 0;
    }
    /* resolveUpdate (count-average path) — Resolve/Shaders.metal:13-57 */
    for (size_t p = 0; p < npx; p++) {
        const uint32_t cnt = out->count[p];
        if (cnt == 0u) {
            if (out->rgba8)
                out->rgba8[p] = orc_pack_rgba(cfg->backgroundColor[0], cfg->backgroundColor[1],
                                              cfg->backgroundColor[2], 0.0f);
            if (out->depthF32) out->depthF32[p] = 0.0f;
            continue;
        }
        const float inv = orc_rcp((float)cnt);
        const float r = ((float)out->red[p] * inv) * ORC_INV255;
        const float g = ((float)out->green[p] * inv) * ORC_INV255;
        const float b = ((float)out->blue[p] * inv) * ORC_INV255;
        if (out->rgba8) out->rgba8[p] = orc_pack_rgba(r, g, b, 1.0f);
        if (out->depthF32) out->depthF32[p] = orc_uint_to_depth(out->depth[p]);
    }
    ORC_LAP(4);
#undef ORC_LAP
    return 0;
}

/* ------------------------------------------------------------------ nearest */
typedef struct orc_nearest_job {
    const orc_cloud* c;
    const scr_visible_batch* vis;
    int first, last, pass;
    const scr_config* cfg;
    const float* view;
    const float* proj;
    int width, height;
    const uint32_t* depthsIn; /* pass 1: the merged depths */
    uint32_t* out;            /* pass 0: depth maxima; pass 1: index minima */
} orc_nearest_job;

/* pass 0: nearestDepthUpdate (NearestDepth/Shaders.metal:14-91)
 * pass 1: nearestIndexUpdate (NearestIndex/Shaders.metal:14-94) */
static void* orc_nearest_main(void* arg) {
    const orc_nearest_job* j = (const orc_nearest_job*)arg;
    const orc_cloud* c = j->c;
    for (int v = j->first; v < j->last; v++) {
        const scr_visible_batch* vb = &j->vis[v];
        const scr_raster_batch* b = &c->batches[vb->batchIndex];
        const scr_raster_file* f = &c->files[b->fileIndex];
        for (uint32_t li = 0; li < vb->activePoints; li++) {
            const uint32_t pi = b->firstPoint + li;
            orc_frag fr;
            if (!orc_front(c, b, f, vb, pi, j->cfg, j->view, j->proj, j->width, j->height, 1, 1, &fr)) continue;
            for (int oy = -fr.radius; oy <= fr.radius; oy++)
                for (int ox = -fr.radius; ox <= fr.radius; ox++) {
                    if (!orc_inside_footprint(ox, oy, fr.radius)) continue;
                    const int tx = fr.px + ox, ty = fr.py + oy;
                    if (tx < 0 || tx >= j->width || ty < 0 || ty >= j->height) continue;
                    const size_t p = (size_t)ty * j->width + tx;
                    if (j->pass == 0) {
                        if (fr.depth > j->out[p]) j->out[p] = fr.depth;
                    } else if (fr.depth == j->depthsIn[p]) {
#if ORC_MUTANT == 3
                        if (j->out[p] == 0xffffffffu || pi > j->out[p]) j->out[p] = pi;
#else
                        if (pi < j->out[p]) j->out[p] = pi;
#endif
                    }
                }
        }
    }
    return NULL;
}

int orc_render_nearest(const orc_cloud* c, const scr_config* cfg, const float* view,
                       const float* proj, int width, int height, uint32_t* depths,
                       uint32_t* indices, uint32_t* rgba8, float* depthF32, int nThreads) {
    const size_t npx = (size_t)width * (size_t)height;
    /* clearWinnerUpdate — ClearWinner/Shaders.metal:7-19 */
    for (size_t p = 0; p < npx; p++) { depths[p] = 0u; indices[p] = 0xffffffffu; }
    scr_visible_batch* vis = (scr_visible_batch*)malloc(sizeof(scr_visible_batch) * (size_t)(c->nSlots > 0 ? c->nSlots : 1));
    if (!vis) return -1;
    const int nv = orc_cull(c, cfg, view, proj, width, height, vis);
    if (nThreads < 1) nThreads = 1;
    if (nThreads > nv) nThreads = nv > 0 ? nv : 1;
    /* threads over visible-batch ranges, each with a private plane; the
     * planes merge by max (depths) and min (indices), so the result is the
     * single-thread one */
    orc_nearest_job* jobs = (orc_nearest_job*)calloc((size_t)nThreads, sizeof(orc_nearest_job));
    pthread_t* th = (pthread_t*)calloc((size_t)nThreads, sizeof(pthread_t));
    int ok = jobs && th;
    for (int pass = 0; ok && pass < 2; pass++) {
        for (int t = 0; t < nThreads; t++) {
            orc_nearest_job* j = &jobs[t];
            j->c = c;
            j->vis = vis;
            j->first = (int)((long long)nv * t / nThreads);
            j->last = (int)((long long)nv * (t + 1) / nThreads);
            j->pass = pass;
            j->cfg = cfg;
            j->view = view;
            j->proj = proj;
            j->width = width;
            j->height = height;
            j->depthsIn = depths;
            if (nThreads == 1) {
                j->out = pass == 0 ? depths : indices;
            } else {
                j->out = (uint32_t*)malloc(npx * 4 + 4);
                if (!j->out) { ok = 0; break; }
                for (size_t p = 0; p < npx; p++) j->out[p] = pass == 0 ? 0u : 0xffffffffu;
            }
        }
        if (!ok) break;
        if (nThreads == 1) {
            orc_nearest_main(&jobs[0]);
            continue;
        }
        for (int t = 0; t < nThreads; t++) pthread_create(&th[t], NULL, orc_nearest_main, &jobs[t]);
        for (int t = 0; t < nThreads; t++) pthread_join(th[t], NULL);
        for (int t = 0; t < nThreads; t++) {
            uint32_t* o = jobs[t].out;
            if (pass == 0) {
                for (size_t p = 0; p < npx; p++) if (o[p] > depths[p]) depths[p] = o[p];
            } else {
                for (size_t p = 0; p < npx; p++) {
#if ORC_MUTANT == 3
                    if (o[p] != 0xffffffffu && (indices[p] == 0xffffffffu || o[p] > indices[p])) indices[p] = o[p];
#else
                    if (o[p] < indices[p]) indices[p] = o[p];
#endif
                }
            }
            free(o);
            jobs[t].out = NULL;
        }
    }
    if (jobs && nThreads > 1)
        for (int t = 0; t < nThreads; t++) free(jobs[t].out);
    free(jobs);
    free(th);
    if (!ok) { free(vis); return -1; }
    free(vis);
    /* nearestResolveUpdate — NearestResolve/Shaders.metal:8-38 */
    for (size_t p = 0; p < npx; p++) {
        const uint32_t d = depths[p], idx = indices[p];
        if (d == 0u || idx == 0xffffffffu) {
            if (rgba8) rgba8[p] = orc_pack_rgba(cfg->backgroundColor[0], cfg->backgroundColor[1],
                                                cfg->backgroundColor[2], 0.0f);
            if (depthF32) depthF32[p] = 0.0f;
            continue;
        }
        const uint32_t col = c->colors[idx];
        if (rgba8)
            rgba8[p] = orc_pack_rgba((float)(col & 0xffu) * ORC_INV255, (float)((col >> 8) & 0xffu) * ORC_INV255,
                                     (float)((col >> 16) & 0xffu) * ORC_INV255, 1.0f);
        if (depthF32) depthF32[p] = orc_uint_to_depth(d);
    }
    return 0;
}

/* NearestPointCPUReference.render — NearestPointCPUReference.swift:16-81:
 * every batch (optional frustum), all numPoints (no CLOD skip), radius 0. */
void orc_nearest_cpu_reference(const orc_cloud* c, int width, int height, const float* view,
                               const float* proj, int enableFrustumCulling, uint32_t* depths,
                               uint32_t* indices) {
    const size_t npx = (size_t)(width > 0 ? width : 0) * (size_t)(height > 0 ? height : 0);
    for (size_t p = 0; p < npx; p++) { depths[p] = 0u; indices[p] = 0xffffffffu; }
    if (width <= 0 || height <= 0) return;
    for (int s = 0; s < c->nSlots; s++) {
        const scr_raster_batch* b = &c->batches[s];
        const scr_raster_file* f = &c->files[b->fileIndex];
        const float mn[3] = {b->minX, b->minY, b->minZ};
        const float mx[3] = {b->maxX, b->maxY, b->maxZ};
        if (enableFrustumCulling && !orc_intersects_frustum(f->transformFrustum, mn, mx)) continue;
        const int level = orc_precision_level(orc_pixel_size_on_screen(b, f, view, proj, width, height));
        for (uint32_t li = 0; li < b->numPoints; li++) {
            const uint32_t pi = b->firstPoint + li;
            float pt[3], clip[4];
            orc_decode_point(pi, b, c->xyzLow, c->xyzMed, c->xyzHigh, level, 0, pt);
            const float p4[4] = {pt[0], pt[1], pt[2], 1.0f};
            orc_mat4_vec(f->transform, p4, clip);
            if (clip[3] <= 0.0f) continue;
            const float rw = orc_rcp(clip[3]);
            const float nx = clip[0] * rw, ny = clip[1] * rw, nz = clip[2] * rw;
            if (nx < -1.0f || nx > 1.0f || ny < -1.0f || ny > 1.0f || nz <= 0.0f || nz > 1.0f) continue;
            int px = orc_f2i((nx * 0.5f + 0.5f) * (float)width);
            int py = orc_f2i((ny * 0.5f + 0.5f) * (float)height);
            if (px < 0 || px >= width || py < 0 || py >= height) continue;
            py = height - 1 - py;
            const size_t p = (size_t)py * width + px;
            const uint32_t d = orc_depth_to_uint(nz);
            if (d > depths[p]) { depths[p] = d; indices[p] = pi; }
            else if (d == depths[p] && pi < indices[p]) indices[p] = pi;
        }
    }
}

/* pickPointIndex — ComputeRasteriser.swift:617-669 (pixel mapping and the
 * nearest non-sentinel search within the circular window) */
int orc_pick_search(const uint32_t* indices, int width, int height, float ndcX, float ndcY,
                    int searchRadius, uint32_t* outIndex) {
    if (width <= 0 || height <= 0) return 0;
    int px = orc_f2i((ndcX * 0.5f + 0.5f) * (float)width);
    if (px < 0) px = 0;
    if (px > width - 1) px = width - 1;
    const int pyb = orc_f2i((ndcY * 0.5f + 0.5f) * (float)height);
    int py = height - 1 - pyb;
    if (py < 0) py = 0;
    if (py > height - 1) py = height - 1;
    const int r = searchRadius > 0 ? searchRadius : 0;
    const int y0 = py - r < 0 ? 0 : py - r;
    const int y1 = py + r > height - 1 ? height - 1 : py + r;
    long long bestDist = 0x7fffffffffffffffLL;
    int found = 0;
    const long long r2 = (long long)r * r;
    for (int y = y0; y <= y1; y++) {
        const int dy = y - py;
        const int x0 = px - r < 0 ? 0 : px - r;
        const int x1 = px + r > width - 1 ? width - 1 : px + r;
        for (int x = x0; x <= x1; x++) {
            const int dx = x - px;
            const long long dist = (long long)dx * dx + (long long)dy * dy;
            if (dist > r2 || dist >= bestDist) continue;
            const uint32_t v = indices[(size_t)y * width + x];
            if (v != 0xffffffffu) { bestDist = dist; *outIndex = v; found = 1; }
        }
    }
    return found;
}

/* ------------------------------------------------------------------ CPU pack */

typedef struct orc_keyidx {
    uint64_t key;
    uint32_t idx;
} orc_keyidx;

static int orc_keyidx_cmp(const void* a, const void* b) {
    const orc_keyidx* x = (const orc_keyidx*)a;
    const orc_keyidx* y = (const orc_keyidx*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0); /* C6: stable */
}

static void orc_bounds(const float* pos, int stride, size_t n, float bmin[3], float bmax[3]) {
    for (int a = 0; a < 3; a++) { bmin[a] = 3.40282347e38f; bmax[a] = -3.40282347e38f; }
    for (size_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            const float v = pos[i * (size_t)stride + a];
            bmin[a] = fminf(bmin[a], v);
            bmax[a] = fmaxf(bmax[a], v);
        }
}

/* mortonOrder keys — PackedPointCloudFixtures.swift:219-235 (30-bit); the
 * 63-bit variant quantises to 21 bits/axis with 2097151 in place of 1023. */
static void orc_keys_with_bounds(const float* pos, int stride, size_t n, int keyBits,
                                 const float bmin[3], const float bmax[3], uint64_t* keys) {
    const float qmax = keyBits == 63 ? 2097151.0f : 1023.0f;
    float scale[3];
    for (int a = 0; a < 3; a++) scale[a] = qmax / fmaxf(bmax[a] - bmin[a], 0.000001f);
    for (size_t i = 0; i < n; i++) {
        uint32_t q[3];
        for (int a = 0; a < 3; a++) {
            float v = (pos[i * (size_t)stride + a] - bmin[a]) * scale[a];
            v = fminf(fmaxf(v, 0.0f), qmax);
            q[a] = orc_f2u(v);
        }
        keys[i] = keyBits == 63 ? orc_morton63(q[0], q[1], q[2]) : (uint64_t)orc_morton30(q[0], q[1], q[2]);
    }
}

void orc_morton_keys(const float* positions, int posStride, size_t count, int keyBits, uint64_t* keys) {
    float bmin[3], bmax[3];
    orc_bounds(positions, posStride, count, bmin, bmax);
    orc_keys_with_bounds(positions, posStride, count, keyBits, bmin, bmax, keys);
}

/* computeLODLevels — PackedPointCloudFixtures.swift:246-335 (open-addressing
 * set of 21-bit/axis voxel keys, Fibonacci hashing, one table for all levels) */
void orc_lod_levels(const float* sp, size_t count, const float bmin[3], const float bmax[3],
                    int lodLevels, int coarse, uint8_t* levels) {
    const uint8_t maxLevel = (uint8_t)(lodLevels - 1);
    for (size_t i = 0; i < count; i++) levels[i] = maxLevel;
    if (lodLevels <= 1 || count == 0) return;
    float ext[3];
    for (int a = 0; a < 3; a++) ext[a] = fmaxf(bmax[a] - bmin[a], 0.000001f);
    const float longest = fmaxf(ext[0], fmaxf(ext[1], ext[2]));
    const float baseVoxel = longest / (float)coarse;
    size_t cap = 1;
    while (cap < count * 2) cap <<= 1;
    const size_t mask = cap - 1;
    int tz = 0;
    while (((size_t)1 << tz) < cap) tz++;
    const unsigned shift = (unsigned)(64 - tz);
    uint64_t* table = (uint64_t*)malloc(cap * sizeof(uint64_t));
    if (!table) return;
    for (int level = 0; level < lodLevels - 1; level++) {
        const float voxel = baseVoxel * powf(0.5f, (float)level);
        const float inv = 1.0f / fmaxf(voxel, 0.000001f);
        for (size_t s = 0; s < cap; s++) table[s] = UINT64_MAX;
        for (size_t i = 0; i < count; i++) {
            if (levels[i] != maxLevel) continue;
            uint64_t cell[3];
            for (int a = 0; a < 3; a++) {
                const float local = (sp[i * 3 + a] - bmin[a]) * inv;
                int32_t ci = orc_f2i(floorf(local));
                if (ci < 0) ci = 0;
                cell[a] = (uint64_t)ci & 0x1fffffu;
            }
            const uint64_t key = (cell[0] << 42) | (cell[1] << 21) | cell[2];
            size_t slot = (size_t)((key * 0x9E3779B97F4A7C15ull) >> shift);
            if (shift == 64) slot = 0;
            for (;;) {
                const uint64_t cur = table[slot];
                if (cur == UINT64_MAX) { table[slot] = key; levels[i] = (uint8_t)level; break; }
                if (cur == key) break;
                slot = (slot + 1) & mask;
            }
        }
    }
    free(table);
}

/* GPUPacker.init's LOD gate — GPUPacker.swift:35, 86-99: the dense grid has
 * at most (coarse * 2^(L - 2) + 2)^3 cells; above 268,435,456 the GPU packer
 * skips LOD and leaves every level at the finest (L - 1). */
int orc_gpu_lod_enabled(int lodLevels, int coarse) {
    if (lodLevels < 1) lodLevels = 1;
    if (lodLevels > 8) lodLevels = 8;
    if (coarse < 1) coarse = 1;
    if (lodLevels <= 1) return 0;
    const long long maxDim = (long long)coarse * (1ll << (lodLevels - 2)) + 2;
    if (maxDim > 1000000) return 0;  /* the cube would exceed the cap many times over */
    return maxDim * maxDim * maxDim <= 268435456ll;
}

static int orc_pack_impl(const float* pos, int stride, const float* col, size_t count, int ppb, int lodLevels,
                         int coarse, int keyBits, int gpuRules, orc_pack_out* out);

/* pack — PackedPointCloudFixtures.swift:30-167 */
int orc_pack(const float* pos, int stride, const float* col, size_t count, int ppb, int lodLevels,
             int coarse, int keyBits, orc_pack_out* out) {
    return orc_pack_impl(pos, stride, col, count, ppb, lodLevels, coarse, keyBits, 0, out);
}

/* The same pack under GPUPacker's LOD gate (orc_gpu_lod_enabled). */
int orc_pack_gpu(const float* pos, int stride, const float* col, size_t count, int ppb, int lodLevels,
                 int coarse, int keyBits, orc_pack_out* out) {
    return orc_pack_impl(pos, stride, col, count, ppb, lodLevels, coarse, keyBits, 1, out);
}

static int orc_pack_impl(const float* pos, int stride, const float* col, size_t count, int ppb, int lodLevels,
                         int coarse, int keyBits, int gpuRules, orc_pack_out* out) {
    if (ppb < 1 || ppb > 65535 || (keyBits != 30 && keyBits != 63) || (stride != 3 && stride != 4)) return -1;
    if (count == 0) {
        for (int a = 0; a < 3; a++) { out->boundsMin[a] = 0.0f; out->boundsMax[a] = 0.0f; }
        return 0;
    }
    if (lodLevels < 1) lodLevels = 1;
    if (lodLevels > 8) lodLevels = 8;
    if (coarse < 1) coarse = 1;
    orc_bounds(pos, stride, count, out->boundsMin, out->boundsMax);

    /* mortonOrder — :219-240 (stable, C6) */
    uint64_t* keys = (uint64_t*)malloc(count * sizeof(uint64_t));
    orc_keyidx* ki = (orc_keyidx*)malloc(count * sizeof(orc_keyidx));
    uint32_t* order = (uint32_t*)malloc(count * sizeof(uint32_t));
    float* sp = (float*)malloc(count * 3 * sizeof(float));
    uint8_t* lv = (uint8_t*)malloc(count);
    uint32_t* perm = (uint32_t*)malloc((size_t)ppb * sizeof(uint32_t));
    float* tmpP = (float*)malloc((size_t)ppb * 3 * sizeof(float));
    uint8_t* tmpL = (uint8_t*)malloc((size_t)ppb);
    uint32_t* tmpO = (uint32_t*)malloc((size_t)ppb * sizeof(uint32_t));
    if (!keys || !ki || !order || !sp || !lv || !perm || !tmpP || !tmpL || !tmpO) {
        free(keys); free(ki); free(order); free(sp); free(lv); free(perm); free(tmpP); free(tmpL); free(tmpO);
        return -2;
    }
    orc_keys_with_bounds(pos, stride, count, keyBits, out->boundsMin, out->boundsMax, keys);
    for (size_t i = 0; i < count; i++) { ki[i].key = keys[i]; ki[i].idx = (uint32_t)i; }
    qsort(ki, count, sizeof(orc_keyidx), orc_keyidx_cmp);
    for (size_t i = 0; i < count; i++) {
        order[i] = ki[i].idx;
        for (int a = 0; a < 3; a++) sp[i * 3 + a] = pos[(size_t)ki[i].idx * stride + a];
    }
    free(keys);
    free(ki);

    if (gpuRules && !orc_gpu_lod_enabled(lodLevels, coarse)) memset(lv, lodLevels - 1, count);
    else orc_lod_levels(sp, count, out->boundsMin, out->boundsMax, lodLevels, coarse, lv);

    /* bucketSortBatchSlicesByLevel — :175-214 */
    for (size_t first = 0; first < count; first += (size_t)ppb) {
        const size_t end = first + (size_t)ppb < count ? first + (size_t)ppb : count;
        size_t cur[9] = {0};
        for (size_t i = first; i < end; i++) cur[(lv[i] & 7) + 1]++;
        for (int l = 1; l < 9; l++) cur[l] += cur[l - 1];
        for (size_t i = first; i < end; i++) perm[cur[lv[i] & 7]++] = (uint32_t)i;
        const size_t n = end - first;
        for (size_t j = 0; j < n; j++) {
            tmpO[j] = order[perm[j]];
            tmpL[j] = lv[perm[j]];
            for (int a = 0; a < 3; a++) tmpP[j * 3 + a] = sp[(size_t)perm[j] * 3 + a];
        }
        for (size_t j = 0; j < n; j++) {
            order[first + j] = tmpO[j];
            lv[first + j] = tmpL[j];
            for (int a = 0; a < 3; a++) sp[(first + j) * 3 + a] = tmpP[j * 3 + a];
        }
    }

    /* batches, quantisation, cumulative counts — :85-144 */
    int nb = 0;
    for (size_t first = 0; first < count; first += (size_t)ppb, nb++) {
        const size_t end = first + (size_t)ppb < count ? first + (size_t)ppb : count;
        float mn[3], mx[3], size[3];
        for (int a = 0; a < 3; a++) { mn[a] = 3.40282347e38f; mx[a] = -3.40282347e38f; }
        for (size_t i = first; i < end; i++)
            for (int a = 0; a < 3; a++) { mn[a] = fminf(mn[a], sp[i * 3 + a]); mx[a] = fmaxf(mx[a], sp[i * 3 + a]); }
        for (int a = 0; a < 3; a++) size[a] = fmaxf(mx[a] - mn[a], 0.000001f);
        for (size_t i = first; i < end; i++) {
            uint32_t q[3];
            for (int a = 0; a < 3; a++) {
                float nrm = (sp[i * 3 + a] - mn[a]) / size[a]; /* Swift: IEEE division */
                nrm = fminf(fmaxf(nrm, 0.0f), 0.99999994f);
                q[a] = orc_f2u(nrm * 1073741824.0f); /* Float(2^30 - 1) == 2^30 */
            }
            out->xyzLow[i] = ((q[0] >> 20) & 1023u) | (((q[1] >> 20) & 1023u) << 10) | (((q[2] >> 20) & 1023u) << 20);
            out->xyzMed[i] = ((q[0] >> 10) & 1023u) | (((q[1] >> 10) & 1023u) << 10) | (((q[2] >> 10) & 1023u) << 20);
            out->xyzHigh[i] = (q[0] & 1023u) | ((q[1] & 1023u) << 10) | ((q[2] & 1023u) << 20);
        }
        scr_raster_batch b;
        memset(&b, 0, sizeof b);
        b.state = 1;
        b.minX = mn[0]; b.minY = mn[1]; b.minZ = mn[2];
        b.maxX = mx[0]; b.maxY = mx[1]; b.maxZ = mx[2];
        b.numPoints = (uint32_t)(end - first);
        b.firstPoint = (uint32_t)first;
        b.fileIndex = 0;
        uint32_t cum[8] = {0};
        for (size_t i = first; i < end; i++) cum[lv[i] & 7]++;
        for (int l = 1; l < 8; l++) cum[l] += cum[l - 1];
        b.lodCum01 = cum[0] | (cum[1] << 16);
        b.lodCum23 = cum[2] | (cum[3] << 16);
        b.lodCum45 = cum[4] | (cum[5] << 16);
        b.lodCum67 = cum[6] | (cum[7] << 16);
        out->batches[nb] = b;
    }

    /* colour pack (truncating) — :146-152 */
    for (size_t i = 0; i < count; i++) {
        const float* c = col + (size_t)order[i] * 4;
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) v |= orc_f2u(fminf(fmaxf(c[k], 0.0f), 1.0f) * 255.0f) << (8 * k);
        out->colors[i] = v;
        out->levels[i] = lv[i];
        out->sourceIndices[i] = order[i];
        if (out->orderedPositions)
            for (int a = 0; a < 3; a++) out->orderedPositions[i * 3 + a] = sp[i * 3 + a];
    }
    free(order); free(sp); free(lv); free(perm); free(tmpP); free(tmpL); free(tmpO);
    return nb;
}

/* cubeGrid — PackedPointCloudFixtures.swift:5-25 */
void orc_cube_grid(int n, float extent, float* positions3, float* colors4) {
    if (n < 2) n = 2;
    size_t i = 0;
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++, i++) {
                const float fx = (float)x / (float)(n - 1);
                const float fy = (float)y / (float)(n - 1);
                const float fz = (float)z / (float)(n - 1);
                positions3[i * 3 + 0] = (fx - 0.5f) * extent;
                positions3[i * 3 + 1] = (fy - 0.5f) * extent;
                positions3[i * 3 + 2] = (fz - 0.5f) * extent;
                colors4[i * 4 + 0] = fx;
                colors4[i * 4 + 1] = fy;
                colors4[i * 4 + 2] = fz;
                colors4[i * 4 + 3] = 1.0f;
            }
}

/* ------------------------------------------------------------------ differentiable HQS
 * No reference counterpart: the HQS passes with the per-point colour taken
 * from float parameters and accumulated as 64-bit fixed point (x 2^16).
 * Visibility (cull, depth pass, the colour pass's depth-tolerance test) is
 * the HQS one above, so depth and count equal the HQS planes. */
int64_t orc_diff_fixed(float v) {
    float x = v * 65536.0f;
    if (!(x == x)) return 0;
    if (x > 9.0e18f) x = 9.0e18f;
    if (x < -9.0e18f) x = -9.0e18f;
    return (int64_t)rintf(x);
}

float orc_diff_pixel(int64_t sum, uint32_t count, float background) {
    if (count == 0u) return background;
    return (float)((double)sum / (65536.0 * (double)count));
}

/* Walks every fragment that passes the colour pass's depth test, calling
 * visit(ctx, pointIndex, pixel) in the order of the single-threaded passes. */
typedef void (*orc_frag_visit)(void* ctx, uint32_t pointIndex, size_t pixel);

static int orc_visible_fragments(const orc_cloud* c, const scr_config* cfg, const float* view, const float* proj,
                                 int width, int height, const uint32_t* depth, const uint8_t* mask,
                                 orc_frag_visit visit, void* vctx) {
    scr_visible_batch* vis = (scr_visible_batch*)malloc(sizeof(scr_visible_batch) * (size_t)(c->nSlots > 0 ? c->nSlots : 1));
    if (!vis) return -1;
    const int nv = orc_cull(c, cfg, view, proj, width, height, vis);
    const float keep = 1.0f - cfg->depthTolerance;
    for (int v = 0; v < nv; v++) {
        const scr_raster_batch* b = &c->batches[vis[v].batchIndex];
        const scr_raster_file* f = &c->files[b->fileIndex];
        for (uint32_t li = 0; li < vis[v].activePoints; li++) {
            const uint32_t pi = b->firstPoint + li;
            if (mask && !mask[pi]) continue;
            orc_frag fr;
            if (!orc_front(c, b, f, &vis[v], pi, cfg, view, proj, width, height, 0, 1, &fr)) continue;
            for (int oy = -fr.radius; oy <= fr.radius; oy++)
                for (int ox = -fr.radius; ox <= fr.radius; ox++) {
                    if (!orc_inside_footprint(ox, oy, fr.radius)) continue;
                    const int tx = fr.px + ox, ty = fr.py + oy;
                    if (tx < 0 || tx >= width || ty < 0 || ty >= height) continue;
                    const size_t p = (size_t)ty * width + tx;
                    const uint32_t d = depth[p];
                    if (d == 0u) continue;
                    if (!(fr.ndcZ >= orc_uint_to_depth(d) * keep)) continue;
                    visit(vctx, pi, p);
                }
        }
    }
    free(vis);
    return 0;
}

typedef struct orc_diff_fwd_ctx {
    const float* params;
    int64_t* sums;
} orc_diff_fwd_ctx;

static void orc_diff_fwd_visit(void* vctx, uint32_t pi, size_t p) {
    orc_diff_fwd_ctx* x = (orc_diff_fwd_ctx*)vctx;
    for (int k = 0; k < 3; k++)
        x->sums[p * 3 + k] = (int64_t)((uint64_t)x->sums[p * 3 + k] + (uint64_t)orc_diff_fixed(x->params[(size_t)pi * 4 + k]));
}

int orc_render_diff(const orc_cloud* c, const scr_config* cfg, const float* view, const float* proj, int width,
                    int height, const float* params, uint32_t* depth, uint32_t* count, int64_t* sums, float* image,
                    int nThreads) {
    const size_t npx = (size_t)width * (size_t)height;
    uint32_t* tmp = (uint32_t*)malloc(npx * 3 * 4 + 4);
    if (!tmp) return -1;
    orc_hqs_planes pl = {depth, tmp, tmp + npx, tmp + 2 * npx, count, NULL, NULL};
    int rc = orc_render_hqs(c, 1, cfg, view, proj, width, height, &pl, nThreads);
    free(tmp);
    if (rc != 0) return rc;
    memset(sums, 0, npx * 3 * sizeof(int64_t));
    orc_diff_fwd_ctx x = {params, sums};
    rc = orc_visible_fragments(c, cfg, view, proj, width, height, depth, NULL, orc_diff_fwd_visit, &x);
    if (rc != 0) return rc;
    for (size_t p = 0; p < npx; p++)
        for (int k = 0; k < 3; k++) image[p * 3 + k] = orc_diff_pixel(sums[p * 3 + k], count[p], cfg->backgroundColor[k]);
    return 0;
}

typedef struct orc_diff_bwd_ctx {
    const uint32_t* count;
    const float* dLdImage;
    double* grad;
} orc_diff_bwd_ctx;

static void orc_diff_bwd_visit(void* vctx, uint32_t pi, size_t p) {
    orc_diff_bwd_ctx* x = (orc_diff_bwd_ctx*)vctx;
    const double inv = 1.0 / (double)x->count[p];
    for (int k = 0; k < 3; k++) x->grad[(size_t)pi * 4 + k] += (double)x->dLdImage[p * 3 + k] * inv;
}

/* dL/dparam = sum over the point's visible fragments of dL/dpixel x (1 / count)
 * (the fixed-point rounding passes the gradient straight through), FP64. */
int orc_diff_backward(const orc_cloud* c, const scr_config* cfg, const float* view, const float* proj, int width,
                      int height, const uint32_t* depth, const uint32_t* count, const float* dLdImage, double* grad,
                      const uint8_t* mask) {
    orc_diff_bwd_ctx x = {count, dLdImage, grad};
    return orc_visible_fragments(c, cfg, view, proj, width, height, depth, mask, orc_diff_bwd_visit, &x);
}
