// Host-side property check of the per-batch bounds the cull pass computes
// (batch_bounds): for random and adversarial batches, cameras and viewports,
// every point that decodes inside a batch is run through the draw passes'
// exact per-point arithmetic (the batch projection of its quantised
// coordinates, project_q, correctly rounded reciprocal, viewport fma,
// truncating conversion), and
//   * with kBoundsProven, every fragment's column / unflipped row lies in the
//     rectangle and its ndc z is <= zmax;
//   * with kBoundsInside, every point passes the clip-w, NDC (nearest's
//     z > 0 included) and viewport tests;
//   * with kBoundsFastRcp, every clip w lies in [2^-125, 2^125].
// It then recomputes the bounds with the rounding allowance removed and
// requires violations to appear on the adversarial set (so the set is sharp
// enough to catch a missing allowance).
// Built by tests/test_native_checks.py with hipcc for the host (no GPU use).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "scr_math.hpp"

using namespace scr;

static void look_at(const double eye[3], const double tgt[3], const double up[3], float* m) {
    double f[3] = {tgt[0] - eye[0], tgt[1] - eye[1], tgt[2] - eye[2]};
    double n = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= n;
    double r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    n = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double& v : r) v /= n;
    const double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    const double rows[3][3] = {{r[0], r[1], r[2]}, {u[0], u[1], u[2]}, {-f[0], -f[1], -f[2]}};
    for (int i = 0; i < 16; i++) m[i] = 0.0f;
    for (int row = 0; row < 3; row++) {
        double t = 0.0;
        for (int c = 0; c < 3; c++) {
            m[c * 4 + row] = (float)rows[row][c];
            t -= rows[row][c] * eye[c];
        }
        m[12 + row] = (float)t;
    }
    m[15] = 1.0f;
}

// Reverse-Z perspective (view z = -near -> ndc 1, z = -far -> 0), column-major.
static void perspective(double fovDeg, double aspect, double near, double far, float* m) {
    const double f = 1.0 / std::tan(fovDeg * 3.141592653589793 / 360.0);
    for (int i = 0; i < 16; i++) m[i] = 0.0f;
    m[0] = (float)(f / aspect);
    m[5] = (float)f;
    m[10] = (float)(near / (far - near));
    m[11] = -1.0f;
    m[14] = (float)(near * far / (far - near));
}

static void ortho(double h, double aspect, double near, double far, float* m) {
    for (int i = 0; i < 16; i++) m[i] = 0.0f;
    m[0] = (float)(2.0 / (h * aspect));
    m[5] = (float)(2.0 / h);
    m[10] = (float)(1.0 / (far - near));
    m[14] = (float)(far / (far - near));
    m[15] = 1.0f;
}

struct Frag {
    bool okColor, okNearest;
    int ix, iy;
    float z, w;
};

// The draw passes' per-point front end for quantised coordinates q.
static Frag front(const float* T, const scr_raster_batch& b, int level, uint32_t qx, uint32_t qy, uint32_t qz, int W,
                  int H) {
    const float inv = level == 0 ? (1.0f / 1073741824.0f) : (level == 1 ? (1.0f / 1048576.0f) : (1.0f / 1024.0f));
    const float sx = fmaxf(b.maxX - b.minX, 0.000001f) * inv;
    const float sy = fmaxf(b.maxY - b.minY, 0.000001f) * inv;
    const float sz = fmaxf(b.maxZ - b.minZ, 0.000001f) * inv;
    const BatchAffine m = batch_affine(T, sx, sy, sz, b.minX, b.minY, b.minZ);
    const f4 clip = project_q(m, (float)qx, (float)qy, (float)qz);
    const float rw = rcp(clip.w);
    const float nx = clip.x * rw, ny = clip.y * rw, nz = clip.z * rw;
    Frag f;
    f.ix = f2i(fmaf(nx, 0.5f, 0.5f) * (float)W);
    f.iy = f2i(fmaf(ny, 0.5f, 0.5f) * (float)H);
    f.z = nz;
    f.w = clip.w;
    const bool xy = !(clip.w <= 0.0f) && !(nx < -1.0f || ny < -1.0f || nx > 1.0f || ny > 1.0f) &&
                    !(f.ix >= W || f.iy >= H);
    f.okColor = xy && !(nz < 0.0f || nz > 1.0f);
    f.okNearest = xy && !(nz <= 0.0f || nz > 1.0f);
    return f;
}

// Bounds with the rounding allowance removed (amax = 0): the plain corner
// extremes, which the per-point rounding can leave.
static BatchBounds bounds_without_allowance(const float* m, const scr_raster_batch& b, int W, int H) {
    double lo[3], hi[3];
    const double a[4] = {0.0, 0.0, 0.0, 0.0};
    batch_box(b, lo, hi);
    CornerExtremes e = corner_extremes(corner_terms(m, lo[0], lo[1], lo[2], a));
    for (int k = 1; k < 8; k++)
        e = combine(e, corner_extremes(corner_terms(m, (k & 1) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1],
                                                    (k & 4) ? hi[2] : lo[2], a)));
    return finish_bounds(e, W, H, -1);
}

struct Stats {
    long cases = 0, proven = 0, inside = 0, fast = 0, points = 0, frags = 0, violations = 0;
};

static void check_case(const float* T, const scr_raster_batch& b, int level, int W, int H, std::mt19937_64& rng,
                       int samples, Stats& st, bool planted) {
    const BatchBounds bb = planted ? bounds_without_allowance(T, b, W, H) : batch_bounds(T, b, W, H, -1);
    st.cases++;
    const bool proven = (bb.flags & kBoundsProven) != 0u, inside = (bb.flags & kBoundsInside) != 0u;
    const bool fast = (bb.flags & kBoundsFastRcp) != 0u;
    st.proven += proven;
    st.inside += inside;
    st.fast += fast;
    const uint32_t qmax = level == 0 ? (1u << 30) - 1u : (level == 1 ? (1u << 20) - 1u : 1023u);
    std::uniform_int_distribution<uint32_t> Q(0u, qmax);
    for (int s = 0; s < samples; s++) {
        uint32_t q[3];
        for (int a = 0; a < 3; a++) {
            // the box corners and faces get their share of samples
            const int kind = (int)(rng() % 8);
            q[a] = kind == 0 ? 0u : (kind == 1 ? qmax : Q(rng));
        }
        const Frag f = front(T, b, level, q[0], q[1], q[2], W, H);
        st.points++;
        bool bad = false;
        if (fast && !(f.w >= 0x1p-125f && f.w <= 0x1p125f)) bad = true;
        if (inside && !(f.okColor && f.okNearest)) bad = true;
        if (proven && f.okColor) {
            st.frags++;
            if (f.ix < bb.x0 || f.ix > bb.x1 || f.iy < bb.y0 || f.iy > bb.y1 || !(f.z <= bb.zmax)) bad = true;
        }
        if (bad) {
            if (!planted && st.violations < 5)
                std::printf("violation: box [%g %g %g]-[%g %g %g] level %d W %d H %d q %u %u %u -> ix %d iy %d z %.9g "
                            "w %g | rect [%d %d]x[%d %d] zmax %.9g flags %u\n",
                            b.minX, b.minY, b.minZ, b.maxX, b.maxY, b.maxZ, level, W, H, q[0], q[1], q[2], f.ix, f.iy,
                            f.z, f.w, bb.x0, bb.x1, bb.y0, bb.y1, bb.zmax, bb.flags);
            st.violations++;
        }
    }
}

// One scenario: a camera, a viewport and a batch near the camera's target.
static void scenario(std::mt19937_64& rng, bool adversarial, float* T, scr_raster_batch& b, int& level, int& W, int& H) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    // far offsets: the world coordinates (and the transform's partial sums)
    // dwarf the distance to the camera
    const double off = adversarial ? (U(rng) < 0.5 ? 5e5 : 2e4) * (U(rng) < 0.5 ? 1 : -1) : (U(rng) - 0.5) * 2000.0;
    const double dist = adversarial ? 0.3 + U(rng) * 3.0 : 5.0 + U(rng) * 4000.0;
    const double size = adversarial ? 0.05 + U(rng) * 2.0 : 1.0 + U(rng) * 300.0;
    const double c[3] = {off + (U(rng) - 0.5) * 10.0, off * 0.5 + (U(rng) - 0.5) * 10.0, (U(rng) - 0.5) * 50.0};
    const double th = U(rng) * 6.283185307179586, ph = (U(rng) - 0.5) * 2.8;
    const double eye[3] = {c[0] + dist * std::cos(th) * std::cos(ph), c[1] + dist * std::sin(th) * std::cos(ph),
                           c[2] + dist * std::sin(ph)};
    // aim at the box, or off to one side so the box straddles a screen edge
    const double aimOff = U(rng) < 0.5 ? 0.0 : size * (U(rng) - 0.5) * 3.0;
    const double tgt[3] = {c[0] + aimOff, c[1] - aimOff, c[2]};
    const double up[3] = {0.0, 0.0, 1.0};
    float V[16], P[16];
    look_at(eye, tgt, up, V);
    const int res = (int)(rng() % 4);
    W = res == 0 ? 256 : (res == 1 ? 512 : (res == 2 ? 1920 : 3840 + (int)(rng() % 300)));
    H = res == 0 ? 256 : (res == 1 ? 384 : (res == 2 ? 1080 : 2160));
    if (U(rng) < 0.15) ortho(dist * 2.0, (double)W / H, 0.01, dist * 3.0, P);
    else perspective(20.0 + U(rng) * 90.0, (double)W / H, adversarial ? 0.01 : 0.1, 1e5, P);
    mat_mul(P, V, T);
    std::memset(&b, 0, sizeof b);
    b.state = 1;
    b.minX = (float)(c[0] - size * U(rng));
    b.minY = (float)(c[1] - size * U(rng));
    b.minZ = (float)(c[2] - size * U(rng));
    b.maxX = b.minX + (float)(size * U(rng));
    b.maxY = b.minY + (float)(size * U(rng));
    b.maxZ = b.minZ + (float)(size * U(rng) * (U(rng) < 0.2 ? 0.0 : 1.0));
    level = (int)(rng() % 3);
}

int main() {
    std::mt19937_64 rng(0x5C0B0);
    Stats plain, adv, planted;
    float T[16];
    scr_raster_batch b;
    int level, W, H;
    for (int t = 0; t < 6000; t++) {
        scenario(rng, false, T, b, level, W, H);
        check_case(T, b, level, W, H, rng, 400, plain, false);
    }
    std::mt19937_64 rngA(0xADD5);
    for (int t = 0; t < 6000; t++) {
        scenario(rngA, true, T, b, level, W, H);
        std::mt19937_64 r2 = rngA;
        check_case(T, b, level, W, H, rngA, 400, adv, false);
        check_case(T, b, level, W, H, r2, 400, planted, true);
    }
    std::printf("plain: %ld cases (%ld proven, %ld inside, %ld fast rcp), %ld points, %ld fragments, %ld violations\n",
                plain.cases, plain.proven, plain.inside, plain.fast, plain.points, plain.frags, plain.violations);
    std::printf("adversarial: %ld cases (%ld proven, %ld inside, %ld fast rcp), %ld points, %ld fragments, %ld "
                "violations\n",
                adv.cases, adv.proven, adv.inside, adv.fast, adv.points, adv.frags, adv.violations);
    std::printf("without the allowance: %ld violations over %ld points\n", planted.violations, planted.points);
    const bool ok = plain.violations == 0 && adv.violations == 0 && plain.proven > plain.cases / 2 &&
                    plain.inside > 100 && adv.proven > 100 && planted.violations > 0;
    std::printf(ok ? "bounds hold\n" : "FAILED\n");
    return ok ? 0 : 1;
}
