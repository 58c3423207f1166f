// Host-side property check of the per-batch footprint-radius bound: whenever
// batch_const_radius() names a radius for a batch, every point that decodes
// inside that batch must compute exactly that radius with footprint_radius().
// Built by tests/test_native_checks.py with hipcc for the host (no GPU use).
#include <cmath>
#include <cstdio>
#include <random>

#include "scr_math.hpp"

using scr::PassParams;

static void look_at(const float eye[3], const float tgt[3], const float up[3], float* m) {
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

int main() {
    std::mt19937_64 rng(0x5C0001);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long hoisted = 0, trials = 0, points = 0;
    for (int t = 0; t < 4000; t++) {
        PassParams p{};
        const float eye[3] = {(float)(U(rng) * 8000 - 4000), (float)(U(rng) * 8000 - 4000), (float)(U(rng) * 500)};
        const float tgt[3] = {(float)(U(rng) * 2000 - 1000), (float)(U(rng) * 2000 - 1000), 0.0f};
        const float up[3] = {0.0f, 0.0f, 1.0f};
        look_at(eye, tgt, up, p.view);
        const float fov = (float)(20 + U(rng) * 80) * 3.14159265f / 180.0f;
        const int H = 256 + (int)(U(rng) * 1200);
        const float f = 1.0f / std::tan(fov * 0.5f);
        for (int i = 0; i < 16; i++) p.proj[i] = 0.0f;
        p.proj[0] = f / 1.5f;
        p.proj[5] = f;
        p.proj[11] = -1.0f;
        p.proj[14] = 0.1f;
        p.width = (int)(H * 1.5);
        p.height = H;
        p.pointSizeMode = U(rng) < 0.5 ? 0 : 1;
        const float mn = (float)(1.0 + U(rng) * 2.0), mx = mn + (float)(U(rng) * 6.0);
        p.lo = std::fmax(std::fmin(mn, mx), 1.0f);
        p.hi = std::fmax(std::fmax(mn, mx), p.lo);
        p.pointSizeScale = p.pointSizeMode == 0 ? (float)(10 + U(rng) * 6000) : (float)(0.01 + U(rng) * 3.0);
        p.constRadius = -1;
        p.ortho = 0;
        p.focal = ((float)H * 0.5f) * p.proj[5];
        float world[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        if (U(rng) < 0.3) {   // scaled, translated model matrix
            const float s = (float)(0.5 + U(rng) * 2.0);
            world[0] = world[5] = world[10] = s;
            world[12] = (float)(U(rng) * 200 - 100);
            world[13] = (float)(U(rng) * 200 - 100);
        }
        scr_raster_batch b{};
        const float cx = (float)(U(rng) * 8000 - 4000), cy = (float)(U(rng) * 8000 - 4000), cz = (float)(U(rng) * 300);
        const float sx = (float)(0.5 + U(rng) * 300), sy = (float)(0.5 + U(rng) * 300), sz = (float)(U(rng) * 80);
        b.minX = cx; b.minY = cy; b.minZ = cz;
        b.maxX = cx + sx; b.maxY = cy + sy; b.maxZ = cz + sz;
        if (t % 2 == 1) {
            // adversarial: put a radius step (size 2k+1) within 1e-3 relative of
            // the box's nearest view distance, so the bound's margins decide
            double lo3[3], hi3[3];
            for (int i = 0; i < 3; i++) {
                double l = p.view[12 + i], h = p.view[12 + i];
                const double bmn[3] = {b.minX * (double)world[0] + world[12], b.minY * (double)world[5] + world[13],
                                       b.minZ * (double)world[10] + world[14]};
                const double bmx[3] = {b.maxX * (double)world[0] + world[12], b.maxY * (double)world[5] + world[13],
                                       b.maxZ * (double)world[10] + world[14]};
                for (int j = 0; j < 3; j++) {
                    const double a = p.view[j * 4 + i] * bmn[j], c = p.view[j * 4 + i] * bmx[j];
                    l += std::fmin(a, c);
                    h += std::fmax(a, c);
                }
                lo3[i] = l;
                hi3[i] = h;
            }
            double dnear;
            if (p.pointSizeMode == 1) {
                dnear = std::fmax(-hi3[2], 1e-6);
            } else {
                double n2 = 0.0;
                for (int i = 0; i < 3; i++) {
                    const double nn = std::fmax(0.0, std::fmax(lo3[i], -hi3[i]));
                    n2 += nn * nn;
                }
                dnear = std::fmax(std::sqrt(n2), 1e-6);
            }
            const int k = 1 + (int)(U(rng) * 2);
            const double delta = (U(rng) * 2.0 - 1.0) * 1e-3;
            const double target = (2 * k + 1) * (1.0 - delta);   // size at the nearest point
            p.lo = 1.0f;
            p.hi = 8.0f;
            if (p.pointSizeMode == 1) p.pointSizeScale = (float)(target * dnear / (2.0 * p.focal));
            else p.pointSizeScale = (float)(target * dnear);
        }
        trials++;
        const int r = scr::batch_const_radius(p, world, b);
        if (r < 0) continue;
        hoisted++;
        // decode points exactly as the draw front end does, at every precision level
        for (int level = 0; level < 3; level++) {
            const float inv = level == 0 ? (1.0f / 1073741824.0f) : (level == 1 ? (1.0f / 1048576.0f) : (1.0f / 1024.0f));
            const uint32_t qmax = level == 0 ? 0x3fffffffu : (level == 1 ? 0xfffffu : 1023u);
            const float ssx = std::fmax(b.maxX - b.minX, 0.000001f) * inv;
            const float ssy = std::fmax(b.maxY - b.minY, 0.000001f) * inv;
            const float ssz = std::fmax(b.maxZ - b.minZ, 0.000001f) * inv;
            for (int k = 0; k < 200; k++) {
                uint32_t q[3];
                for (int a = 0; a < 3; a++) {
                    const double u = U(rng);
                    q[a] = k < 8 ? (((k >> a) & 1) ? qmax : 0u) : (uint32_t)(u * (double)qmax);
                }
                const float x = (float)q[0] * ssx + b.minX;
                const float y = (float)q[1] * ssy + b.minY;
                const float z = (float)q[2] * ssz + b.minZ;
                const int rp = scr::footprint_radius(p, world, x, y, z);
                points++;
                if (rp != r) {
                    std::printf("VIOLATION trial %d level %d: batch radius %d, point radius %d\n", t, level, r, rp);
                    return 1;
                }
            }
        }
    }
    std::printf("trials %ld hoisted %ld points %ld\n", trials, hoisted, points);
    return hoisted > trials / 10 ? 0 : 2;
}
