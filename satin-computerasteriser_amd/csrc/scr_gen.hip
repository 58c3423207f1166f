// scr_gen.hip — deterministic synthetic LiDAR-like scene, generated on the GPU
// (benchmark and test input; the reference has no generator).
//
// Point i draws its random numbers from a counter-based SplitMix64 stream
// keyed by (seed, i, k), so the scene is independent of launch shape.  The
// first round(terrainFraction * count) points lie on a multi-octave value-noise
// heightfield; the rest lie on the four sides and the top of numBoxes boxes,
// picked area-weighted through an integer CDF built on the host.
#include <cmath>
#include <cstring>
#include <vector>

#include "scr_internal.hpp"

namespace scr {
namespace {

__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint32_t rng_u32(uint64_t seed, uint64_t i, uint32_t k) {
    return (uint32_t)(splitmix64(seed + 0x9e3779b97f4a7c15ull * (i + 1) + 0xd6e8feb86659fd93ull * (k + 1)) >> 32);
}

__host__ __device__ inline float rng_unit(uint64_t seed, uint64_t i, uint32_t k) {
    return (float)(rng_u32(seed, i, k) >> 8) * (1.0f / 16777216.0f);
}

__host__ __device__ inline float lattice_value(int ix, int iy, int octave, uint64_t seed) {
    const uint64_t h = splitmix64(seed ^ ((uint64_t)(uint32_t)ix * 0x8cb92ba72f3d8dd7ull) ^
                                  ((uint64_t)(uint32_t)iy * 0xa0761d6478bd642full) ^
                                  ((uint64_t)(uint32_t)octave * 0xe7037ed1a0b428dbull));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

__host__ __device__ inline float value_noise(float x, float y, int octave, uint64_t seed) {
    const float fx0 = floorf(x), fy0 = floorf(y);
    const int ix = (int)fx0, iy = (int)fy0;
    const float fx = x - fx0, fy = y - fy0;
    const float sx = fx * fx * (3.0f - 2.0f * fx);
    const float sy = fy * fy * (3.0f - 2.0f * fy);
    const float a = lattice_value(ix, iy, octave, seed), b = lattice_value(ix + 1, iy, octave, seed);
    const float c = lattice_value(ix, iy + 1, octave, seed), d = lattice_value(ix + 1, iy + 1, octave, seed);
    const float ab = a + (b - a) * sx;
    const float cd = c + (d - c) * sx;
    return ab + (cd - ab) * sy;
}

struct GenParams {
    uint64_t seed;
    float half, amplitude, invWavelength, boxMax, colorNoise;
    int octaves, quantize;
    uint64_t nTerrain, count;
    uint32_t nFaces;
};

__host__ __device__ inline float terrain_height(const GenParams& g, float x, float y) {
    float h = 0.0f, amp = g.amplitude, freq = g.invWavelength;
    for (int o = 0; o < g.octaves; o++) {
        h = h + amp * (2.0f * value_noise(x * freq, y * freq, o, g.seed) - 1.0f);
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    return h;
}

// Box face: origin + s*u + t*v (u, v edge vectors).
struct Face {
    float ox, oy, oz;
    float ux, uy, uz;
    float vx, vy, vz;
};

__device__ inline void colormap(const GenParams& g, float z, uint64_t i, float4& out) {
    const float lo = -g.amplitude, hi = g.amplitude + g.boxMax;
    const float t = fminf(fmaxf((z - lo) / (hi - lo), 0.0f), 1.0f);
    // stops: deep blue-green → green → brown → grey → white
    const float st[5] = {0.0f, 0.3f, 0.55f, 0.8f, 1.0f};
    const float cr[5] = {0.10f, 0.22f, 0.55f, 0.75f, 1.00f};
    const float cg[5] = {0.30f, 0.55f, 0.42f, 0.75f, 1.00f};
    const float cb[5] = {0.45f, 0.20f, 0.25f, 0.78f, 1.00f};
    int k = 0;
    while (k < 3 && t > st[k + 1]) k++;
    const float f = (t - st[k]) / (st[k + 1] - st[k]);
    float c[3] = {cr[k] + (cr[k + 1] - cr[k]) * f, cg[k] + (cg[k + 1] - cg[k]) * f, cb[k] + (cb[k + 1] - cb[k]) * f};
    for (int a = 0; a < 3; a++) {
        const float noise = (2.0f * rng_unit(g.seed, i, 10 + a) - 1.0f) * g.colorNoise;
        if (g.quantize) {
            // RGB8 value v stored as (v + 0.5) / 255 so the packer's truncating
            // uint(c * 255) recovers v exactly
            float v = rintf(c[a] * 255.0f + noise);
            v = fminf(fmaxf(v, 0.0f), 255.0f);
            c[a] = (v + 0.5f) / 255.0f;
        } else {
            c[a] = fminf(fmaxf(c[a] + noise / 255.0f, 0.0f), 1.0f);
        }
    }
    out = make_float4(c[0], c[1], c[2], 1.0f);
}

__global__ __launch_bounds__(256) void k_gen(GenParams g, const Face* __restrict__ faces,
                                             const uint32_t* __restrict__ cdf, float4* __restrict__ pos,
                                             float4* __restrict__ col) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= g.count) return;
    float x, y, z;
    if (i < g.nTerrain || g.nFaces == 0) {
        x = (2.0f * rng_unit(g.seed, i, 0) - 1.0f) * g.half;
        y = (2.0f * rng_unit(g.seed, i, 1) - 1.0f) * g.half;
        z = terrain_height(g, x, y);
    } else {
        const uint32_t r = rng_u32(g.seed, i, 2);
        uint32_t lo = 0, hi = g.nFaces - 1;
        while (lo < hi) {  // first face whose cumulative ticket exceeds r
            const uint32_t mid = (lo + hi) >> 1;
            if (cdf[mid] > r) hi = mid;
            else lo = mid + 1;
        }
        const Face f = faces[lo];
        const float s = rng_unit(g.seed, i, 3), t = rng_unit(g.seed, i, 4);
        x = f.ox + s * f.ux + t * f.vx;
        y = f.oy + s * f.uy + t * f.vy;
        z = f.oz + s * f.uz + t * f.vz;
    }
    pos[i] = make_float4(x, y, z, 0.0f);
    float4 c;
    colormap(g, z, i, c);
    col[i] = c;
}

}  // namespace
}  // namespace scr

extern "C" {

void scr_scene_params_default(scr_scene_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->seed = 0x5C0002ull;
    p->halfExtent = 1000.0f;
    p->numBoxes = 2000;
    p->terrainFraction = 0.7f;
    p->amplitude = 120.0f;
    p->baseWavelength = 500.0f;
    p->octaves = 6;
    p->boxMinEdge = 10.0f;
    p->boxMaxEdge = 60.0f;
    p->colorNoise = 10.0f;
    p->quantizeColors = 1;
}

namespace {

scr::GenParams gen_params(const scr_scene_params* prm, size_t count) {
    scr::GenParams g;
    g.seed = prm->seed;
    g.half = prm->halfExtent;
    g.amplitude = prm->amplitude;
    g.invWavelength = 1.0f / prm->baseWavelength;
    g.boxMax = prm->boxMaxEdge;
    g.colorNoise = prm->colorNoise;
    g.octaves = prm->octaves;
    g.quantize = prm->quantizeColors;
    g.count = count;
    g.nTerrain = prm->numBoxes > 0 ? (uint64_t)std::llround((double)prm->terrainFraction * (double)count) : count;
    g.nFaces = 0;
    return g;
}

}  // namespace

float scr_scene_height(const scr_scene_params* prm, float x, float y) {
    if (!prm || prm->baseWavelength <= 0.0f || prm->octaves < 0 || prm->octaves > 16) return 0.0f;
    return scr::terrain_height(gen_params(prm, 0), x, y);
}

int scr_gen_scene(scr_ctx* ctx, const scr_scene_params* prm, size_t count, float* positions4, float* colors4,
                  int outputOnDevice) {
    using namespace scr;
    if (!ctx || !prm || (count && (!positions4 || !colors4))) return SCR_ERR_INVALID_ARGUMENT;
    if (prm->halfExtent <= 0.0f || prm->octaves < 0 || prm->octaves > 16 || prm->numBoxes < 0 ||
        !(prm->terrainFraction >= 0.0f && prm->terrainFraction <= 1.0f) || prm->boxMaxEdge < prm->boxMinEdge ||
        prm->baseWavelength <= 0.0f)
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "bad scene parameters");
    if (count == 0) return SCR_OK;
    GenParams g = gen_params(prm, count);
    // boxes and their faces (host, deterministic in the seed)
    std::vector<Face> faces;
    std::vector<double> area;
    const uint64_t boxSeed = prm->seed ^ 0xb0b0b0b0b0b0b0b0ull;
    for (int b = 0; b < prm->numBoxes; b++) {
        const float cx = (2.0f * rng_unit(boxSeed, (uint64_t)b, 0) - 1.0f) * prm->halfExtent;
        const float cy = (2.0f * rng_unit(boxSeed, (uint64_t)b, 1) - 1.0f) * prm->halfExtent;
        const float span = prm->boxMaxEdge - prm->boxMinEdge;
        const float sx = prm->boxMinEdge + span * rng_unit(boxSeed, (uint64_t)b, 2);
        const float sy = prm->boxMinEdge + span * rng_unit(boxSeed, (uint64_t)b, 3);
        const float sz = prm->boxMinEdge + span * rng_unit(boxSeed, (uint64_t)b, 4);
        const float base = terrain_height(g, cx, cy);
        const float x0 = cx - 0.5f * sx, y0 = cy - 0.5f * sy;
        faces.push_back({x0, y0, base, 0, sy, 0, 0, 0, sz});        // x = x0
        faces.push_back({x0 + sx, y0, base, 0, sy, 0, 0, 0, sz});   // x = x1
        faces.push_back({x0, y0, base, sx, 0, 0, 0, 0, sz});        // y = y0
        faces.push_back({x0, y0 + sy, base, sx, 0, 0, 0, 0, sz});   // y = y1
        faces.push_back({x0, y0, base + sz, sx, 0, 0, 0, sy, 0});   // top
        area.push_back((double)sy * sz);
        area.push_back((double)sy * sz);
        area.push_back((double)sx * sz);
        area.push_back((double)sx * sz);
        area.push_back((double)sx * sy);
    }
    std::vector<uint32_t> cdf(faces.size());
    double total = 0.0;
    for (double a : area) total += a;
    double acc = 0.0;
    for (size_t k = 0; k < faces.size(); k++) {
        acc += area[k];
        const double t = acc / total * 4294967296.0;
        cdf[k] = k + 1 == faces.size() ? 0xffffffffu : (uint32_t)std::fmin(t, 4294967295.0);
    }
    g.nFaces = (uint32_t)faces.size();
    hipStream_t s = ctx->stream;
    Face* dFaces = nullptr;
    uint32_t* dCdf = nullptr;
    float4* dPos = outputOnDevice ? (float4*)positions4 : nullptr;
    float4* dCol = outputOnDevice ? (float4*)colors4 : nullptr;
    hipError_t e = hipSuccess;
    if (!faces.empty()) {
        e = hipMalloc((void**)&dFaces, faces.size() * sizeof(Face));
        if (e == hipSuccess) e = hipMalloc((void**)&dCdf, cdf.size() * 4);
        if (e == hipSuccess) e = hipMemcpy(dFaces, faces.data(), faces.size() * sizeof(Face), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dCdf, cdf.data(), cdf.size() * 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !outputOnDevice) e = hipMalloc((void**)&dPos, count * sizeof(float4));
    if (e == hipSuccess && !outputOnDevice) e = hipMalloc((void**)&dCol, count * sizeof(float4));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_gen, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, g, dFaces, dCdf, dPos, dCol);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && !outputOnDevice) e = hipMemcpy(positions4, dPos, count * sizeof(float4), hipMemcpyDeviceToHost);
    if (e == hipSuccess && !outputOnDevice) e = hipMemcpy(colors4, dCol, count * sizeof(float4), hipMemcpyDeviceToHost);
    if (dFaces) (void)hipFree(dFaces);
    if (dCdf) (void)hipFree(dCdf);
    if (!outputOnDevice) {
        if (dPos) (void)hipFree(dPos);
        if (dCol) (void)hipFree(dCol);
    }
    if (e != hipSuccess) return hip_error(ctx, e, "scr_gen_scene");
    return SCR_OK;
}

}  // extern "C"
