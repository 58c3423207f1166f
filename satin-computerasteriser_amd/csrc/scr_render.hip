// scr_render.hip — per-frame render passes for gfx950.
//
//   k_cull            cullUpdate + cullFinalizeUpdate   Cull/Shaders.metal:19-83, CullFinalize :3-14
//   k_depth           depthPassUpdate                   DepthPass/Shaders.metal:21-116
//   k_color           colorPassUpdate (opaque path)     ColorPass/Shaders.metal:30-219
//   k_resolve         resolveUpdate (count average)     Resolve/Shaders.metal:13-57
//   k_nearest         nearestDepthUpdate+nearestIndexUpdate in one 64-bit
//                     atomic max of depth<<32 | ~pointIndex
//                                                       NearestDepth :14-91, NearestIndex :14-94
//   k_nearest_resolve nearestResolveUpdate              NearestResolve/Shaders.metal:8-38
//
// HIP has no indirect dispatch: the draw kernels are persistent grids sized to
// the machine; each workgroup walks the visible list the cull kernel wrote
// (count read on the device), so no host round trip separates cull and draw.
#include "scr_internal.hpp"

namespace scr {
namespace {

constexpr int kDrawThreads = 256;

// ---------------------------------------------------------------- cull helpers

// makePlane + intersectsFrustum — Common.metal:119-153
__device__ bool intersects_frustum(const float* m, const scr_raster_batch& b) {
#define M(c, r) m[(c) * 4 + (r)]
    const float pl[6][4] = {
        {M(0, 3) - M(0, 0), M(1, 3) - M(1, 0), M(2, 3) - M(2, 0), M(3, 3) - M(3, 0)},
        {M(0, 3) + M(0, 0), M(1, 3) + M(1, 0), M(2, 3) + M(2, 0), M(3, 3) + M(3, 0)},
        {M(0, 3) + M(0, 1), M(1, 3) + M(1, 1), M(2, 3) + M(2, 1), M(3, 3) + M(3, 1)},
        {M(0, 3) - M(0, 1), M(1, 3) - M(1, 1), M(2, 3) - M(2, 1), M(3, 3) - M(3, 1)},
        {M(0, 3) - M(0, 2), M(1, 3) - M(1, 2), M(2, 3) - M(2, 2), M(3, 3) - M(3, 2)},
        {M(0, 2), M(1, 2), M(2, 2), M(3, 2)},
    };
#undef M
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const float len = fmaxf(len3(pl[i][0], pl[i][1], pl[i][2]), 0.000001f);
        const float r = rcp(len);
        const float nx = pl[i][0] * r, ny = pl[i][1] * r, nz = pl[i][2] * r;
        const float off = pl[i][3] * r;
        const float px = nx > 0.0f ? b.maxX : b.minX;
        const float py = ny > 0.0f ? b.maxY : b.minY;
        const float pz = nz > 0.0f ? b.maxZ : b.minZ;
        const float d = (nx * px + ny * py) + nz * pz;
        if (d + off < 0.0f) return false;
    }
    return true;
}

// pixelSizeOnScreen — Common.metal:155-172
__device__ float pixel_size_on_screen(const scr_raster_batch& b, const float* world, const PassParams& p) {
    const float cx = (b.minX + b.maxX) * 0.5f;
    const float cy = (b.minY + b.maxY) * 0.5f;
    const float cz = (b.minZ + b.maxZ) * 0.5f;
    const float radius = len3(b.minX - b.maxX, b.minY - b.maxY, b.minZ - b.maxZ);
    float vw[16];
    mat_mul(p.view, world, vw);
    const f4 vc = mul(vw, f4{cx, cy, cz, 1.0f});
    const f4 ve{vc.x + radius, vc.y + 0.0f, vc.z + 0.0f, vc.w + 0.0f};
    f4 pc = mul(p.proj, vc);
    f4 pe = mul(p.proj, ve);
    const float rc = rcp(fmaxf(fabsf(pc.w), 0.000001f));
    const float re = rcp(fmaxf(fabsf(pe.w), 0.000001f));
    pc.x = pc.x * rc;
    pc.y = pc.y * rc;
    pe.x = pe.x * re;
    pe.y = pe.y * re;
    const float sw = (float)p.width, sh = (float)p.height;
    const float scx = (sw * (pc.x + 1.0f)) * 0.5f, scy = (sh * (pc.y + 1.0f)) * 0.5f;
    const float sex = (sw * (pe.x + 1.0f)) * 0.5f, sey = (sh * (pe.y + 1.0f)) * 0.5f;
    const float dx = sex - scx, dy = sey - scy;
    return sqrtf(dx * dx + dy * dy);
}

__global__ __launch_bounds__(256) void k_cull(CloudDev c, PassParams p) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= c.nSlots) return;
    const scr_raster_batch b = c.batches[gid];
    if (b.state == 0) return;  // residency gate, Cull/Shaders.metal:34-36
    if ((int)b.fileIndex >= c.nFiles) return;
    const scr_raster_file* f = c.files + b.fileIndex;
    if (p.enableFrustumCulling != 0 && !intersects_frustum(f->transformFrustum, b)) return;
    const float px = pixel_size_on_screen(b, f->world, p);
    int level;
    if (px < 100.0f) level = 4;
    else if (px < 200.0f) level = 3;
    else if (px < 500.0f) level = 2;
    else if (px < 10000.0f) level = 1;
    else level = 0;
    const float thr = (p.enableCLOD != 0)
                          ? fmaxf(0.0f, (scr_log2f(fmaxf(px, 1.0f) * (1.0f / 25.0f)) + 3.0f) + (float)p.lodBias)
                          : 99.0f;
    uint32_t active = b.numPoints;
    if (b.lodCum67 != 0u) {
        int lmax = f2i(floorf(thr + 0.5f));
        lmax = lmax < 0 ? 0 : (lmax > 7 ? 7 : lmax);
        const uint32_t w = lmax < 2 ? b.lodCum01 : (lmax < 4 ? b.lodCum23 : (lmax < 6 ? b.lodCum45 : b.lodCum67));
        active = (lmax & 1) ? (w >> 16) : (w & 0xffffu);
    }
    if (active == 0u) return;
    const uint32_t slot = atomicAdd(c.counter, 1u);
    scr_visible_batch vb;
    vb.batchIndex = (uint32_t)gid;
    vb.level = level;
    vb.lodThreshold = thr;
    vb.activePoints = active;
    c.visible[slot] = vb;
}

// ---------------------------------------------------------------- point front end

struct BatchCtx {
    float mnx, mny, mnz;
    float sx, sy, sz;   // batch size * 2^-bits
    int level;
    float lodThreshold;
    bool keepAll;       // no point of the batch can fail the LOD/dither test
    uint32_t first, active, batchIndex;
    float T[16];        // file.transform (held in registers: the atomics below
                        // would otherwise force reloads through memory)
    const float* W;     // file.world (only read for per-point sizes)
};

__device__ inline BatchCtx batch_ctx(const CloudDev& c, uint32_t v) {
    const scr_visible_batch vb = c.visible[v];
    const scr_raster_batch b = c.batches[vb.batchIndex];
    BatchCtx bc;
    bc.level = vb.level;
    bc.lodThreshold = vb.lodThreshold;
    // dither is in [0, 1] and (thr - L) + 0.5 decreases with L <= 7
    bc.keepAll = ((vb.lodThreshold - 7.0f) + 0.5f) > 1.0f;
    bc.active = vb.activePoints;
    bc.batchIndex = vb.batchIndex;
    bc.first = b.firstPoint;
    const float inv = vb.level == 0 ? (1.0f / 1073741824.0f) : (vb.level == 1 ? (1.0f / 1048576.0f) : (1.0f / 1024.0f));
    bc.mnx = b.minX;
    bc.mny = b.minY;
    bc.mnz = b.minZ;
    bc.sx = fmaxf(b.maxX - b.minX, 0.000001f) * inv;
    bc.sy = fmaxf(b.maxY - b.minY, 0.000001f) * inv;
    bc.sz = fmaxf(b.maxZ - b.minZ, 0.000001f) * inv;
    const scr_raster_file* f = c.files + b.fileIndex;
#pragma unroll
    for (int i = 0; i < 16; i++) bc.T[i] = f->transform[i];
    bc.W = f->world;
    return bc;
}

// ---------------------------------------------------------------- chunked draw
//
// Each visible batch is processed in chunks of kChunk points: every thread
// owns kPPT consecutive points (vector loads of the position words), keeps
// their fragments in registers, and the block reduces the chunk's actual
// screen rectangle.  When that rectangle is small next to the fragment count,
// the chunk's fragments meet in an LDS tile (LDS atomics) and the global
// atomics see one update per touched pixel; otherwise they go straight to the
// global planes.  Integer max / add are order independent, so both routes
// give the same planes.

constexpr int kPPT = 8;
constexpr int kChunk = kDrawThreads * kPPT;

struct ChunkFrags {
    int px[kPPT];      // -1: no fragment
    int py[kPPT];      // y-flipped row
    int r[kPPT];       // footprint radius
    float z[kPPT];     // ndc z
};

// Loads the position words (and levels) of the thread's kPPT points and runs
// the front end of every point pass on each of them.
template <bool kNearestZ, bool kR0>
__device__ inline void chunk_front(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t li0,
                                   bool vec, ChunkFrags& f) {
    uint32_t lo[kPPT], me[kPPT], hi[kPPT], lv[kPPT];
    const uint32_t base = bc.first + li0;
    const uint32_t n = li0 < bc.active ? min((uint32_t)kPPT, bc.active - li0) : 0u;
    if (vec && n == (uint32_t)kPPT) {
        // base is a multiple of kPPT: 32-byte aligned words, 8-byte aligned levels
        const uint4* pl = reinterpret_cast<const uint4*>(c.xyzLow + base);
        const uint4 a = pl[0], b = pl[1];
        lo[0] = a.x; lo[1] = a.y; lo[2] = a.z; lo[3] = a.w; lo[4] = b.x; lo[5] = b.y; lo[6] = b.z; lo[7] = b.w;
        if (bc.level <= 1) {
            const uint4* pm = reinterpret_cast<const uint4*>(c.xyzMed + base);
            const uint4 ma = pm[0], mb = pm[1];
            me[0] = ma.x; me[1] = ma.y; me[2] = ma.z; me[3] = ma.w; me[4] = mb.x; me[5] = mb.y; me[6] = mb.z; me[7] = mb.w;
        }
        if (bc.level == 0) {
            const uint4* ph = reinterpret_cast<const uint4*>(c.xyzHigh + base);
            const uint4 ha = ph[0], hb = ph[1];
            hi[0] = ha.x; hi[1] = ha.y; hi[2] = ha.z; hi[3] = ha.w; hi[4] = hb.x; hi[5] = hb.y; hi[6] = hb.z; hi[7] = hb.w;
        }
        if (!bc.keepAll) {
            const uint2 l8 = *reinterpret_cast<const uint2*>(c.levels + base);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                lv[j] = (l8.x >> (8 * j)) & 0xffu;
                lv[4 + j] = (l8.y >> (8 * j)) & 0xffu;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < kPPT; j++) {
            const uint32_t pi = base + (uint32_t)j;
            const bool ok = (uint32_t)j < n;
            lo[j] = ok ? c.xyzLow[pi] : 0u;
            me[j] = (ok && bc.level <= 1) ? c.xyzMed[pi] : 0u;
            hi[j] = (ok && bc.level == 0) ? c.xyzHigh[pi] : 0u;
            lv[j] = (ok && !bc.keepAll) ? (uint32_t)c.levels[pi] : 0u;
        }
    }
    // Branch-free per point: every test of the reference folds into `ok`
    // (each written as the negation of its reject condition, so NaNs behave
    // as in the Metal source) and the fragment is kept or dropped once.
#pragma unroll
    for (int j = 0; j < kPPT; j++) {
        const uint32_t pi = base + (uint32_t)j;
        bool ok = (uint32_t)j < n;
        if (!bc.keepAll) {
            const float pointLevel = (float)(lv[j] & 7u);
            const float dither = (p.lodDither != 0) ? hash_unit(pi) : 0.5f;
            ok = ok && !(dither >= (bc.lodThreshold - pointLevel) + 0.5f);
        }
        uint32_t qx, qy, qz;
        const uint32_t l = lo[j];
        if (bc.level == 0) {
            const uint32_t m = me[j], h = hi[j];
            qx = ((l & 1023u) << 20) | ((m & 1023u) << 10) | (h & 1023u);
            qy = (((l >> 10) & 1023u) << 20) | (((m >> 10) & 1023u) << 10) | ((h >> 10) & 1023u);
            qz = (((l >> 20) & 1023u) << 20) | (((m >> 20) & 1023u) << 10) | ((h >> 20) & 1023u);
        } else if (bc.level == 1) {
            const uint32_t m = me[j];
            qx = ((l & 1023u) << 10) | (m & 1023u);
            qy = (((l >> 10) & 1023u) << 10) | ((m >> 10) & 1023u);
            qz = (((l >> 20) & 1023u) << 10) | ((m >> 20) & 1023u);
        } else {
            qx = l & 1023u;
            qy = (l >> 10) & 1023u;
            qz = (l >> 20) & 1023u;
        }
        const float x = (float)qx * bc.sx + bc.mnx;
        const float y = (float)qy * bc.sy + bc.mny;
        const float z = (float)qz * bc.sz + bc.mnz;
        const f4 clip = mul_point(bc.T, x, y, z);
        ok = ok && !(clip.w <= 0.0f);
        const float rw = rcp(clip.w);
        const float nx = clip.x * rw, ny = clip.y * rw, nz = clip.z * rw;
        ok = ok && !(nx < -1.0f || ny < -1.0f || nx > 1.0f || ny > 1.0f);
        if (kNearestZ) ok = ok && !(nz <= 0.0f || nz > 1.0f);
        else ok = ok && !(nz < 0.0f || nz > 1.0f);
        const int ix = f2i_hw((nx * 0.5f + 0.5f) * (float)p.width);
        const int iy = f2i_hw((ny * 0.5f + 0.5f) * (float)p.height);
        ok = ok && !(ix < 0 || ix >= p.width || iy < 0 || iy >= p.height);
        f.px[j] = ok ? ix : -1;
        f.py[j] = p.height - 1 - iy;
        f.z[j] = nz;
        f.r[j] = (kR0 || !ok) ? 0 : footprint_radius(p, bc.W, x, y, z);
    }
}

// Screen rectangle of the chunk's fragments (footprints included, clamped to
// the screen) and the number of fragment centres.
struct ChunkBox {
    int x0, y0, w, h, frags;
    uint32_t magic;   // i / w == umulhi(i, magic) for i < 2^16
    bool lds;
    // every on-screen footprint pixel of the chunk lies inside the box (it is
    // the clamped union of the footprints), so no bounds test is needed
    __device__ inline int local(int tx, int ty) const { return (ty - y0) * w + (tx - x0); }
    __device__ inline size_t pixel(int k, int width) const {
        const int row = (int)__umulhi((uint32_t)k, magic);
        const int col = k - row * w;
        return (size_t)(y0 + row) * (size_t)width + (size_t)(x0 + col);
    }
};

__device__ inline int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ inline int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
__device__ inline int wave_sum(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block-wide reduction; red must hold 5 * (kDrawThreads / 64) ints.  Ends
// with a barrier, so the caller may reuse LDS right after.
__device__ inline ChunkBox chunk_box(const ChunkFrags& f, const PassParams& p, int cap, int* red) {
    int x0 = 0x7fffffff, y0 = 0x7fffffff, x1 = -1, y1 = -1, n = 0;
#pragma unroll
    for (int j = 0; j < kPPT; j++) {
        if (f.px[j] < 0) continue;
        x0 = min(x0, f.px[j] - f.r[j]);
        x1 = max(x1, f.px[j] + f.r[j]);
        y0 = min(y0, f.py[j] - f.r[j]);
        y1 = max(y1, f.py[j] + f.r[j]);
        n++;
    }
    x0 = wave_min(x0);
    y0 = wave_min(y0);
    x1 = wave_max(x1);
    y1 = wave_max(y1);
    n = wave_sum(n);
    const int w = threadIdx.x >> 6;
    constexpr int kW = kDrawThreads / 64;
    if (__lane_id() == 0) {
        red[w] = x0;
        red[kW + w] = y0;
        red[2 * kW + w] = x1;
        red[3 * kW + w] = y1;
        red[4 * kW + w] = n;
    }
    __syncthreads();
    x0 = red[0];
    y0 = red[kW];
    x1 = red[2 * kW];
    y1 = red[3 * kW];
    n = red[4 * kW];
#pragma unroll
    for (int k = 1; k < kW; k++) {
        x0 = min(x0, red[k]);
        y0 = min(y0, red[kW + k]);
        x1 = max(x1, red[2 * kW + k]);
        y1 = max(y1, red[3 * kW + k]);
        n += red[4 * kW + k];
    }
    __syncthreads();
    ChunkBox b;
    b.frags = n;
    b.x0 = max(x0, 0);
    b.y0 = max(y0, 0);
    x1 = min(x1, p.width - 1);
    y1 = min(y1, p.height - 1);
    b.w = x1 >= b.x0 ? x1 - b.x0 + 1 : 0;
    b.h = y1 >= b.y0 ? y1 - b.y0 + 1 : 0;
    const long long area = (long long)b.w * b.h;
    // LDS pays when the tile is small next to the fragments it absorbs
    b.lds = n > 0 && area > 0 && area <= cap && area <= 8LL * n;
    b.magic = b.lds ? (uint32_t)(0xffffffffu / (uint32_t)b.w) + 1u : 0u;
    return b;
}

__device__ inline bool vec_ok(const BatchCtx& bc) { return (bc.first & (uint32_t)(kPPT - 1)) == 0u; }

// ---------------------------------------------------------------- depth pass

constexpr int kDepthTile = 4096;  // u32 per pixel: 16 KB

template <bool kR0>
__global__ __launch_bounds__(kDrawThreads) void k_depth(CloudDev c, PassParams p, uint32_t* __restrict__ depth) {
    __shared__ uint32_t lds[kDepthTile];
    __shared__ int red[5 * (kDrawThreads / 64)];
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, v);
        const bool vec = vec_ok(bc);
        for (uint32_t cb = 0; cb < bc.active; cb += kChunk) {
            ChunkFrags f;
            chunk_front<false, kR0>(c, p, bc, cb + threadIdx.x * kPPT, vec, f);
            const ChunkBox box = chunk_box(f, p, kDepthTile, red);
            if (box.frags == 0) continue;
            const int area = box.w * box.h;
            if (box.lds) {
                for (int i = threadIdx.x; i < area; i += kDrawThreads) lds[i] = 0u;
                __syncthreads();
            }
#pragma unroll
            for (int j = 0; j < kPPT; j++) {
                if (f.px[j] < 0) continue;
                const uint32_t d = depth_to_uint_hw(f.z[j]);
                if (kR0) {
                    if (box.lds) atomicMax(lds + box.local(f.px[j], f.py[j]), d);
                    else atomicMax(depth + (size_t)f.py[j] * p.width + f.px[j], d);
                    continue;
                }
                const int r = f.r[j];
                for (int oy = -r; oy <= r; oy++)
                    for (int ox = -r; ox <= r; ox++) {
                        if (!inside_footprint(ox, oy, r)) continue;
                        const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                        if (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height) continue;
                        if (box.lds) atomicMax(lds + box.local(tx, ty), d);
                        else atomicMax(depth + (size_t)ty * p.width + tx, d);
                    }
            }
            if (box.lds) {
                __syncthreads();
                for (int i = threadIdx.x; i < area; i += kDrawThreads) {
                    const uint32_t d = lds[i];
                    if (d == 0u) continue;
                    uint32_t* g = depth + box.pixel(i, p.width);
                    // depth only grows during the pass: a stale read can only
                    // cause a redundant atomic, never a missed one
                    if (d > *g) atomicMax(g, d);
                }
                __syncthreads();
            }
        }
    }
}

// ---------------------------------------------------------------- colour pass

constexpr int kColorTile = 2048;  // depth u32 + (r | g << 32) u64 + (b | count << 32) u64: 40 KB

template <bool kR0>
__global__ __launch_bounds__(kDrawThreads) void k_color(CloudDev c, PassParams p, const uint32_t* __restrict__ depth,
                                                        uint32_t* red, uint32_t* green, uint32_t* blue,
                                                        uint32_t* count) {
    __shared__ uint32_t ldsDepth[kColorTile];
    __shared__ unsigned long long ldsRG[kColorTile];
    __shared__ unsigned long long ldsBC[kColorTile];
    __shared__ int redbuf[5 * (kDrawThreads / 64)];
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, v);
        const bool vec = vec_ok(bc);
        for (uint32_t cb = 0; cb < bc.active; cb += kChunk) {
            const uint32_t li0 = cb + threadIdx.x * kPPT;
            ChunkFrags f;
            chunk_front<false, kR0>(c, p, bc, li0, vec, f);
            // colour stage (ColorPass/Shaders.metal:95-133): packed r | g << 8 | b << 16
            uint32_t rgb[kPPT];
            const uint32_t base = bc.first + li0;
            if (vec && li0 + (uint32_t)kPPT <= bc.active) {
                const uint4* pc = reinterpret_cast<const uint4*>(c.colors + base);
                const uint4 ca = pc[0], cb2 = pc[1];
                rgb[0] = ca.x; rgb[1] = ca.y; rgb[2] = ca.z; rgb[3] = ca.w;
                rgb[4] = cb2.x; rgb[5] = cb2.y; rgb[6] = cb2.z; rgb[7] = cb2.w;
            } else {
#pragma unroll
                for (int j = 0; j < kPPT; j++) rgb[j] = f.px[j] >= 0 ? c.colors[base + (uint32_t)j] : 0u;
            }
#pragma unroll
            for (int j = 0; j < kPPT; j++) {
                const uint32_t pi = base + (uint32_t)j;
                uint32_t color = rgb[j];
                rgb[j] = 0u;
                if (f.px[j] < 0) continue;
                if (p.colorizeChunks != 0) color = bc.batchIndex * 1234567u;
                else if (p.colorizeOverdraw != 0) color = 0x00010101u;
                uint32_t r8 = color & 0xffu, g8 = (color >> 8) & 0xffu, b8 = (color >> 16) & 0xffu;
                if (p.applyTint != 0) {
                    const float4 t = c.tints ? c.tints[pi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    if (t.w < 0.0f) {  // discard sentinel, ColorPass/Shaders.metal:119
                        f.px[j] = -1;
                        continue;
                    }
                    const float w = saturate(t.w);
                    const float ox = (float)r8 * kInv255, oy = (float)g8 * kInv255, oz = (float)b8 * kInv255;
                    const float mx = ox + (saturate(t.x) - ox) * w;
                    const float my = oy + (saturate(t.y) - oy) * w;
                    const float mz = oz + (saturate(t.z) - oz) * w;
                    r8 = f2u(saturate(mx) * 255.0f + 0.5f);
                    g8 = f2u(saturate(my) * 255.0f + 0.5f);
                    b8 = f2u(saturate(mz) * 255.0f + 0.5f);
                }
                rgb[j] = r8 | (g8 << 8) | (b8 << 16);
            }
            const ChunkBox box = chunk_box(f, p, kColorTile, redbuf);
            if (box.frags == 0) continue;
            const int area = box.w * box.h;
            if (box.lds) {
                for (int i = threadIdx.x; i < area; i += kDrawThreads) {
                    ldsDepth[i] = depth[box.pixel(i, p.width)];
                    ldsRG[i] = 0ull;
                    ldsBC[i] = 0ull;
                }
                __syncthreads();
            }
#pragma unroll
            for (int j = 0; j < kPPT; j++) {
                if (f.px[j] < 0) continue;
                const uint32_t r8 = rgb[j] & 0xffu, g8 = (rgb[j] >> 8) & 0xffu, b8 = (rgb[j] >> 16) & 0xffu;
                const unsigned long long rg = (unsigned long long)r8 | ((unsigned long long)g8 << 32);
                const unsigned long long bn = (unsigned long long)b8 | (1ull << 32);
                const int rad = kR0 ? 0 : f.r[j];
                for (int oy = -rad; oy <= rad; oy++)
                    for (int ox = -rad; ox <= rad; ox++) {
                        if (!kR0 && !inside_footprint(ox, oy, rad)) continue;
                        const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                        if (!kR0 && (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height)) continue;
                        const size_t pix = (size_t)ty * p.width + tx;
                        const int k = box.lds ? box.local(tx, ty) : 0;
                        const uint32_t d = box.lds ? ldsDepth[k] : depth[pix];
                        if (d == 0u) continue;
                        if (!(f.z[j] >= uint_to_depth(d) * p.depthKeep)) continue;
                        if (box.lds) {
                            atomicAdd(ldsRG + k, rg);
                            atomicAdd(ldsBC + k, bn);
                        } else {
                            atomicAdd(red + pix, r8);
                            atomicAdd(green + pix, g8);
                            atomicAdd(blue + pix, b8);
                            atomicAdd(count + pix, 1u);
                        }
                    }
            }
            if (box.lds) {
                __syncthreads();
                // per-chunk sums stay below 2^32 (<= 2048 fragments x 255 per
                // pixel), so the packed halves never carry into each other
                for (int i = threadIdx.x; i < area; i += kDrawThreads) {
                    const unsigned long long bn = ldsBC[i];
                    const uint32_t n = (uint32_t)(bn >> 32);
                    if (n == 0u) continue;
                    const unsigned long long rg = ldsRG[i];
                    const size_t pix = box.pixel(i, p.width);
                    atomicAdd(red + pix, (uint32_t)rg);
                    atomicAdd(green + pix, (uint32_t)(rg >> 32));
                    atomicAdd(blue + pix, (uint32_t)bn);
                    atomicAdd(count + pix, n);
                }
                __syncthreads();
            }
        }
    }
}

// ---------------------------------------------------------------- resolve

__device__ inline uint32_t unorm8(float x) { return f2u(rintf(saturate(x) * 255.0f)); }

__device__ inline uint32_t pack_rgba(float r, float g, float b, float a) {
    return unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (unorm8(a) << 24);
}

__global__ __launch_bounds__(256) void k_resolve(PassParams p, const uint32_t* __restrict__ planes, size_t stride,
                                                 uint32_t* __restrict__ rgba, float* __restrict__ depthf) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t cnt = planes[4 * stride + i];
    if (cnt == 0u) {
        rgba[i] = pack_rgba(p.bg[0], p.bg[1], p.bg[2], 0.0f);
        depthf[i] = 0.0f;
        return;
    }
    const float inv = rcp((float)cnt);
    const float r = ((float)planes[stride + i] * inv) * kInv255;
    const float g = ((float)planes[2 * stride + i] * inv) * kInv255;
    const float b = ((float)planes[3 * stride + i] * inv) * kInv255;
    rgba[i] = pack_rgba(r, g, b, 1.0f);
    depthf[i] = uint_to_depth(planes[i]);
}

// ---------------------------------------------------------------- nearest

constexpr int kNearestTile = 2048;  // u64 per pixel: 16 KB

template <bool kR0>
__global__ __launch_bounds__(kDrawThreads) void k_nearest(CloudDev c, PassParams p,
                                                          unsigned long long* __restrict__ packed) {
    __shared__ unsigned long long lds[kNearestTile];
    __shared__ int red[5 * (kDrawThreads / 64)];
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, v);
        const bool vec = vec_ok(bc);
        for (uint32_t cb = 0; cb < bc.active; cb += kChunk) {
            const uint32_t li0 = cb + threadIdx.x * kPPT;
            ChunkFrags f;
            chunk_front<true, kR0>(c, p, bc, li0, vec, f);
            const ChunkBox box = chunk_box(f, p, kNearestTile, red);
            if (box.frags == 0) continue;
            const int area = box.w * box.h;
            if (box.lds) {
                for (int i = threadIdx.x; i < area; i += kDrawThreads) lds[i] = 0ull;
                __syncthreads();
            }
#pragma unroll
            for (int j = 0; j < kPPT; j++) {
                if (f.px[j] < 0) continue;
                const uint32_t pi = bc.first + li0 + (uint32_t)j;
                // max of depth << 32 | ~index == the two-pass rule: deepest
                // reverse-Z depth, then the lowest point index among equal depths
                const unsigned long long key =
                    ((unsigned long long)depth_to_uint_hw(f.z[j]) << 32) | (unsigned long long)(~pi);
                const int r = kR0 ? 0 : f.r[j];
                for (int oy = -r; oy <= r; oy++)
                    for (int ox = -r; ox <= r; ox++) {
                        if (!kR0 && !inside_footprint(ox, oy, r)) continue;
                        const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                        if (!kR0 && (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height)) continue;
                        if (box.lds) atomicMax(lds + box.local(tx, ty), key);
                        else atomicMax(packed + (size_t)ty * p.width + tx, key);
                    }
            }
            if (box.lds) {
                __syncthreads();
                for (int i = threadIdx.x; i < area; i += kDrawThreads) {
                    const unsigned long long k = lds[i];
                    if (k == 0ull) continue;
                    unsigned long long* g = packed + box.pixel(i, p.width);
                    if (k > *g) atomicMax(g, k);
                }
                __syncthreads();
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_nearest_resolve(PassParams p, const unsigned long long* __restrict__ packed,
                                                         const uint32_t* __restrict__ colors,
                                                         uint32_t* __restrict__ rgba, float* __restrict__ depthf) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = packed[i];
    const uint32_t d = (uint32_t)(k >> 32);
    const uint32_t idx = ~(uint32_t)k;
    if (d == 0u || idx == 0xffffffffu) {
        rgba[i] = pack_rgba(p.bg[0], p.bg[1], p.bg[2], 0.0f);
        depthf[i] = 0.0f;
        return;
    }
    const uint32_t col = colors[idx];
    rgba[i] = pack_rgba((float)(col & 0xffu) * kInv255, (float)((col >> 8) & 0xffu) * kInv255,
                        (float)((col >> 16) & 0xffu) * kInv255, 1.0f);
    depthf[i] = uint_to_depth(d);
}

// ---------------------------------------------------------------- fills

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i + 4 <= n && (((uintptr_t)p) & 15u) == 0) {
        *(uint4*)(p + i) = make_uint4(v, v, v, v);
    } else {
        for (size_t j = i; j < n && j < i + 4; j++) p[j] = v;
    }
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* __restrict__ p, size_t n,
                                                  unsigned long long v) {
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 2;
    for (size_t j = i; j < n && j < i + 2; j++) p[j] = v;
}

unsigned grid_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

unsigned draw_grid(int numCUs, int nSlots) {
    const int g = numCUs * 8;
    return (unsigned)(nSlots < g ? (nSlots > 0 ? nSlots : 1) : g);
}

}  // namespace

hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 1024)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, size_t n, unsigned long long v) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 512)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_cull(hipStream_t s, const CloudDev& c, const PassParams& p) {
    hipError_t e = launch_fill_u32(s, c.counter, 1, 0u);
    if (e != hipSuccess) return e;
    if (c.nSlots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cull, dim3(grid_for((size_t)c.nSlots, 256)), dim3(256), 0, s, c, p);
    return hipGetLastError();
}

hipError_t launch_depth(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, uint32_t* depth) {
    if (p.constRadius == 0)
        hipLaunchKernelGGL(k_depth<true>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kDrawThreads), 0, s, c, p, depth);
    else
        hipLaunchKernelGGL(k_depth<false>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kDrawThreads), 0, s, c, p, depth);
    return hipGetLastError();
}

hipError_t launch_color(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, const uint32_t* depth,
                        uint32_t* red, uint32_t* green, uint32_t* blue, uint32_t* count) {
    if (p.constRadius == 0)
        hipLaunchKernelGGL(k_color<true>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kDrawThreads), 0, s, c, p, depth,
                           red, green, blue, count);
    else
        hipLaunchKernelGGL(k_color<false>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kDrawThreads), 0, s, c, p, depth,
                           red, green, blue, count);
    return hipGetLastError();
}

hipError_t launch_resolve(hipStream_t s, const PassParams& p, const uint32_t* planes, size_t stride, uint32_t* rgba,
                          float* depthf) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_resolve, dim3(grid_for(n, 256)), dim3(256), 0, s, p, planes, stride, rgba, depthf);
    return hipGetLastError();
}

hipError_t launch_nearest(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p,
                          unsigned long long* packed) {
    if (p.constRadius == 0)
        hipLaunchKernelGGL(k_nearest<true>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kDrawThreads), 0, s, c, p, packed);
    else
        hipLaunchKernelGGL(k_nearest<false>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kDrawThreads), 0, s, c, p, packed);
    return hipGetLastError();
}

hipError_t launch_nearest_resolve(hipStream_t s, const PassParams& p, const unsigned long long* packed,
                                  const uint32_t* colors, uint32_t* rgba, float* depthf) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nearest_resolve, dim3(grid_for(n, 256)), dim3(256), 0, s, p, packed, colors, rgba,
                       depthf);
    return hipGetLastError();
}

}  // namespace scr
