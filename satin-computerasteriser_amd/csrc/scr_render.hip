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
    uint32_t first, active, batchIndex;
    const float* T;     // file.transform
    const float* W;     // file.world
};

__device__ inline BatchCtx batch_ctx(const CloudDev& c, uint32_t v) {
    const scr_visible_batch vb = c.visible[v];
    const scr_raster_batch b = c.batches[vb.batchIndex];
    BatchCtx bc;
    bc.level = vb.level;
    bc.lodThreshold = vb.lodThreshold;
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
    bc.T = c.files[b.fileIndex].transform;
    bc.W = c.files[b.fileIndex].world;
    return bc;
}

struct Frag {
    float x, y, z;   // decoded point
    float ndcZ;
    int px, py;      // pixel, y already flipped
};

// dither/LOD skip, decodePoint (Common.metal:73-117), projection, clip and
// viewport tests, y flip — the shared prologue of every point pass.
template <bool kNearestZ>
__device__ inline bool point_front(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t pi,
                                   Frag& fr) {
    const float pointLevel = (float)((uint32_t)c.levels[pi] & 7u);
    const float dither = (p.lodDither != 0) ? hash_unit(pi) : 0.5f;
    if (dither >= (bc.lodThreshold - pointLevel) + 0.5f) return false;
    const uint32_t lo = c.xyzLow[pi];
    uint32_t qx, qy, qz;
    if (bc.level == 0) {
        const uint32_t me = c.xyzMed[pi], hi = c.xyzHigh[pi];
        qx = ((lo & 1023u) << 20) | ((me & 1023u) << 10) | (hi & 1023u);
        qy = (((lo >> 10) & 1023u) << 20) | (((me >> 10) & 1023u) << 10) | ((hi >> 10) & 1023u);
        qz = (((lo >> 20) & 1023u) << 20) | (((me >> 20) & 1023u) << 10) | ((hi >> 20) & 1023u);
    } else if (bc.level == 1) {
        const uint32_t me = c.xyzMed[pi];
        qx = ((lo & 1023u) << 10) | (me & 1023u);
        qy = (((lo >> 10) & 1023u) << 10) | ((me >> 10) & 1023u);
        qz = (((lo >> 20) & 1023u) << 10) | ((me >> 20) & 1023u);
    } else {
        qx = lo & 1023u;
        qy = (lo >> 10) & 1023u;
        qz = (lo >> 20) & 1023u;
    }
    fr.x = (float)qx * bc.sx + bc.mnx;
    fr.y = (float)qy * bc.sy + bc.mny;
    fr.z = (float)qz * bc.sz + bc.mnz;
    const f4 clip = mul_point(bc.T, fr.x, fr.y, fr.z);
    if (clip.w <= 0.0f) return false;
    const float rw = rcp(clip.w);
    const float nx = clip.x * rw, ny = clip.y * rw, nz = clip.z * rw;
    if (nx < -1.0f || ny < -1.0f || nx > 1.0f || ny > 1.0f) return false;
    if (kNearestZ) {
        if (nz <= 0.0f || nz > 1.0f) return false;
    } else {
        if (nz < 0.0f || nz > 1.0f) return false;
    }
    const int px = f2i((nx * 0.5f + 0.5f) * (float)p.width);
    const int py = f2i((ny * 0.5f + 0.5f) * (float)p.height);
    if (px < 0 || px >= p.width || py < 0 || py >= p.height) return false;
    fr.px = px;
    fr.py = p.height - 1 - py;
    fr.ndcZ = nz;
    return true;
}

// ---------------------------------------------------------------- depth pass

__global__ __launch_bounds__(kDrawThreads) void k_depth(CloudDev c, PassParams p, uint32_t* __restrict__ depth) {
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, v);
        for (uint32_t li = threadIdx.x; li < bc.active; li += kDrawThreads) {
            const uint32_t pi = bc.first + li;
            Frag fr;
            if (!point_front<false>(c, p, bc, pi, fr)) continue;
            const uint32_t d = depth_to_uint(fr.ndcZ);
            const int r = footprint_radius(p, bc.W, fr.x, fr.y, fr.z);
            if (r == 0) {
                atomicMax(depth + (size_t)fr.py * p.width + fr.px, d);
                continue;
            }
            for (int oy = -r; oy <= r; oy++)
                for (int ox = -r; ox <= r; ox++) {
                    if (!inside_footprint(ox, oy, r)) continue;
                    const int tx = fr.px + ox, ty = fr.py + oy;
                    if (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height) continue;
                    atomicMax(depth + (size_t)ty * p.width + tx, d);
                }
        }
    }
}

// ---------------------------------------------------------------- colour pass

__global__ __launch_bounds__(kDrawThreads) void k_color(CloudDev c, PassParams p, const uint32_t* __restrict__ depth,
                                                        uint32_t* red, uint32_t* green, uint32_t* blue,
                                                        uint32_t* count) {
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, v);
        for (uint32_t li = threadIdx.x; li < bc.active; li += kDrawThreads) {
            const uint32_t pi = bc.first + li;
            Frag fr;
            if (!point_front<false>(c, p, bc, pi, fr)) continue;
            uint32_t color = c.colors[pi];
            if (p.colorizeChunks != 0) color = bc.batchIndex * 1234567u;
            else if (p.colorizeOverdraw != 0) color = 0x00010101u;
            uint32_t r8 = color & 0xffu, g8 = (color >> 8) & 0xffu, b8 = (color >> 16) & 0xffu;
            if (p.applyTint != 0) {
                const float4 t = c.tints ? c.tints[pi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (t.w < 0.0f) continue;  // discard sentinel, ColorPass/Shaders.metal:119
                const float w = saturate(t.w);
                const float ox = (float)r8 * kInv255, oy = (float)g8 * kInv255, oz = (float)b8 * kInv255;
                const float mx = ox + (saturate(t.x) - ox) * w;
                const float my = oy + (saturate(t.y) - oy) * w;
                const float mz = oz + (saturate(t.z) - oz) * w;
                r8 = f2u(saturate(mx) * 255.0f + 0.5f);
                g8 = f2u(saturate(my) * 255.0f + 0.5f);
                b8 = f2u(saturate(mz) * 255.0f + 0.5f);
            }
            const int rad = footprint_radius(p, bc.W, fr.x, fr.y, fr.z);
            for (int oy = -rad; oy <= rad; oy++)
                for (int ox = -rad; ox <= rad; ox++) {
                    if (!inside_footprint(ox, oy, rad)) continue;
                    const int tx = fr.px + ox, ty = fr.py + oy;
                    if (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height) continue;
                    const size_t pix = (size_t)ty * p.width + tx;
                    const uint32_t d = depth[pix];
                    if (d == 0u) continue;
                    if (!(fr.ndcZ >= uint_to_depth(d) * p.depthKeep)) continue;
                    atomicAdd(red + pix, r8);
                    atomicAdd(green + pix, g8);
                    atomicAdd(blue + pix, b8);
                    atomicAdd(count + pix, 1u);
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

__global__ __launch_bounds__(kDrawThreads) void k_nearest(CloudDev c, PassParams p,
                                                          unsigned long long* __restrict__ packed) {
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, v);
        for (uint32_t li = threadIdx.x; li < bc.active; li += kDrawThreads) {
            const uint32_t pi = bc.first + li;
            Frag fr;
            if (!point_front<true>(c, p, bc, pi, fr)) continue;
            const unsigned long long key =
                ((unsigned long long)depth_to_uint(fr.ndcZ) << 32) | (unsigned long long)(~pi);
            const int r = footprint_radius(p, bc.W, fr.x, fr.y, fr.z);
            for (int oy = -r; oy <= r; oy++)
                for (int ox = -r; ox <= r; ox++) {
                    if (!inside_footprint(ox, oy, r)) continue;
                    const int tx = fr.px + ox, ty = fr.py + oy;
                    if (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height) continue;
                    atomicMax(packed + (size_t)ty * p.width + tx, key);
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
    hipLaunchKernelGGL(k_depth, dim3(draw_grid(numCUs, c.nSlots)), dim3(kDrawThreads), 0, s, c, p, depth);
    return hipGetLastError();
}

hipError_t launch_color(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, const uint32_t* depth,
                        uint32_t* red, uint32_t* green, uint32_t* blue, uint32_t* count) {
    hipLaunchKernelGGL(k_color, dim3(draw_grid(numCUs, c.nSlots)), dim3(kDrawThreads), 0, s, c, p, depth, red,
                       green, blue, count);
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
    hipLaunchKernelGGL(k_nearest, dim3(draw_grid(numCUs, c.nSlots)), dim3(kDrawThreads), 0, s, c, p, packed);
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
