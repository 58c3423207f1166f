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
//   k_frame_front     the frame's clear (accumulation or winner planes), the
//                     composed file tables (updateFiles) and the clouds'
//                     culls in one launch
//
// HIP has no indirect dispatch: the draw kernels are persistent grids sized to
// the machine; each workgroup walks the visible list the cull kernel wrote
// (count read on the device), so no host round trip separates cull and draw.
#include <algorithm>

#include "scr_draw.hpp"

namespace scr {
namespace {

using namespace draw;

constexpr int kThreads = 256;   // draw workgroup
constexpr int kPPT = 8;         // points per thread
constexpr bool kWaveTiles = false;  // chunks and LDS tiles shared by the workgroup (measured faster than per-wave tiles)

// The cull in 256-thread blocks of kCullSlotsPerBlock slots: waves 2 and 3
// prove each slot's bounds (cull_bounds, the double arithmetic) while waves 0
// and 1 take the rest of its decision, then read the bounds from LDS.  A
// visible slot is appended with one counter update per wave (all lanes reach
// the ballot), each visible lane taking the place its rank among the wave's
// visible lanes gives it.  `blk` is the block's index among the cull blocks.
__device__ __forceinline__ void cull_block(const CloudDev& c, const PassParams& p, int blk) {
    __shared__ BatchBounds sBounds[kCullSlotsPerBlock];
    const int t = (int)threadIdx.x;
    const bool proves = t >= kCullSlotsPerBlock;
    const int gid = blk * kCullSlotsPerBlock + (proves ? t - kCullSlotsPerBlock : t);
    scr_visible_batch vb;
    DrawRecord dr;
    bool vis = false;
    if (proves) {
        if (gid < c.nSlots) sBounds[t - kCullSlotsPerBlock] = cull_bounds(c, p, gid);
    } else {
        // the rest of the decision meanwhile; the bounds are filled in below
        const BatchBounds later{};
        vis = gid < c.nSlots && cull_slot(c, p, gid, vb, dr, &later);
    }
    __syncthreads();
    if (proves) return;
    if (vis && p.applyDisplacement == 0) dr.bb = sBounds[t];
    const uint64_t m = __ballot(vis);
    if (m == 0ull) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0u;
    if ((int)__lane_id() == leader) base = atomicAdd(c.counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (!vis) return;
    const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
    c.visible[slot] = vb;
    c.bounds[slot] = dr;
}

__global__ __launch_bounds__(256) void k_cull(CloudDev c, PassParams p) {
    cull_block(c, p, (int)blockIdx.x);
}

// The draw kernels are persistent grids: workgroup b takes visible batches
// b, b + grid, ... up to the count the cull kernel left on the device.

// kDisp: points drawn at position + displacement (a kernel of its own, so
// the undisplaced kernels keep their code).
template <bool kR0, bool kDisp>
__global__ __launch_bounds__(kThreads, draw_waves(kPPT)) void k_depth(CloudDev c, PassParams p, uint32_t* __restrict__ depth) {
    __shared__ uint32_t lds[kDepthTile + kSpare];
    __shared__ int red[5 * kMaxWaves];
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x)
        depth_batch<kR0, kWaveTiles, kThreads, kPPT, kDepthTile, kDisp>(c, p, c.visible[v], c.bounds + v, 0u,
                                                                         0xffffffffu, depth, lds, red);
}

template <bool kR0, bool kDisp>
__global__ __launch_bounds__(kThreads, kR0 ? 4 : 3) void k_color(CloudDev c, PassParams p, const uint32_t* __restrict__ depth,
                                                    uint32_t* red, uint32_t* green, uint32_t* blue,
                                                    uint32_t* count) {
    __shared__ uint32_t ldsDepth[kColorTile];
    __shared__ unsigned long long ldsRGN[kColorTile];
    __shared__ uint32_t ldsB[kColorTile];
    __shared__ int redbuf[5 * kMaxWaves];
    const ColorLds L = color_lds(ldsDepth, ldsRGN, ldsB, redbuf);
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x)
        color_batch<kR0, kWaveTiles, kThreads, kPPT, kDisp>(c, p, c.visible[v], c.bounds + v, 0u, 0xffffffffu, depth,
                                                           red, green, blue, count, L);
}

__global__ __launch_bounds__(256) void k_resolve(PassParams p, const uint32_t* __restrict__ planes, size_t stride,
                                                 uint32_t* __restrict__ rgba, float* __restrict__ depthf) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    resolve_pixel(p, planes, stride, i, rgba, depthf);
}

template <bool kR0>
__global__ __launch_bounds__(kThreads, draw_waves(kPPT)) void k_nearest(CloudDev c, PassParams p,
                                                      unsigned long long* __restrict__ packed) {
    __shared__ unsigned long long lds[kNearestTile + kSpare];
    __shared__ int red[5 * kMaxWaves];
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x)
        nearest_batch<kR0, kWaveTiles, kThreads, kPPT>(c, p, c.visible[v], c.bounds + v, 0u, 0xffffffffu, packed, lds, red);
}

__global__ __launch_bounds__(256) void k_nearest_resolve(PassParams p, const unsigned long long* __restrict__ packed,
                                                         const uint32_t* __restrict__ colors,
                                                         uint32_t* __restrict__ rgba, float* __restrict__ depthf) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    nearest_resolve_pixel(p, packed, colors, i, rgba, depthf);
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

// launch_frame_front: the cull blocks first (their chains of dependent loads
// and double arithmetic start while the clear streams); the clear blocks store
// four 16-byte words per thread (4,096 words per block), and the first of them
// also writes the tail words, the counter words and the file tables.
__global__ __launch_bounds__(256) void k_frame_front(FrameSetup su, CullSet cs, PassParams p) {
    const int nCull = cs.blockStart[cs.nClouds];
    if ((int)blockIdx.x < nCull) {
        int k = 0;
        while (k + 1 < cs.nClouds && (int)blockIdx.x >= cs.blockStart[k + 1]) k++;
        CloudDev c = cs.c[k];
        c.files = su.files + su.first[k];
        cull_block(c, p, (int)blockIdx.x - cs.blockStart[k]);
        return;
    }
    const size_t blk = blockIdx.x - (unsigned)nCull;
    const size_t nv = su.words / 4;
    uint4* p4 = reinterpret_cast<uint4*>(su.planes);
    const size_t b = blk * 1024 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (b + (size_t)k * 256 < nv) p4[b + (size_t)k * 256] = make_uint4(0u, 0u, 0u, 0u);
    if (blk != 0) return;
    if (threadIdx.x < (unsigned)(su.words & 3)) su.planes[nv * 4 + threadIdx.x] = 0u;
    constexpr int kWords = (int)(sizeof(scr_raster_file) / 4);
    for (int c = 0; c < su.nClouds; c++) {
        if (threadIdx.x == 0) *su.zero[c] = 0u;
        const uint32_t* src = reinterpret_cast<const uint32_t*>(su.files + su.first[c]);
        uint32_t* dst = reinterpret_cast<uint32_t*>(su.dst[c]);
        for (int i = (int)threadIdx.x; i < su.count[c] * kWords; i += (int)blockDim.x) dst[i] = src[i];
    }
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* __restrict__ p, size_t n,
                                                  unsigned long long v) {
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 2;
    for (size_t j = i; j < n && j < i + 2; j++) p[j] = v;
}

// Global atomic throughput as the draw passes' tile flushes use it: each wave
// updates 64 consecutive words (one per lane) at a pseudo-random 256-byte
// aligned spot of the buffer, `iters` times.
__global__ __launch_bounds__(256) void k_atomic_rate(uint32_t* __restrict__ buf, uint32_t spots, int add, int iters) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = __lane_id();
    uint32_t h = wave * 0x9E3779B9u + 0x7F4A7C15u;
    for (int i = 0; i < iters; i++) {
        h ^= h >> 16;
        h *= 0x7FEB352Du;
        h ^= h >> 15;
        uint32_t* p = buf + (size_t)(h % spots) * 64u + lane;
        if (add) atomicAdd(p, 1u);
        else atomicMax(p, h);
    }
}

unsigned grid_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

unsigned draw_grid(int numCUs, int nSlots) {
    // about one workgroup per visible batch up to 24 per CU: more workgroups
    // than are resident keep the tail of the pass balanced (measured)
    const int g = numCUs * 24;
    return (unsigned)(nSlots < g ? (nSlots > 0 ? nSlots : 1) : g);
}

}  // namespace

hipError_t launch_frame_front(hipStream_t s, const FrameSetup& su, const CullSet& cs, const PassParams& p) {
    static_assert(sizeof(FrameSetup) + sizeof(CullSet) + sizeof(PassParams) <= 4096,
                  "the front travels in the kernel arguments");
    if (su.nClouds < 0 || su.nClouds > kSetupClouds || cs.nClouds < 0 || cs.nClouds > su.nClouds)
        return hipErrorInvalidValue;
    int files = 0;
    for (int c = 0; c < su.nClouds; c++) {
        if (su.first[c] < 0 || su.count[c] < 0 || su.first[c] + su.count[c] > kSetupFiles) return hipErrorInvalidValue;
        files += su.count[c];
    }
    if (files > kSetupFiles || (((uintptr_t)su.planes) & 15u) != 0) return hipErrorInvalidValue;
    // cloud k's blocks cover its slots, and its table is the one it reads
    if (cs.blockStart[0] != 0) return hipErrorInvalidValue;
    for (int c = 0; c < cs.nClouds; c++)
        if (cs.blockStart[c + 1] - cs.blockStart[c] !=
                (int)grid_for((size_t)std::max(cs.c[c].nSlots, 0), kCullSlotsPerBlock) ||
            cs.c[c].nFiles != su.count[c])
            return hipErrorInvalidValue;
    const unsigned grid = (unsigned)cs.blockStart[cs.nClouds] + std::max(1u, grid_for(su.words / 4, 1024));
    hipLaunchKernelGGL(k_frame_front, dim3(grid), dim3(256), 0, s, su, cs, p);
    return hipGetLastError();
}

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

// ---------------------------------------------------------------- coverage (OIT) passes
// Translucent defocus (applyTint with tintAlphaIsCoverage) and motion blur
// replace the depth-tested colour sum by a weighted blend; these passes are
// one point per thread, one visible batch per workgroup iteration, with
// global atomics (outside the measured opaque path).

namespace {
using namespace draw;

// One point through the shared front end (decode, projection, tests,
// footprint radius); false when it has no fragment.
// kCovPPT points per thread: their loads are issued together
constexpr int kCovPPT = 4;

template <int kLevel, bool kDisp>
__device__ __forceinline__ void coverage_front(const CloudDev& c, const PassParams& p, const BatchCtx& bc,
                                               uint32_t li0, bool vec, ChunkFrags<kCovPPT>& f) {
    ChunkRaw<kCovPPT> raw;
    chunk_fetch<kLevel, kCovPPT>(c, bc, li0, vec, raw);
    chunk_decode<false, false, kLevel, kCovPPT, true, kDisp>(c, p, bc, li0, raw, f);
}

template <bool kDisp>
__device__ __forceinline__ void coverage_front_level(const CloudDev& c, const PassParams& p, const BatchCtx& bc,
                                                     uint32_t li0, bool vec, ChunkFrags<kCovPPT>& f) {
    if (bc.level == 0) coverage_front<0, kDisp>(c, p, bc, li0, vec, f);
    else if (bc.level == 1) coverage_front<1, kDisp>(c, p, bc, li0, vec, f);
    else coverage_front<2, kDisp>(c, p, bc, li0, vec, f);
}

__device__ __forceinline__ void coverage_front_any(const CloudDev& c, const PassParams& p, const BatchCtx& bc,
                                                   uint32_t li0, bool vec, ChunkFrags<kCovPPT>& f) {
    if (p.applyDisplacement != 0) coverage_front_level<true>(c, p, bc, li0, vec, f);
    else coverage_front_level<false>(c, p, bc, li0, vec, f);
}

// depthPassUpdate in translucent-defocus mode (DepthPass/Shaders.metal:60-67):
// a point whose tint alpha is the discard sentinel (< 0) claims no depth.
// Maxima meet in an LDS tile over the batch rectangle, flushed once.
// Its tile stays at 4,096 words: clearing and flushing the single-view depth
// tile's 8,192 per batch measured 9 % slower on translucent defocus.
constexpr int kCoverageDepthTile = 4096;

__global__ __launch_bounds__(256) void k_depth_coverage(CloudDev c, PassParams p, uint32_t* __restrict__ depth) {
    __shared__ uint32_t tile[kCoverageDepthTile];
    __shared__ int redbuf[5 * kMaxWaves];
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, p, c.visible[v], c.bounds + v, 0xffffffffu);
        const ChunkBox seg = segment_box(p, bc, bc.active, kCoverageDepthTile);
        const int area = seg.w * seg.h;
        if (seg.lds) {
            for (int i = (int)threadIdx.x; i < area; i += (int)blockDim.x) lds_store(tile + i, 0u);
            __syncthreads();
        }
        const bool vec = vec_ok<kCovPPT>(bc);
        for (uint32_t cb = 0; cb < bc.active; cb += blockDim.x * kCovPPT) {
            const uint32_t li0 = cb + threadIdx.x * kCovPPT;
            ChunkFrags<kCovPPT> f;
            coverage_front_any(c, p, bc, li0, vec, f);
            // a batch too wide for one tile: the chunk's own rectangle
            ChunkBox box = seg;
            if (!seg.lds) {
                box = chunk_box<false, 256, kCovPPT>(f, p, kCoverageDepthTile, redbuf);
                if (box.lds) {
                    for (int i = (int)threadIdx.x; i < box.w * box.h; i += (int)blockDim.x) lds_store(tile + i, 0u);
                    __syncthreads();
                }
            }
#pragma unroll
            for (int j = 0; j < kCovPPT; j++) {
            if (!f.ok[j]) continue;
            if (c.tints && c.tints[bc.first + li0 + (uint32_t)j].w < 0.0f) continue;
            const uint32_t d = depth_to_uint_hw(f.z[j]);
            const int rad = f.r[j];
            for (int oy = -rad; oy <= rad; oy++)
                for (int ox = -rad; ox <= rad; ox++) {
                    if (!inside_footprint(ox, oy, rad)) continue;
                    const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                    if (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height) continue;
                    if (box.lds && box_holds(box, tx, ty)) lds_max(tile + box.local(tx, ty), d);
                    else gmax(depth + (size_t)ty * p.width + tx, d);
                }
            }
            if (!seg.lds && box.lds) {
                __syncthreads();
                for (int i = (int)threadIdx.x; i < box.w * box.h; i += (int)blockDim.x) {
                    const uint32_t d = lds_load(tile + i);
                    if (d != 0u) gmax(depth + box.pixel(i, p.width), d);
                }
                __syncthreads();
            }
        }
        if (seg.lds) {
            __syncthreads();
            for (int i = (int)threadIdx.x; i < area; i += (int)blockDim.x) {
                const uint32_t d = lds_load(tile + i);
                if (d != 0u) gmax(depth + seg.pixel(i, p.width), d);
            }
            __syncthreads();
        }
    }
}

// colorPassUpdate, weighted-blend path (ColorPass/Shaders.metal:96-209): no
// depth test; each fragment adds colour * a * 4096 and a * 4096 (a = 1 -
// tint alpha in defocus mode, else 1), split over the motion-blur sweep from
// the point's pixel back toward its previous frame's pixel.
// The sums of a batch meet in an LDS tile over the batch rectangle (widened
// by the sweep's reach) and are flushed once, as in the opaque colour pass:
// per pixel and batch they stay below 2^32 (65,535 points, each adding at
// most 4096 + 0.5 per sample to a pixel over a sweep that splits it), and
// integer sums do not depend on the order they are made in.
// Adds a blend tile's sums to the planes (a pixel with no coverage has no
// colour either: each fragment's colour sums are at most its coverage sum).
__device__ __forceinline__ void coverage_tile_flush(const ChunkBox& box, const PassParams& p,
                                                    const unsigned long long* tileRG,
                                                    const unsigned long long* tileBA, uint32_t* red,
                                                    uint32_t* green, uint32_t* blue, uint32_t* count) {
    const int area = box.w * box.h;
    for (int i = (int)threadIdx.x; i < area; i += (int)blockDim.x) {
        const unsigned long long ba = lds_load(tileBA + i);
        if ((uint32_t)(ba >> 32) == 0u) continue;
        const unsigned long long rg = lds_load(tileRG + i);
        const size_t pix = box.pixel(i, p.width);
        gadd(red + pix, (uint32_t)rg);
        gadd(green + pix, (uint32_t)(rg >> 32));
        gadd(blue + pix, (uint32_t)ba);
        gadd(count + pix, (uint32_t)(ba >> 32));
    }
}

__device__ __forceinline__ void coverage_tile_clear(const ChunkBox& box, unsigned long long* tileRG,
                                                    unsigned long long* tileBA) {
    const int area = box.w * box.h;
    for (int i = (int)threadIdx.x; i < area; i += (int)blockDim.x) {
        lds_store(tileRG + i, 0ull);
        lds_store(tileBA + i, 0ull);
    }
}

__global__ __launch_bounds__(256) void k_color_coverage(CloudDev c, PassParams p, CoverageParams q,
                                                        uint32_t* __restrict__ red, uint32_t* __restrict__ green,
                                                        uint32_t* __restrict__ blue, uint32_t* __restrict__ count) {
    __shared__ unsigned long long tileRG[kBlendTile];
    __shared__ unsigned long long tileBA[kBlendTile];
    __shared__ int redbuf[5 * kMaxWaves];
    // a sweep reaches at most maxSpread + 1 pixels from the point's pixel
    const int reach = q.motionBlur > 0.0f ? ((q.maxSpread >= 0.0f && q.maxSpread < 4096.0f)
                                                 ? (int)ceilf(q.maxSpread) + 1 : (1 << 20))
                                          : 0;
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const scr_visible_batch vb = c.visible[v];
        const BatchCtx bc = batch_ctx(c, p, vb, c.bounds + v, 0xffffffffu);
        const scr_raster_file* file = c.files + c.batches[vb.batchIndex].fileIndex;
        const ChunkBox seg = segment_box(p, bc, bc.active, kBlendTile, reach);
        if (seg.lds) {
            coverage_tile_clear(seg, tileRG, tileBA);
            __syncthreads();
        }
        const bool vec = vec_ok<kCovPPT>(bc);
        for (uint32_t cb = 0; cb < bc.active; cb += blockDim.x * kCovPPT) {
            const uint32_t li0 = cb + threadIdx.x * kCovPPT;
            ChunkFrags<kCovPPT> f;
            coverage_front_any(c, p, bc, li0, vec, f);
            // a batch too wide for one tile: the chunk's own rectangle, when
            // it fits (as the opaque colour pass does)
            ChunkBox box = seg;
            if (!seg.lds) {
                box = chunk_box<false, 256, kCovPPT>(f, p, kBlendTile, redbuf);
                if (box.lds) {
                    coverage_tile_clear(box, tileRG, tileBA);
                    __syncthreads();
                }
            }
            // runs of the thread's consecutive single-pixel fragments on one
            // tile pixel (Morton order) are summed in registers and added once
            int runK = -1;
            unsigned long long runRG = 0ull, runBA = 0ull;
#pragma unroll
            for (int j = 0; j < kCovPPT; j++) {
            if (!f.ok[j]) continue;
            const uint32_t pi = bc.first + li0 + (uint32_t)j;
            uint32_t color = c.colors[pi];
            if (p.colorizeChunks != 0) color = vb.batchIndex * 1234567u;
            else if (p.colorizeOverdraw != 0) color = 0x00010101u;
            uint32_t r = color & 0xffu, g = (color >> 8) & 0xffu, b = (color >> 16) & 0xffu;
            float a = 1.0f;
            if (p.applyTint != 0) {
                const float4 t = c.tints ? c.tints[pi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (t.w < 0.0f) continue;
                if (q.coverageMode != 0) {
                    a = saturate(1.0f - t.w);
                } else {
                    const float w = saturate(t.w);
                    const float ox = (float)r * kInv255, oy = (float)g * kInv255, oz = (float)b * kInv255;
                    r = f2u(saturate(ox + (saturate(t.x) - ox) * w) * 255.0f + 0.5f);
                    g = f2u(saturate(oy + (saturate(t.y) - oy) * w) * 255.0f + 0.5f);
                    b = f2u(saturate(oz + (saturate(t.z) - oz) * w) * 255.0f + 0.5f);
                }
            }
            if (a <= 0.0f) continue;
            // motion blur (:146-167): pixel differences wrap as the shader's int2 does
            int samples = 1;
            float sx = 0.0f, sy = 0.0f;
            if (q.motionBlur > 0.0f) {
                // from the decoded position + the previous displacement
                // (ColorPass/Shaders.metal:150-153)
                float bx = f.wx[j], by = f.wy[j], bz = f.wz[j];
                if (p.applyDisplacement != 0 && c.prevDisp) {
                    const float4 d = c.prevDisp[pi];
                    bx = bx + d.x;
                    by = by + d.y;
                    bz = bz + d.z;
                } else if (p.applyDisplacement != 0) {
                    bx = bx + 0.0f;
                    by = by + 0.0f;
                    bz = bz + 0.0f;
                }
                const f4 prev = mul_point(file->prevTransform, bx, by, bz);
                if (prev.w > 0.0f) {
                    const float prw = rcp(prev.w);
                    const int ppx = f2i_hw(fmaf(prev.x * prw, 0.5f, 0.5f) * (float)p.width);
                    const int ppy = p.height - 1 - f2i_hw(fmaf(prev.y * prw, 0.5f, 0.5f) * (float)p.height);
                    float vx = (float)(int)((uint32_t)f.px[j] - (uint32_t)ppx) * q.motionBlur;
                    float vy = (float)(int)((uint32_t)f.py[j] - (uint32_t)ppy) * q.motionBlur;
                    float len = sqrtf(vx * vx + vy * vy);
                    if (len > q.maxSpread) {
                        const float k = q.maxSpread * rcp(len);
                        vx = vx * k;
                        vy = vy * k;
                        len = q.maxSpread;
                    }
                    if (len > 0.75f) {
                        samples = min(max(f2i_hw(ceilf(len)), 1), q.samples);
                        const float d = rcp((float)max(samples - 1, 1));
                        sx = vx * d;
                        sy = vy * d;
                    }
                }
            }
            const float aN = a * rcp((float)samples);
            const uint32_t cr = f2u((((float)r * kInv255) * aN) * 4096.0f + 0.5f);
            const uint32_t cg = f2u((((float)g * kInv255) * aN) * 4096.0f + 0.5f);
            const uint32_t cb = f2u((((float)b * kInv255) * aN) * 4096.0f + 0.5f);
            const uint32_t ca = f2u(aN * 4096.0f + 0.5f);
            const int rad = f.r[j];
            if (rad == 0 && samples == 1 && box.lds && box_holds(box, f.px[j], f.py[j])) {
                const int k = box.local(f.px[j], f.py[j]);
                const unsigned long long rg = (unsigned long long)cr | ((unsigned long long)cg << 32);
                const unsigned long long ba = (unsigned long long)cb | ((unsigned long long)ca << 32);
                if (k == runK) {
                    runRG += rg;
                    runBA += ba;
                    continue;
                }
                if (runK >= 0) {
                    lds_add(tileRG + runK, runRG);
                    lds_add(tileBA + runK, runBA);
                }
                runK = k;
                runRG = rg;
                runBA = ba;
                continue;
            }
            for (int k = 0; k < samples; k++) {
                const int cx = (int)((uint32_t)f.px[j] - (uint32_t)f2i_hw(roundf(sx * (float)k)));
                const int cy = (int)((uint32_t)f.py[j] - (uint32_t)f2i_hw(roundf(sy * (float)k)));
                for (int oy = -rad; oy <= rad; oy++)
                    for (int ox = -rad; ox <= rad; ox++) {
                        if (!inside_footprint(ox, oy, rad)) continue;
                        const int tx = (int)((uint32_t)cx + (uint32_t)ox), ty = (int)((uint32_t)cy + (uint32_t)oy);
                        if (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height) continue;
                        if (box.lds && box_holds(box, tx, ty)) {
                            const int k = box.local(tx, ty);
                            lds_add(tileRG + k, (unsigned long long)cr | ((unsigned long long)cg << 32));
                            lds_add(tileBA + k, (unsigned long long)cb | ((unsigned long long)ca << 32));
                            continue;
                        }
                        const size_t pix = (size_t)ty * p.width + tx;
                        gadd(red + pix, cr);
                        gadd(green + pix, cg);
                        gadd(blue + pix, cb);
                        gadd(count + pix, ca);
                    }
            }
            }
            if (runK >= 0) {
                lds_add(tileRG + runK, runRG);
                lds_add(tileBA + runK, runBA);
            }
            if (!seg.lds && box.lds) {
                __syncthreads();
                coverage_tile_flush(box, p, tileRG, tileBA, red, green, blue, count);
                __syncthreads();
            }
        }
        if (seg.lds) {
            __syncthreads();
            coverage_tile_flush(seg, p, tileRG, tileBA, red, green, blue, count);
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(256) void k_resolve_coverage(PassParams p, const uint32_t* __restrict__ planes,
                                                          size_t stride, uint32_t* __restrict__ rgba,
                                                          float* __restrict__ depthf) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    resolve_coverage_pixel(p, planes, stride, i, rgba, depthf);
}

}  // namespace

hipError_t launch_depth_coverage(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, uint32_t* depth) {
    hipLaunchKernelGGL(k_depth_coverage, dim3(draw_grid(numCUs, c.nSlots)), dim3(256), 0, s, c, p, depth);
    return hipGetLastError();
}

hipError_t launch_color_coverage(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p,
                                 const CoverageParams& q, uint32_t* red, uint32_t* green, uint32_t* blue,
                                 uint32_t* count) {
    hipLaunchKernelGGL(k_color_coverage, dim3(draw_grid(numCUs, c.nSlots)), dim3(256), 0, s, c, p, q, red, green,
                       blue, count);
    return hipGetLastError();
}

hipError_t launch_resolve_coverage(hipStream_t s, const PassParams& p, const uint32_t* planes, size_t stride,
                                   uint32_t* rgba, float* depthf) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    hipLaunchKernelGGL(k_resolve_coverage, dim3(grid_for(n, 256)), dim3(256), 0, s, p, planes, stride, rgba, depthf);
    return hipGetLastError();
}

// One hole-fill pass over a width x height RGBA8 plane, 16 x 16 pixels per
// workgroup (HoleFillProcessor.swift:37-44).
// The block's 18 x 18 texels (its 16 x 16 pixels and a one-pixel ring) are
// staged in LDS; a texel outside the image is staged as alpha 0, which no
// neighbour sum counts, the same as skipping it.
__global__ __launch_bounds__(256) void k_hole_fill(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                   int w, int h) {
    __shared__ uint32_t t[18 * 18];
    const int bx = (int)(blockIdx.x * 16u) - 1, by = (int)(blockIdx.y * 16u) - 1;
    for (int i = (int)threadIdx.x; i < 18 * 18; i += 256) {
        const int sx = bx + i % 18, sy = by + i / 18;
        t[i] = (sx >= 0 && sx < w && sy >= 0 && sy < h) ? in[(size_t)sy * w + sx] : 0u;
    }
    __syncthreads();
    const int lx = (int)(threadIdx.x & 15u), ly = (int)(threadIdx.x >> 4);
    const int x = bx + 1 + lx, y = by + 1 + ly;
    if (x >= w || y >= h) return;
    out[(size_t)y * w + x] = draw::hole_fill_texels(t + (ly + 1) * 18 + (lx + 1), 18);
}

// ComputeRasteriser.draw's composite (Post.metal:23-50), one thread per pixel.
__global__ __launch_bounds__(256) void k_composite(const uint32_t* __restrict__ rgba, const float* __restrict__ depthf,
                                                   int w, int h, scr_composite_options o, uint32_t* __restrict__ target,
                                                   float* __restrict__ targetDepth, int pitch) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * (size_t)h) return;
    const size_t x = i % (size_t)w, y = i / (size_t)w;
    const size_t ti = y * (size_t)pitch + x;
    draw::composite_pixel(rgba[i], depthf[i], target + ti, targetDepth ? targetDepth + ti : nullptr, o);
}

hipError_t launch_composite(hipStream_t s, const uint32_t* rgba, const float* depthf, int w, int h,
                            const scr_composite_options& o, uint32_t* target, float* targetDepth, int pitch) {
    const size_t n = (size_t)w * (size_t)h;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_composite, dim3(grid_for(n, 256)), dim3(256), 0, s, rgba, depthf, w, h, o, target,
                       targetDepth, pitch);
    return hipGetLastError();
}

hipError_t launch_hole_fill(hipStream_t s, const uint32_t* in, uint32_t* out, int w, int h) {
    const dim3 grid((unsigned)(w + 15) / 16u, (unsigned)(h + 15) / 16u);
    hipLaunchKernelGGL(k_hole_fill, grid, dim3(256), 0, s, in, out, w, h);
    return hipGetLastError();
}

// Every float in [2^-126, 2^126): rcp_fast against the correctly rounded
// reciprocal, bit for bit.
__global__ __launch_bounds__(256) void k_check_fast_rcp(unsigned long long* __restrict__ mismatches,
                                                        uint32_t* __restrict__ first) {
    constexpr uint64_t kLo = 0x00800000u, kHi = 0x7e800000u;
    uint32_t bad = 0u, lowest = 0xffffffffu;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = kLo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < kHi; i += step) {
        const float w = __uint_as_float((uint32_t)i);
        if (__float_as_uint(rcp_fast(w)) != __float_as_uint(rcp(w))) {
            bad++;
            lowest = min(lowest, (uint32_t)i);
        }
    }
    if (bad != 0u) {
        atomicAdd(mismatches, (unsigned long long)bad);
        atomicMin(first, lowest);
    }
}

hipError_t launch_check_fast_rcp(hipStream_t s, int numCUs, unsigned long long* mismatches, uint32_t* first) {
    hipLaunchKernelGGL(k_check_fast_rcp, dim3((unsigned)numCUs * 8u), dim3(256), 0, s, mismatches, first);
    return hipGetLastError();
}

hipError_t launch_atomic_rate(hipStream_t s, int numCUs, uint32_t* buf, uint32_t spots, int add, int iters,
                              uint64_t* atomics) {
    const unsigned grid = (unsigned)numCUs * 16u;
    hipLaunchKernelGGL(k_atomic_rate, dim3(grid), dim3(256), 0, s, buf, spots, add, iters);
    *atomics = (uint64_t)grid * 256u * (uint64_t)iters;
    return hipGetLastError();
}

hipError_t launch_cull(hipStream_t s, const CloudDev& c, const PassParams& p) {
    hipError_t e = launch_fill_u32(s, c.counter, 1, 0u);
    if (e != hipSuccess) return e;
    if (c.nSlots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cull, dim3(grid_for((size_t)c.nSlots, kCullSlotsPerBlock)), dim3(256), 0, s, c, p);
    return hipGetLastError();
}

hipError_t launch_depth(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, uint32_t* depth) {
    const dim3 grid(draw_grid(numCUs, c.nSlots));
    const bool r0 = p.constRadius == 0, disp = p.applyDisplacement != 0;
    if (disp && r0) hipLaunchKernelGGL((k_depth<true, true>), grid, dim3(kThreads), 0, s, c, p, depth);
    else if (disp) hipLaunchKernelGGL((k_depth<false, true>), grid, dim3(kThreads), 0, s, c, p, depth);
    else if (r0) hipLaunchKernelGGL((k_depth<true, false>), grid, dim3(kThreads), 0, s, c, p, depth);
    else hipLaunchKernelGGL((k_depth<false, false>), grid, dim3(kThreads), 0, s, c, p, depth);
    return hipGetLastError();
}

hipError_t launch_color(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, const uint32_t* depth,
                        uint32_t* red, uint32_t* green, uint32_t* blue, uint32_t* count) {
    const dim3 grid(draw_grid(numCUs, c.nSlots));
    const bool r0 = p.constRadius == 0, disp = p.applyDisplacement != 0;
    if (disp && r0)
        hipLaunchKernelGGL((k_color<true, true>), grid, dim3(kThreads), 0, s, c, p, depth, red, green, blue, count);
    else if (disp)
        hipLaunchKernelGGL((k_color<false, true>), grid, dim3(kThreads), 0, s, c, p, depth, red, green, blue, count);
    else if (r0)
        hipLaunchKernelGGL((k_color<true, false>), grid, dim3(kThreads), 0, s, c, p, depth, red, green, blue, count);
    else
        hipLaunchKernelGGL((k_color<false, false>), grid, dim3(kThreads), 0, s, c, p, depth, red, green, blue, count);
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
        hipLaunchKernelGGL(k_nearest<true>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kThreads), 0, s, c, p, packed);
    else
        hipLaunchKernelGGL(k_nearest<false>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kThreads), 0, s, c, p, packed);
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
