// scr_mv.hpp — device side of batched multi-view rendering shared by the
// draw-kernel translation units (scr_mv_depth.hip, scr_mv_color.hip,
// scr_mv_nearest.hip) and the host side (scr_multiview.hip): the per-view
// record, the work-item table, the work-queue drain loop and the draw kernel
// templates.  Each draw phase is compiled in its own translation unit, which
// instantiates every shape of it and exports the table of kernels.
#pragma once

#include "scr_diff_draw.hpp"

namespace scr {
namespace mv {

using namespace draw;

struct MVView {
    PassParams p;
    CloudDev c;                     // files -> this view's composed table
    uint32_t* planes;               // depth, red, green, blue, count (stride apart)
    unsigned long long* nearest;
    uint32_t* rgba;
    float* depthf;
    unsigned long long stride;
    uint32_t itemBase, itemStride;  // item of slot s = itemBase + s * itemStride
    // differentiable mode
    unsigned long long* sums;       // 3 int64 planes, stride apart
    float* image;                   // 3 floats per pixel
    const float* dLdImage;          // 3 floats per pixel
};

struct MVItem {
    uint32_t view;
    uint32_t pad[3];
    scr_visible_batch vb;           // activePoints == 0: culled (the record is then not written)
    DrawRecord bounds;              // what the batch box proves under this view, its draw context (cull)
};
static_assert(sizeof(MVItem) == 160, "MVItem is ten 16-byte words");

constexpr int kQueues = 4;          // queue heads: depth, colour, nearest, differentiable colour

// Takes work units from heads[queue] until the table is drained, kClaim
// consecutive units per claim, whose item headers the first threads read
// together into LDS (culled items, about half of the render workload's
// table, then cost neither a claim's round trip nor a dependent header read;
// a short table takes one unit per claim, which keeps the pass's tail even).  The bodies
// passed in are marked always_inline: a large body is otherwise called as a
// function, with a stack frame, measured up to twice as slow.  A unit is
// (item, part); part k of S draws the k-th of S equal, chunk-aligned pieces of
// the item's active prefix.  Every workgroup leaves the loop once the counter
// passes nUnits, so the grid drains.
constexpr uint32_t kMaxClaim = 4;
// The claim the host passes: kMaxClaim when the table holds at least 32 units
// per workgroup (the pass's tail then stays fine-grained), else 1.
__host__ __device__ inline uint32_t claim_for(uint32_t nUnits, uint32_t grid) {
    return nUnits >= 32u * grid ? kMaxClaim : 1u;
}
template <int NT, int PPT, typename Body>
__device__ __forceinline__ void drain(const MVView* __restrict__ views, const MVItem* __restrict__ items, uint32_t nUnits,
                             int splits, uint32_t claim, uint32_t* head, Body body) {
    static_assert(NT >= (int)kMaxClaim, "a thread per claimed header");
    __shared__ uint32_t sUnit;
    __shared__ uint32_t sView[kMaxClaim];
    __shared__ scr_visible_batch sVb[kMaxClaim];
    constexpr uint32_t kChunk = NT * PPT;
    claim = min(max(claim, 1u), kMaxClaim);
    for (;;) {
        if (threadIdx.x == 0) sUnit = atomicAdd(head, claim);
        __syncthreads();
        const uint32_t u0 = __builtin_amdgcn_readfirstlane(sUnit);
        if (u0 >= nUnits) break;
        const uint32_t uEnd = min(u0 + claim, nUnits);
        if (threadIdx.x < claim && u0 + threadIdx.x < uEnd) {
            const uint32_t item = (u0 + threadIdx.x) / (uint32_t)splits;
            sView[threadIdx.x] = items[item].view;
            sVb[threadIdx.x] = items[item].vb;
        }
        __syncthreads();
        for (uint32_t u = u0; u < uEnd; u++) {
            const uint32_t item = u / (uint32_t)splits, part = u - item * (uint32_t)splits;
            const uint32_t view = __builtin_amdgcn_readfirstlane(sView[u - u0]);
            scr_visible_batch vb = sVb[u - u0];
            vb.batchIndex = __builtin_amdgcn_readfirstlane(vb.batchIndex);
            vb.level = (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)vb.level);
            vb.lodThreshold = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(vb.lodThreshold)));
            vb.activePoints = __builtin_amdgcn_readfirstlane(vb.activePoints);
            const uint32_t active = vb.activePoints;
            if (active == 0u) continue;
            const uint32_t chunks = (active + kChunk - 1) / kChunk;
            const uint32_t per = (chunks + (uint32_t)splits - 1) / (uint32_t)splits;
            const uint32_t lo = part * per * kChunk;
            if (lo >= active) continue;
            const uint32_t hi = min(active, lo + per * kChunk);
            body(views[view], vb, &items[item].bounds, lo, hi);
        }
        // every wave has read the headers before the next claim rewrites them
        __syncthreads();
    }
}

template <bool kR0, bool kWave, int NT, int PPT>
__global__ __launch_bounds__(NT, draw_waves(PPT)) void k_mv_depth(const MVView* __restrict__ views, const MVItem* __restrict__ items,
                                                 uint32_t nUnits, int splits, uint32_t claim, uint32_t* heads) {
    __shared__ uint32_t lds[kMVDepthTile + kSpare];
    __shared__ int red[5 * kMaxWaves];
    drain<NT, PPT>(views, items, nUnits, splits, claim, heads + 0,
                   [&](const MVView& V, const scr_visible_batch& vb, const DrawRecord* bb, uint32_t lo, uint32_t hi) __attribute__((always_inline)) {
                       depth_batch<kR0, kWave, NT, PPT, kMVDepthTile>(V.c, V.p, vb, bb, lo, hi, V.planes, lds, red);
                   });
}

// The multi-view colour tile per workgroup width: the most pixels (16 bytes
// each) that keep four waves per SIMD within a CU's 160 KB of LDS.
__host__ __device__ constexpr int mv_color_tile(int nt) { return nt <= 128 ? 1256 : (nt <= 256 ? kColorTile : 5088); }

template <bool kR0, bool kWave, int NT, int PPT>
__global__ __launch_bounds__(NT, PPT <= 8 ? (kR0 ? 4 : 3) : 2) void k_mv_color(const MVView* __restrict__ views, const MVItem* __restrict__ items,
                                                 uint32_t nUnits, int splits, uint32_t claim, uint32_t* heads) {
    constexpr int kTile = mv_color_tile(NT);
    __shared__ uint32_t ldsDepth[kTile];
    __shared__ unsigned long long ldsRGN[kTile];
    __shared__ uint32_t ldsB[kTile];
    __shared__ int redbuf[5 * kMaxWaves];
    drain<NT, PPT>(views, items, nUnits, splits, claim, heads + 1,
                   [&](const MVView& V, const scr_visible_batch& vb, const DrawRecord* bb, uint32_t lo, uint32_t hi) __attribute__((always_inline)) {
                       const ColorLds L = color_lds(ldsDepth, ldsRGN, ldsB, redbuf);
                       const size_t s = (size_t)V.stride;
                       color_batch<kR0, kWave, NT, PPT, false, kTile>(V.c, V.p, vb, bb, lo, hi, V.planes, V.planes + s, V.planes + 2 * s,
                                                 V.planes + 3 * s, V.planes + 4 * s, L);
                   });
}

template <bool kR0, bool kWave, int NT, int PPT>
__global__ __launch_bounds__(NT, draw_waves(PPT)) void k_mv_nearest(const MVView* __restrict__ views,
                                                   const MVItem* __restrict__ items, uint32_t nUnits, int splits,
                                                   uint32_t claim, uint32_t* heads) {
    __shared__ unsigned long long lds[kMVNearestTile + kSpare];
    __shared__ int red[5 * kMaxWaves];
    drain<NT, PPT>(views, items, nUnits, splits, claim, heads + 2,
                   [&](const MVView& V, const scr_visible_batch& vb, const DrawRecord* bb, uint32_t lo, uint32_t hi) __attribute__((always_inline)) {
                       nearest_batch<kR0, kWave, NT, PPT, kMVNearestTile>(V.c, V.p, vb, bb, lo, hi, V.nearest, lds, red);
                   });
}

using DrawKernel = void (*)(const MVView*, const MVItem*, uint32_t, int, uint32_t, uint32_t*);

// Draw shapes (threads x points per thread) the autotuner chooses from.
struct Shape {
    int threads, ppt;
};
constexpr Shape kShapes[] = {{128, 8}, {256, 4}, {256, 8}, {256, 16}, {512, 4}, {512, 8}};
constexpr int kNumShapes = (int)(sizeof(kShapes) / sizeof(kShapes[0]));

enum Phase { kDepth = 0, kColor = 1, kNearest = 2 };

// kernels[kR0][kWave] of one (phase, shape).  scr_mv_draw.hip is compiled
// once per pair (SCR_MV_PHASE, SCR_MV_SHAPE), so the library builds the
// 18 kernel sets in parallel; each object defines its pair's accessor.
using ShapeKernels = DrawKernel[2][2];
#define SCR_MV_ACCESSOR(P, S) shape_kernels_##P##_##S
#define SCR_MV_DECLARE(P) \
    const ShapeKernels& SCR_MV_ACCESSOR(P, 0)(); \
    const ShapeKernels& SCR_MV_ACCESSOR(P, 1)(); \
    const ShapeKernels& SCR_MV_ACCESSOR(P, 2)(); \
    const ShapeKernels& SCR_MV_ACCESSOR(P, 3)(); \
    const ShapeKernels& SCR_MV_ACCESSOR(P, 4)(); \
    const ShapeKernels& SCR_MV_ACCESSOR(P, 5)();
SCR_MV_DECLARE(0)
SCR_MV_DECLARE(1)
SCR_MV_DECLARE(2)
static_assert(kNumShapes == 6, "one accessor per shape");

}  // namespace mv
}  // namespace scr
