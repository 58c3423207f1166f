// scr_diff_draw.hpp — device code of the differentiable HQS passes, shared by
// the single-view kernels (scr_diff.hip) and the batched views of
// scr_multiview.hip.
//
// Colour of pool point p = params[p].xyz, accumulated per pixel as int64
// rint(param * 2^16) over the fragments the HQS colour pass keeps
// (ColorPass/Shaders.metal:201-214 visibility).  The backward recomputes each
// point's fragments and adds dL/dimage x (1 / count) per fragment in FP64; one
// thread owns a point, so a point's sum needs no atomics.
#pragma once

#include "scr_draw.hpp"

namespace scr {
namespace draw {

constexpr int kDiffThreads = 256;  // diff kernels: workgroup size
constexpr int kDiffPPT = 8;        // points per thread
constexpr int kDiffTile = 1024;  // depth u32 + 3 sums u64 + count u32 = 32 B per pixel: 32 KB

__device__ __forceinline__ unsigned long long diff_fixed(float v) {
    float x = v * 65536.0f;
    if (!(x == x)) return 0ull;
    x = fminf(fmaxf(x, -9.0e18f), 9.0e18f);
    return (unsigned long long)(long long)rintf(x);
}

struct DiffLds {
    uint32_t depth[kDiffTile];
    unsigned long long s[3][kDiffTile];
    uint32_t n[kDiffTile];
    int red[5 * kMaxWaves];
};

// A chunk's fixed-point sums (the routes of depth_frags: kCheck tests each
// fragment against a batch rectangle, otherwise the box is the chunk's own).
template <bool kR0, bool kCheck>
__device__ __forceinline__ void diff_frags(const ChunkFrags<kDiffPPT>& f, uint32_t base, const ChunkBox& box,
                                           const PassParams& p, const float4* __restrict__ params,
                                           const uint32_t* __restrict__ depth, unsigned long long* sums,
                                           size_t stride, uint32_t* count, DiffLds& L) {
#pragma unroll
    for (int j = 0; j < kDiffPPT; j++) {
        if (!f.ok[j]) continue;
        const float4 t = gload_rec(params + base + (uint32_t)j);
        const unsigned long long q0 = diff_fixed(t.x), q1 = diff_fixed(t.y), q2 = diff_fixed(t.z);
        const int rad = kR0 ? 0 : f.r[j];
        for (int oy = -rad; oy <= rad; oy++)
            for (int ox = -rad; ox <= rad; ox++) {
                if (!kR0 && !inside_footprint(ox, oy, rad)) continue;
                const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                if (!kR0 && (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height)) continue;
                const size_t pix = (size_t)ty * p.width + tx;
                const bool inTile = box.lds && (!kCheck || box_holds(box, tx, ty));
                const int k = inTile ? box.local(tx, ty) : 0;
                const uint32_t d = inTile ? lds_load(L.depth + k) : gload(depth + pix);
                if (d == 0u) continue;
                if (!(f.z[j] >= uint_to_depth(d) * p.depthKeep)) continue;
                if (inTile) {
                    lds_add(&L.s[0][k], q0);
                    lds_add(&L.s[1][k], q1);
                    lds_add(&L.s[2][k], q2);
                    lds_add(&L.n[k], 1u);
                } else {
                    gadd(sums + pix, q0);
                    gadd(sums + stride + pix, q1);
                    gadd(sums + 2 * stride + pix, q2);
                    gadd(count + pix, 1u);
                }
            }
    }
}

__device__ __forceinline__ void diff_stage(const ChunkBox& box, const PassParams& p,
                                           const uint32_t* __restrict__ depth, DiffLds& L) {
    const int area = box.w * box.h;
    for (int i = threadIdx.x; i < area; i += kDiffThreads) {
        lds_store(&L.depth[i], gload(depth + box.pixel(i, p.width)));
        lds_store(&L.s[0][i], 0ull);
        lds_store(&L.s[1][i], 0ull);
        lds_store(&L.s[2][i], 0ull);
        lds_store(&L.n[i], 0u);
    }
}

__device__ __forceinline__ void diff_tile_flush(const ChunkBox& box, const PassParams& p, unsigned long long* sums,
                                                size_t stride, uint32_t* count, const DiffLds& L) {
    const int area = box.w * box.h;
    for (int i = threadIdx.x; i < area; i += kDiffThreads) {
        const uint32_t n = lds_load(&L.n[i]);
        if (n == 0u) continue;
        const size_t pix = box.pixel(i, p.width);
        gadd(sums + pix, lds_load(&L.s[0][i]));
        gadd(sums + stride + pix, lds_load(&L.s[1][i]));
        gadd(sums + 2 * stride + pix, lds_load(&L.s[2][i]));
        gadd(count + pix, n);
    }
}

// Local points [lo, bc.active) of one visible batch (lo a multiple of the
// chunk): one tile for the whole range when the batch rectangle fits it,
// else one per chunk.
template <bool kR0, int kLevel>
__device__ __forceinline__ void diff_color_chunks_level(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t lo,
                                         const float4* __restrict__ params, const uint32_t* __restrict__ depth,
                                         unsigned long long* sums, size_t stride, uint32_t* count, DiffLds& L,
                                              const ChunkBox& seg) {
    constexpr int kChunk = kDiffThreads * kDiffPPT;
    const bool vec = vec_ok<kDiffPPT>(bc);
    if (seg.lds) {
        diff_stage(seg, p, depth, L);
        __syncthreads();
        for (uint32_t cb = lo; cb < bc.active; cb += kChunk) {
            const uint32_t li0 = cb + threadIdx.x * kDiffPPT;
            ChunkFrags<kDiffPPT> f;
            chunk_front<false, kR0, kLevel, kDiffPPT>(c, p, bc, li0, vec, f);
            diff_frags<kR0, true>(f, bc.first + li0, seg, p, params, depth, sums, stride, count, L);
        }
        __syncthreads();
        diff_tile_flush(seg, p, sums, stride, count, L);
        __syncthreads();
        return;
    }
    for (uint32_t cb = lo; cb < bc.active; cb += kChunk) {
        const uint32_t li0 = cb + threadIdx.x * kDiffPPT;
        ChunkFrags<kDiffPPT> f;
        chunk_front<false, kR0, kLevel, kDiffPPT>(c, p, bc, li0, vec, f);
        const ChunkBox box = chunk_box<false, kDiffThreads, kDiffPPT>(f, p, kDiffTile, L.red);
        if (box.frags == 0) continue;
        if (box.lds) {
            diff_stage(box, p, depth, L);
            __syncthreads();
        }
        diff_frags<kR0, false>(f, bc.first + li0, box, p, params, depth, sums, stride, count, L);
        if (box.lds) {
            __syncthreads();
            diff_tile_flush(box, p, sums, stride, count, L);
            __syncthreads();
        }
    }
}

template <bool kR0>
__device__ __forceinline__ void diff_color_chunks(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t lo,
                                         const float4* __restrict__ params, const uint32_t* __restrict__ depth,
                                         unsigned long long* sums, size_t stride, uint32_t* count, DiffLds& L) {
    const ChunkBox seg = segment_box(p, bc, bc.active > lo ? bc.active - lo : 0u, kDiffTile);
    if (bc.level == 0) diff_color_chunks_level<kR0, 0>(c, p, bc, lo, params, depth, sums, stride, count, L, seg);
    else if (bc.level == 1) diff_color_chunks_level<kR0, 1>(c, p, bc, lo, params, depth, sums, stride, count, L, seg);
    else diff_color_chunks_level<kR0, 2>(c, p, bc, lo, params, depth, sums, stride, count, L, seg);
}

// The gradient of the thread's PPT points starting at local index li0 under
// one view, added to acc (loaded from grad when a point is first touched;
// touched marks the points to write back): per fragment that passes the
// colour pass's depth test, dL/dpixel x (1 / count) in FP64, summed over the
// point's footprint and then added to the point's running sum.
template <bool kR0>
__device__ __forceinline__ void diff_backward_acc(const CloudDev& c, const PassParams& p, const BatchCtx& bc,
                                                  uint32_t li0, bool vec, const uint32_t* __restrict__ depth,
                                                  const uint32_t* __restrict__ count,
                                                  const float* __restrict__ dLdImage, const double* grad,
                                                  double (&acc)[kDiffPPT][3], bool (&touched)[kDiffPPT]) {
    ChunkFrags<kDiffPPT> f;
    if (bc.level == 0) chunk_front<false, kR0, 0, kDiffPPT>(c, p, bc, li0, vec, f);
    else if (bc.level == 1) chunk_front<false, kR0, 1, kDiffPPT>(c, p, bc, li0, vec, f);
    else chunk_front<false, kR0, 2, kDiffPPT>(c, p, bc, li0, vec, f);
#pragma unroll
    for (int j = 0; j < kDiffPPT; j++) {
        if (!f.ok[j]) continue;
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        bool any = false;
        const int rad = kR0 ? 0 : f.r[j];
        for (int oy = -rad; oy <= rad; oy++)
            for (int ox = -rad; ox <= rad; ox++) {
                if (!kR0 && !inside_footprint(ox, oy, rad)) continue;
                const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                if (!kR0 && (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height)) continue;
                const size_t pix = (size_t)ty * p.width + tx;
                const uint32_t d = gload(depth + pix);
                if (d == 0u) continue;
                if (!(f.z[j] >= uint_to_depth(d) * p.depthKeep)) continue;
                const double inv = 1.0 / (double)gload(count + pix);
                g0 += (double)gload(dLdImage + pix * 3 + 0) * inv;
                g1 += (double)gload(dLdImage + pix * 3 + 1) * inv;
                g2 += (double)gload(dLdImage + pix * 3 + 2) * inv;
                any = true;
            }
        if (!any) continue;
        if (!touched[j]) {
            const double* gp = grad + (size_t)(bc.first + li0 + (uint32_t)j) * 4;
            acc[j][0] = gload(gp);
            acc[j][1] = gload(gp + 1);
            acc[j][2] = gload(gp + 2);
            touched[j] = true;
        }
        acc[j][0] += g0;
        acc[j][1] += g1;
        acc[j][2] += g2;
    }
}

// Writes back the touched points' sums.
__device__ __forceinline__ void diff_backward_store(uint32_t first, const double (&acc)[kDiffPPT][3],
                                                    const bool (&touched)[kDiffPPT], double* grad) {
#pragma unroll
    for (int j = 0; j < kDiffPPT; j++) {
        if (!touched[j]) continue;
        SCR_GLOBAL double* gp = gptr(grad + (size_t)(first + (uint32_t)j) * 4);
        gp[0] = acc[j][0];
        gp[1] = acc[j][1];
        gp[2] = acc[j][2];
    }
}

// Local points [lo, bc.active) of one visible batch (lo a multiple of the
// chunk) under one view: each point's gradient gets this view's sum added.
template <bool kR0>
__device__ __forceinline__ void diff_backward_chunks(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t lo,
                                            const uint32_t* __restrict__ depth, const uint32_t* __restrict__ count,
                                            const float* __restrict__ dLdImage, double* __restrict__ grad) {
    constexpr int kChunk = kDiffThreads * kDiffPPT;
    const bool vec = vec_ok<kDiffPPT>(bc);
    for (uint32_t cb = lo; cb < bc.active; cb += kChunk) {
        const uint32_t li0 = cb + threadIdx.x * kDiffPPT;
        double acc[kDiffPPT][3];
        bool touched[kDiffPPT];
#pragma unroll
        for (int j = 0; j < kDiffPPT; j++) touched[j] = false;
        diff_backward_acc<kR0>(c, p, bc, li0, vec, depth, count, dLdImage, grad, acc, touched);
        diff_backward_store(bc.first + li0, acc, touched, grad);
    }
}

// Pixel i of the differentiable image: sum / (2^16 * count) in FP64, or the
// background where count is 0.
__device__ __forceinline__ void diff_resolve_pixel(const PassParams& p, const unsigned long long* __restrict__ sums,
                                          size_t stride, const uint32_t* __restrict__ count, size_t i,
                                          float* __restrict__ image) {
    const uint32_t cnt = count[i];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float v = p.bg[k];
        if (cnt != 0u) v = (float)((double)(long long)sums[(size_t)k * stride + i] / (65536.0 * (double)cnt));
        image[i * 3 + (size_t)k] = v;
    }
}

}  // namespace draw
}  // namespace scr
