// scr_diff.hip — differentiable HQS: the HQS passes with per-point float
// colours (parameters) accumulated as 64-bit fixed point, and a backward pass
// that recomputes each point's fragments from the saved depth / count planes.
//
//   forward   clear → cull → depth (the HQS kernel) → k_diff_color → k_diff_resolve
//             sum_k(pixel) = Σ rint(param_k · 2^16) over the fragments the HQS
//             colour pass keeps (ColorPass/Shaders.metal:201-214 visibility);
//             image = sum / (2^16 · count) in FP64, rounded to float
//   backward  cull → k_diff_backward: dL/dparam_k(p) += Σ_fragments dL/dimage_k / count
//             (rounding passes the gradient straight through), FP64, one thread
//             per point (no atomics, so the result does not depend on timing)
//
// Integer sums make the forward deterministic; the backward adds per point in
// the fixed footprint order, and views accumulate in call order.
#include <cstring>

#include "scr_draw.hpp"

namespace scr {
namespace {

using namespace draw;

constexpr int kThreads = 256;
constexpr int kPPT = 8;
constexpr int kDiffTile = 1024;  // depth u32 + 3 sums u64 + count u32 = 32 B per pixel: 32 KB

__device__ inline unsigned long long diff_fixed(float v) {
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

template <bool kR0>
__device__ inline void diff_color_chunks(const CloudDev& c, const PassParams& p, const BatchCtx& bc,
                                         const float4* __restrict__ params, const uint32_t* __restrict__ depth,
                                         unsigned long long* sums, size_t stride, uint32_t* count, DiffLds& L) {
    constexpr int kChunk = kThreads * kPPT;
    const bool vec = vec_ok<kPPT>(bc);
    for (uint32_t cb = 0; cb < bc.active; cb += kChunk) {
        const uint32_t li0 = cb + threadIdx.x * kPPT;
        ChunkFrags<kPPT> f;
        chunk_front<false, kR0, kPPT>(c, p, bc, li0, vec, f);
        const ChunkBox box = chunk_box<false, kThreads, kPPT>(f, p, kDiffTile, L.red);
        if (box.frags == 0) continue;
        const int area = box.w * box.h;
        if (box.lds) {
            for (int i = threadIdx.x; i < area; i += kThreads) {
                L.depth[i] = depth[box.pixel(i, p.width)];
                L.s[0][i] = L.s[1][i] = L.s[2][i] = 0ull;
                L.n[i] = 0u;
            }
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < kPPT; j++) {
            if (f.px[j] < 0) continue;
            const float4 t = params[bc.first + li0 + (uint32_t)j];
            const unsigned long long q0 = diff_fixed(t.x), q1 = diff_fixed(t.y), q2 = diff_fixed(t.z);
            const int rad = kR0 ? 0 : f.r[j];
            for (int oy = -rad; oy <= rad; oy++)
                for (int ox = -rad; ox <= rad; ox++) {
                    if (!kR0 && !inside_footprint(ox, oy, rad)) continue;
                    const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                    if (!kR0 && (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height)) continue;
                    const size_t pix = (size_t)ty * p.width + tx;
                    const int k = box.lds ? box.local(tx, ty) : 0;
                    const uint32_t d = box.lds ? L.depth[k] : depth[pix];
                    if (d == 0u) continue;
                    if (!(f.z[j] >= uint_to_depth(d) * p.depthKeep)) continue;
                    if (box.lds) {
                        atomicAdd(&L.s[0][k], q0);
                        atomicAdd(&L.s[1][k], q1);
                        atomicAdd(&L.s[2][k], q2);
                        atomicAdd(&L.n[k], 1u);
                    } else {
                        atomicAdd(sums + pix, q0);
                        atomicAdd(sums + stride + pix, q1);
                        atomicAdd(sums + 2 * stride + pix, q2);
                        atomicAdd(count + pix, 1u);
                    }
                }
        }
        if (box.lds) {
            __syncthreads();
            for (int i = threadIdx.x; i < area; i += kThreads) {
                const uint32_t n = L.n[i];
                if (n == 0u) continue;
                const size_t pix = box.pixel(i, p.width);
                atomicAdd(sums + pix, L.s[0][i]);
                atomicAdd(sums + stride + pix, L.s[1][i]);
                atomicAdd(sums + 2 * stride + pix, L.s[2][i]);
                atomicAdd(count + pix, n);
            }
            __syncthreads();
        }
    }
}

template <bool kR0>
__global__ __launch_bounds__(kThreads) void k_diff_color(CloudDev c, PassParams p, const float4* __restrict__ params,
                                                         const uint32_t* __restrict__ depth,
                                                         unsigned long long* sums, size_t stride, uint32_t* count) {
    __shared__ DiffLds L;
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, p, c.visible[v], 0xffffffffu);
        if (!kR0 && bc.rconst == 0) diff_color_chunks<true>(c, p, bc, params, depth, sums, stride, count, L);
        else diff_color_chunks<kR0>(c, p, bc, params, depth, sums, stride, count, L);
    }
}

__global__ __launch_bounds__(256) void k_diff_resolve(PassParams p, const unsigned long long* __restrict__ sums,
                                                      size_t stride, const uint32_t* __restrict__ count,
                                                      float* __restrict__ image) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t cnt = count[i];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float v = p.bg[k];
        if (cnt != 0u) v = (float)((double)(long long)sums[(size_t)k * stride + i] / (65536.0 * (double)cnt));
        image[i * 3 + (size_t)k] = v;
    }
}

template <bool kR0>
__device__ inline void diff_backward_chunks(const CloudDev& c, const PassParams& p, const BatchCtx& bc,
                                            const uint32_t* __restrict__ depth, const uint32_t* __restrict__ count,
                                            const float* __restrict__ dLdImage, double* __restrict__ grad) {
    constexpr int kChunk = kThreads * kPPT;
    const bool vec = vec_ok<kPPT>(bc);
    for (uint32_t cb = 0; cb < bc.active; cb += kChunk) {
        const uint32_t li0 = cb + threadIdx.x * kPPT;
        ChunkFrags<kPPT> f;
        chunk_front<false, kR0, kPPT>(c, p, bc, li0, vec, f);
#pragma unroll
        for (int j = 0; j < kPPT; j++) {
            if (f.px[j] < 0) continue;
            double g0 = 0.0, g1 = 0.0, g2 = 0.0;
            bool any = false;
            const int rad = kR0 ? 0 : f.r[j];
            for (int oy = -rad; oy <= rad; oy++)
                for (int ox = -rad; ox <= rad; ox++) {
                    if (!kR0 && !inside_footprint(ox, oy, rad)) continue;
                    const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                    if (!kR0 && (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height)) continue;
                    const size_t pix = (size_t)ty * p.width + tx;
                    const uint32_t d = depth[pix];
                    if (d == 0u) continue;
                    if (!(f.z[j] >= uint_to_depth(d) * p.depthKeep)) continue;
                    const double n = (double)count[pix];
                    g0 += (double)dLdImage[pix * 3 + 0] / n;
                    g1 += (double)dLdImage[pix * 3 + 1] / n;
                    g2 += (double)dLdImage[pix * 3 + 2] / n;
                    any = true;
                }
            if (!any) continue;
            double* gp = grad + (size_t)(bc.first + li0 + (uint32_t)j) * 4;
            gp[0] += g0;
            gp[1] += g1;
            gp[2] += g2;
        }
    }
}

template <bool kR0>
__global__ __launch_bounds__(kThreads) void k_diff_backward(CloudDev c, PassParams p,
                                                            const uint32_t* __restrict__ depth,
                                                            const uint32_t* __restrict__ count,
                                                            const float* __restrict__ dLdImage,
                                                            double* __restrict__ grad) {
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, p, c.visible[v], 0xffffffffu);
        if (!kR0 && bc.rconst == 0) diff_backward_chunks<true>(c, p, bc, depth, count, dLdImage, grad);
        else diff_backward_chunks<kR0>(c, p, bc, depth, count, dLdImage, grad);
    }
}

unsigned grid_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

unsigned draw_grid(int numCUs, int nSlots) {
    const int g = numCUs * 8;
    return (unsigned)(nSlots < g ? (nSlots > 0 ? nSlots : 1) : g);
}

}  // namespace

hipError_t launch_diff_color(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, const float4* params,
                             const uint32_t* depth, unsigned long long* sums, size_t stride, uint32_t* count) {
    if (p.constRadius == 0)
        hipLaunchKernelGGL(k_diff_color<true>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kThreads), 0, s, c, p, params,
                           depth, sums, stride, count);
    else
        hipLaunchKernelGGL(k_diff_color<false>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kThreads), 0, s, c, p,
                           params, depth, sums, stride, count);
    return hipGetLastError();
}

hipError_t launch_diff_resolve(hipStream_t s, const PassParams& p, const unsigned long long* sums, size_t stride,
                               const uint32_t* count, float* image) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_diff_resolve, dim3(grid_for(n, 256)), dim3(256), 0, s, p, sums, stride, count, image);
    return hipGetLastError();
}

hipError_t launch_diff_backward(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p,
                                const uint32_t* depth, const uint32_t* count, const float* dLdImage, double* grad) {
    if (p.constRadius == 0)
        hipLaunchKernelGGL(k_diff_backward<true>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kThreads), 0, s, c, p,
                           depth, count, dLdImage, grad);
    else
        hipLaunchKernelGGL(k_diff_backward<false>, dim3(draw_grid(numCUs, c.nSlots)), dim3(kThreads), 0, s, c, p,
                           depth, count, dLdImage, grad);
    return hipGetLastError();
}

}  // namespace scr
