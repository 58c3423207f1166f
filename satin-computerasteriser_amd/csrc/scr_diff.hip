// scr_diff.hip — differentiable HQS: the HQS passes with per-point float
// colours (parameters) accumulated as 64-bit fixed point, and a backward pass
// that recomputes each point's fragments from the saved depth / count planes.
//
//   forward   clear → cull → depth (the HQS kernel) → k_diff_color → k_diff_resolve
//             sum_k(pixel) = Σ rint(param_k · 2^16) over the fragments the HQS
//             colour pass keeps (ColorPass/Shaders.metal:201-214 visibility);
//             image = sum / (2^16 · count) in FP64, rounded to float
//   backward  cull → k_diff_backward: dL/dparam_k(p) += Σ_fragments dL/dimage_k · (1 / count)
//             (rounding passes the gradient straight through), FP64, one thread
//             per point (no atomics, so the result does not depend on timing)
//
// Integer sums make the forward deterministic; the backward adds per point in
// the fixed footprint order, and views accumulate in call order.
#include <cstring>

#include "scr_diff_draw.hpp"

namespace scr {
namespace {

using namespace draw;

constexpr int kThreads = kDiffThreads;

template <bool kR0>
__global__ __launch_bounds__(kThreads) void k_diff_color(CloudDev c, PassParams p, const float4* __restrict__ params,
                                                         const uint32_t* __restrict__ depth,
                                                         unsigned long long* sums, size_t stride, uint32_t* count) {
    __shared__ DiffLds L;
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, p, c.visible[v], c.bounds + v, 0xffffffffu);
        if (!kR0 && bc.rconst == 0) diff_color_chunks<true>(c, p, bc, 0u, params, depth, sums, stride, count, L);
        else diff_color_chunks<kR0>(c, p, bc, 0u, params, depth, sums, stride, count, L);
    }
}

__global__ __launch_bounds__(256) void k_diff_resolve(PassParams p, const unsigned long long* __restrict__ sums,
                                                      size_t stride, const uint32_t* __restrict__ count,
                                                      float* __restrict__ image) {
    const size_t n = (size_t)p.width * (size_t)p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    diff_resolve_pixel(p, sums, stride, count, i, image);
}

template <bool kR0>
__global__ __launch_bounds__(kThreads) void k_diff_backward(CloudDev c, PassParams p,
                                                            const uint32_t* __restrict__ depth,
                                                            const uint32_t* __restrict__ count,
                                                            const float* __restrict__ dLdImage,
                                                            double* __restrict__ grad) {
    const uint32_t nvis = *c.counter;
    for (uint32_t v = blockIdx.x; v < nvis; v += gridDim.x) {
        const BatchCtx bc = batch_ctx(c, p, c.visible[v], c.bounds + v, 0xffffffffu);
        if (!kR0 && bc.rconst == 0) diff_backward_chunks<true>(c, p, bc, 0u, depth, count, dLdImage, grad);
        else diff_backward_chunks<kR0>(c, p, bc, 0u, depth, count, dLdImage, grad);
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
