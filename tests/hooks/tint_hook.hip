// A user tint kernel written against scr_device.h, built by build_hooks.py
// into libuser_hooks.so for the hook tests: points of one height band are
// dropped, higher points take their inverted colour, lower ones a tint from
// their x and slot offset.  The tests restate it on the CPU.
#include "scr_device.h"

__global__ void tint_by_height(scr_cloud_view v, float lo, float hi) {
    scr_raster_batch b;
    uint32_t i, local;
    if (!scr_resolve_tint_thread(blockIdx.x * blockDim.x + threadIdx.x, &v, &b, &i, &local)) return;
    const float3 p = scr_decode_point_at(i, &b, &v);
    const float4 c = scr_decode_color_at(i, &v);
    float4 t;
    if (p.y > lo && p.y < hi)
        t = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    else if (p.y >= hi)
        t = make_float4(1.0f - c.x, 1.0f - c.y, 1.0f - c.z, 0.75f);
    else
        t = make_float4(p.x * 0.1f, 0.5f, (float)local * (1.0f / 1024.0f), 0.5f);
    reinterpret_cast<float4*>(v.tints)[i] = t;
}

extern "C" int tint_hook_run(const scr_cloud_view* v, float lo, float hi) {
    if (!v || !v->tints || v->pointsPerBatch == 0u) return -1;
    const uint64_t n = (uint64_t)v->slots * v->pointsPerBatch;
    hipLaunchKernelGGL(tint_by_height, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)v->stream, *v,
                       lo, hi);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
