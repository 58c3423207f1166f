// A user displacement kernel written against scr_device.h, built with
// tint_hook.hip by build_hooks.py into libuser_hooks.so for the hook tests:
// every point is lifted by a quarter of its x and pushed along y by its slot
// offset; every 17th point of a batch gets a NaN displacement (removed from
// the render).  The tests restate it on the CPU.
#include "scr_device.h"

__global__ void displace_wave(scr_cloud_view v) {
    scr_raster_batch b;
    uint32_t i, local;
    if (!scr_resolve_displacement_thread(blockIdx.x * blockDim.x + threadIdx.x, &v, &b, &i, &local)) return;
    const float3 p = scr_decode_point_at(i, &b, &v);
    const float nan = __builtin_nanf("");
    const float4 d = (local % 17u == 0u) ? make_float4(0.0f, nan, 0.0f, 0.0f)
                                         : make_float4(0.0f, (float)local * (1.0f / 4096.0f), p.x * 0.25f, 0.0f);
    reinterpret_cast<float4*>(v.displacements)[i] = d;
}

extern "C" int displace_hook_run(const scr_cloud_view* v) {
    if (!v || !v->displacements || v->pointsPerBatch == 0u) return -1;
    const uint64_t n = (uint64_t)v->slots * v->pointsPerBatch;
    hipLaunchKernelGGL(displace_wave, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)v->stream, *v);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
