/*
 * scr_device.h — device helpers for user HIP kernels over a cloud's pool.
 *
 * The reference lets users write TintPass / DisplacementPass kernels in
 * Metal, prefixed with a preamble of helpers (TintPass.swift:215-338,
 * DisplacementPass.swift:120-172).  Here a user kernel is ordinary HIP code,
 * built with hipcc --offload-arch=gfx950, that includes this header, takes the
 * scr_cloud_view of scr_cloud_device_view and runs on view.stream over
 * view.slots * view.pointsPerBatch threads, one per pool point:
 *
 *   __global__ void my_tint(scr_cloud_view v) {
 *       scr_raster_batch b;
 *       uint32_t i, local;
 *       if (!scr_resolve_tint_thread(blockIdx.x * blockDim.x + threadIdx.x, &v, &b, &i, &local)) return;
 *       const float3 p = scr_decode_point_at(i, &b, &v);
 *       reinterpret_cast<float4*>(v.tints)[i] = make_float4(1.0f, 0.0f, 0.0f, p.z > 10.0f ? 0.5f : 0.0f);
 *   }
 *
 * Tints are read by a render whose configuration sets applyTint: the colour
 * is mixed toward saturate(tint.rgb) by saturate(tint.a); a negative alpha
 * drops the point (ColorPass/Shaders.metal:95-133).
 *
 * A displacement kernel (DisplacementPass.swift:120-172) is written the same
 * way against the view of scr_cloud_device_view(cloud, SCR_VIEW_DISPLACEMENTS,
 * &v): it resolves its point with scr_resolve_displacement_thread and writes
 * reinterpret_cast<float4*>(v.displacements)[i] = make_float4(dx, dy, dz, 0).
 * A render with applyDisplacement draws position + displacement in the
 * depth, colour and weighted-blend passes; a NaN component removes the point
 * from all of them (DepthPass/Shaders.metal:69-74, ColorPass/Shaders.metal:72-78).
 */
#ifndef SCR_DEVICE_H
#define SCR_DEVICE_H

#include <hip/hip_runtime.h>

#include "scr.h"

/* scr_resolveTintThread: the slot of thread `id` and the pool index of its
 * point; false for a thread past the pool, an empty slot, or a lane past the
 * batch's numPoints. */
__device__ inline bool scr_resolve_tint_thread(uint32_t id, const scr_cloud_view* v, scr_raster_batch* batch,
                                               uint32_t* pointIndex, uint32_t* localOffset) {
    const uint32_t slot = id / v->pointsPerBatch;
    *localOffset = id - slot * v->pointsPerBatch;
    if (slot >= v->slots) return false;
    *batch = v->batches[slot];
    if (batch->state == 0 || *localOffset >= batch->numPoints) return false;
    *pointIndex = batch->firstPoint + *localOffset;
    return true;
}

/* scr_resolveDisplacementThread (DisplacementPass.swift:305-318): the same
 * thread-to-point mapping as the tint preamble's. */
__device__ inline bool scr_resolve_displacement_thread(uint32_t id, const scr_cloud_view* v, scr_raster_batch* batch,
                                                       uint32_t* pointIndex, uint32_t* localOffset) {
    return scr_resolve_tint_thread(id, v, batch, pointIndex, localOffset);
}

/* scr_decodePointAt: the point's position at the precision its level byte
 * selects (0: 30-bit words, 1: 20-bit, otherwise 10-bit), as the reference's
 * preamble chooses it, with the draw passes' arithmetic: q * (size * 2^-bits)
 * + min as one fused multiply-add per axis, size clamped below at 1e-6. */
__device__ inline float3 scr_decode_point_at(uint32_t pointIndex, const scr_raster_batch* b, const scr_cloud_view* v) {
    const uint32_t level = (uint32_t)v->levels[pointIndex] & 7u;
    const uint32_t lo = v->xyzLow[pointIndex];
    uint32_t qx, qy, qz;
    float inv;
    if (level == 0u) {
        const uint32_t me = v->xyzMed[pointIndex], hi = v->xyzHigh[pointIndex];
        qx = ((lo & 1023u) << 20) | ((me & 1023u) << 10) | (hi & 1023u);
        qy = (((lo >> 10) & 1023u) << 20) | (((me >> 10) & 1023u) << 10) | ((hi >> 10) & 1023u);
        qz = (((lo >> 20) & 1023u) << 20) | (((me >> 20) & 1023u) << 10) | ((hi >> 20) & 1023u);
        inv = 1.0f / 1073741824.0f;
    } else if (level == 1u) {
        const uint32_t me = v->xyzMed[pointIndex];
        qx = ((lo & 1023u) << 10) | (me & 1023u);
        qy = (((lo >> 10) & 1023u) << 10) | ((me >> 10) & 1023u);
        qz = (((lo >> 20) & 1023u) << 10) | ((me >> 20) & 1023u);
        inv = 1.0f / 1048576.0f;
    } else {
        qx = lo & 1023u;
        qy = (lo >> 10) & 1023u;
        qz = (lo >> 20) & 1023u;
        inv = 1.0f / 1024.0f;
    }
    const float sx = fmaxf(b->maxX - b->minX, 0.000001f) * inv;
    const float sy = fmaxf(b->maxY - b->minY, 0.000001f) * inv;
    const float sz = fmaxf(b->maxZ - b->minZ, 0.000001f) * inv;
    return make_float3(fmaf((float)qx, sx, b->minX), fmaf((float)qy, sy, b->minY), fmaf((float)qz, sz, b->minZ));
}

/* scr_decodeColorAt: the point's packed RGBA8 colour as 0..1 floats (each
 * byte times the binary32 value of 1/255). */
__device__ inline float4 scr_decode_color_at(uint32_t pointIndex, const scr_cloud_view* v) {
    const uint32_t c = v->colors[pointIndex];
    const float k = 1.0f / 255.0f;
    return make_float4((float)(c & 0xffu) * k, (float)((c >> 8) & 0xffu) * k, (float)((c >> 16) & 0xffu) * k,
                       (float)(c >> 24) * k);
}

#endif
