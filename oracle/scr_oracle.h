/*
 * scr_oracle.h — CPU oracle (test infrastructure only).
 *
 * A scalar C restatement of the reference's compute-rasteriser passes and CPU
 * packer, against which the HIP path is tested.  It is linked only by the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark; the
 * product library never calls it.
 *
 * Arithmetic contract (shared by this file and the HIP kernels, written
 * independently in each):
 *  C1  IEEE-754 binary32, round-to-nearest-even, no FP contraction, every
 *      expression evaluated in the order written in the Metal/Swift source.
 *      Matrix·vector is ((c0·x + c1·y) + c2·z) + c3·w; matrix·matrix applies
 *      that to each column; length() is sqrt((x·x + y·y) + z·z).
 *  C2  In the Metal render passes a division a / b by a non-constant b is
 *      a · rcp(b) with rcp(b) = the correctly rounded 1.0f / b (Metal's
 *      default fast-math reciprocal division, made exact); a division by a
 *      constant c is a · (1.0f / c).  The CPU packer (Swift, IEEE) keeps true
 *      correctly rounded division.  sqrt is correctly rounded.
 *  C3  log2 (cull only) is scr_log2f below: a deterministic double-precision
 *      evaluation (frexp-style split + atanh series, only + - * /), rounded
 *      once to float.
 *  C4  float→u32 / float→i32 truncate toward zero and saturate; NaN → 0.
 *  C5  RGBA8 output byte = rint(saturate(x) · 255) (round half to even).
 *  C6  Equal Morton keys keep input order (stable sort).
 *
 * "Parity unpinned" functions (no reference vector can pin them, see the
 * design notes): the Metal fast-math arithmetic itself is replaced by C1-C5.
 */
#ifndef SCR_ORACLE_H
#define SCR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/scr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- arithmetic helpers (exported for pinning tests) */
float orc_log2f(float x);
float orc_expf(float x);
float orc_hash_unit(uint32_t x);
uint32_t orc_depth_to_uint(float ndcZ);
float orc_uint_to_depth(uint32_t d);
uint32_t orc_f2u(float x);
int32_t orc_f2i(float x);
void orc_mat4_mul(const float* a, const float* b, float* out);       /* out = a * b */
void orc_mat4_vec(const float* m, const float* v4, float* out4);     /* out = m * v */
void orc_transform_point(const float* m, const float p[3], float out4[4]); /* m * (p, 1), fused */
uint32_t orc_morton30(uint32_t qx, uint32_t qy, uint32_t qz);
uint64_t orc_morton63(uint32_t qx, uint32_t qy, uint32_t qz);
int orc_footprint_radius(const float p[3], const scr_raster_file* file, const float* view,
                         const float* proj, int width, int height, const scr_config* cfg);
int orc_inside_footprint(int ox, int oy, int radius);
void orc_decode_point(uint32_t pointIndex, const scr_raster_batch* b, const uint32_t* low,
                      const uint32_t* med, const uint32_t* high, int level, int fused, float out[3]);
void orc_project_q(const float* T, const float mn[3], const float step[3], const float qf[3],
                   float out4[4]); /* batch projection of quantised coordinates (C1) */
int orc_intersects_frustum(const float* m, const float mn[3], const float mx[3]);
float orc_pixel_size_on_screen(const scr_raster_batch* b, const scr_raster_file* f,
                               const float* view, const float* proj, int width, int height);
int orc_precision_level(float pixelSize);
float orc_lod_threshold(float pixelSize, int lodBias);

/* ---------------- point cloud view (pool layout) */
typedef struct orc_cloud {
    const scr_raster_batch* batches; /* nSlots entries */
    int nSlots;
    const scr_raster_file* files;    /* already composed for the camera (updateFiles) */
    int nFiles;
    const uint32_t* xyzLow;
    const uint32_t* xyzMed;
    const uint32_t* xyzHigh;
    const uint32_t* colors;
    const uint8_t* levels;
    const float* tints;              /* float4 per pool point, or NULL */
    const float* disp;               /* displacement, float4 per pool point (xyz), or NULL: zeros */
    const float* prevDisp;           /* previous frame's displacement, or NULL: `disp` */
} orc_cloud;

/* updateFiles — ComputeRasteriserPointCloud.swift:487-515 */
void orc_update_files(const scr_raster_file* in, int n, const float* viewProjection,
                      const float* model, const float* frustumOrNull, const float* prevVPOrNull,
                      scr_raster_file* out);

/* cullUpdate — Pipelines/Cull/Shaders.metal:19-83 (visible list in slot order) */
int orc_cull(const orc_cloud* c, const scr_config* cfg, const float* view, const float* proj,
             int width, int height, scr_visible_batch* visibleOut);

/* HQS frame: clear, cull, depth, colour, resolve over nClouds clouds
 * (ComputeRasteriser.swift:322-456).  Planes are width*height each; any
 * output pointer may be NULL except depth/red/green/blue/count. */
typedef struct orc_hqs_planes {
    uint32_t* depth;
    uint32_t* red;
    uint32_t* green;
    uint32_t* blue;
    uint32_t* count;
    uint32_t* rgba8;
    float* depthF32;
} orc_hqs_planes;
/* Wall-clock seconds of the phases of the process's last orc_render_hqs
 * (clear, cull, depth, colour, resolve). */
void orc_last_hqs_pass_seconds(double out[5]);
int orc_render_hqs(const orc_cloud* clouds, int nClouds, const scr_config* cfg, const float* view,
                   const float* proj, int width, int height, orc_hqs_planes* out, int nThreads);

/* Nearest mode: clearWinner, cull, nearestDepth, nearestIndex, nearestResolve
 * (ComputeRasteriser.swift:483-553) for one cloud, on nThreads threads over
 * visible-batch ranges (private planes merged by max, then min: the result
 * does not depend on nThreads). */
int orc_render_nearest(const orc_cloud* cloud, const scr_config* cfg, const float* view,
                       const float* proj, int width, int height, uint32_t* depths,
                       uint32_t* indices, uint32_t* rgba8OrNull, float* depthF32OrNull, int nThreads);

/* NearestPointCPUReference.render — NearestPointCPUReference.swift:16-81 */
void orc_nearest_cpu_reference(const orc_cloud* cloud, int width, int height, const float* view,
                               const float* proj, int enableFrustumCulling, uint32_t* depths,
                               uint32_t* indices);

/* pickPointIndex band search — ComputeRasteriser.swift:617-669 */
int orc_pick_search(const uint32_t* indices, int width, int height, float ndcX, float ndcY,
                    int searchRadius, uint32_t* outIndex);

/* ---------------- CPU pack — PackedPointCloudFixtures.swift:30-344 */
typedef struct orc_pack_out {
    scr_raster_batch* batches; /* ceil(count / ppb) */
    uint32_t* xyzLow;
    uint32_t* xyzMed;
    uint32_t* xyzHigh;
    uint32_t* colors;
    uint8_t* levels;
    uint32_t* sourceIndices;
    float* orderedPositions;   /* 3 floats per point, may be NULL */
    float boundsMin[3];
    float boundsMax[3];
} orc_pack_out;
/* positions: posStride floats per point (3 or 4); colors: 4 floats per point.
 * keyBits 30 (reference) or 63.  Returns number of batches (or < 0). */
int orc_pack(const float* positions, int posStride, const float* colors, size_t count,
             int pointsPerBatch, int lodLevels, int coarseVoxelDivisions, int keyBits,
             orc_pack_out* out);
/* GPUPacker's LOD gate (GPUPacker.swift:35, 86-99): 1 when the dense grid's
 * cell bound (coarse * 2^(L - 2) + 2)^3 is within 268,435,456, else 0 (the GPU
 * packer then leaves every level at L - 1). */
int orc_gpu_lod_enabled(int lodLevels, int coarseVoxelDivisions);
/* orc_pack under that gate: what GPUPacker.pack produces. */
int orc_pack_gpu(const float* positions, int posStride, const float* colors, size_t count,
                 int pointsPerBatch, int lodLevels, int coarseVoxelDivisions, int keyBits,
                 orc_pack_out* out);
/* Morton key per point (the sort key the packer uses), keyBits 30 or 63. */
void orc_morton_keys(const float* positions, int posStride, size_t count, int keyBits,
                     uint64_t* keys);
/* computeLODLevels on already Morton-ordered positions (3 floats/pt). */
void orc_lod_levels(const float* sortedPositions, size_t count, const float bmin[3],
                    const float bmax[3], int lodLevels, int coarseVoxelDivisions, uint8_t* levels);

/* cubeGrid(pointsPerAxis:) — PackedPointCloudFixtures.swift:5-25, with the
 * lattice scaled to [-extent/2, extent/2]^3 (extent 1 = reference). */
void orc_cube_grid(int pointsPerAxis, float extent, float* positions3, float* colors4);

/* ---------------- differentiable HQS (no reference counterpart)
 * Colour of point p = params[4p .. 4p+2] (float), accumulated per pixel as
 * int64 fixed point rint(param * 2^16) over the fragments the HQS colour
 * pass keeps; image = sum / (2^16 * count) rounded to float, background where
 * count == 0.  depth / count equal the HQS planes.  Buffers: depth, count
 * W*H; sums, image 3 per pixel (interleaved, row 0 = top). */
int64_t orc_diff_fixed(float v);
float orc_diff_pixel(int64_t sum, uint32_t count, float background);
int orc_render_diff(const orc_cloud* cloud, const scr_config* cfg, const float* view, const float* proj, int width,
                    int height, const float* params, uint32_t* depth, uint32_t* count, int64_t* sums, float* image,
                    int nThreads);
/* grad[4p + k] += sum over p's kept fragments of dLdImage[3 * pixel + k] * (1 / count[pixel])
 * in FP64, for points with mask[p] != 0 (mask NULL: all points). */
int orc_diff_backward(const orc_cloud* cloud, const scr_config* cfg, const float* view, const float* proj, int width,
                      int height, const uint32_t* depth, const uint32_t* count, const float* dLdImage,
                      double* grad, const uint8_t* mask);

#ifdef __cplusplus
}
#endif
#endif
