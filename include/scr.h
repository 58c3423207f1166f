/*
 * scr.h — C ABI of the MI355X compute point-cloud rasteriser (libscr.so).
 *
 * Every entry point replaces a piece of the Swift API of
 * mnmly/Satin-ComputeRasteriser; the reference interface each one stands in
 * for is cited beside it (paths relative to the reference repository root,
 * under Sources/SatinComputeRasteriser/ unless stated otherwise).
 *
 * Conventions
 *  - Plain C types only; matrices are float[16], column-major (m[c*4 + r]),
 *    the same memory layout as simd_float4x4 / Metal float4x4.
 *  - Every function returns an int status: SCR_OK (0) or a negative SCR_ERR_*
 *    code; scr_last_error() returns the message of the last failure on a
 *    context.  Reference preconditions (traps) become error codes.
 *  - Host arrays passed in are copied before the call returns (the reference
 *    memcpy's caller data into its pool, ComputeRasteriserPointCloud.swift:646-664);
 *    a flag selects device-resident inputs where noted.
 *  - One HIP stream per context.  Render/pack calls enqueue work and return;
 *    scr_sync and the read functions synchronise.  A context is not
 *    thread-safe (the reference is driven from one render thread).
 */
#ifndef SCR_H
#define SCR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants: Pipelines/ComputeRasteriserTypes.h:4-9, ComputeRasteriserTypes.swift:4-6 */
#define SCR_THREADS_PER_GROUP 128
#define SCR_DEFAULT_POINTS_PER_BATCH 10240 /* computeRasteriserThreadsPerGroup * 80 */
#define SCR_MAX_POINTS_PER_BATCH 65535     /* uint16 LOD prefix counts, PackedPointCloudFixtures.swift:75 */
#define SCR_MAX_DEPTH 0xffffffffu
#define SCR_FILES_RING_SLOTS 6             /* ComputeRasteriserPointCloud.swift:153-157 */

/* ---- status codes */
#define SCR_OK 0
#define SCR_ERR_INVALID_ARGUMENT (-1)
#define SCR_ERR_NO_FREE_SLOTS (-2)      /* addBatches precondition, ComputeRasteriserPointCloud.swift:419 */
#define SCR_ERR_BATCH_TOO_LARGE (-3)    /* numPoints > slot capacity, :432 */
#define SCR_ERR_POINTS_PER_BATCH (-4)   /* ppb outside 1..65535 */
#define SCR_ERR_HIP (-5)
#define SCR_ERR_OUT_OF_MEMORY (-6)
#define SCR_ERR_SOURCE_TOO_SMALL (-7)   /* copySlice precondition, :657 */
#define SCR_ERR_UNSUPPORTED (-8)
#define SCR_ERR_SLOT_RANGE (-9)         /* removeBatches precondition, :462 */
#define SCR_ERR_NO_DEVICE (-10)
/* PLYPointCloudLoaderError (PLYPointCloudLoader.swift:4-11) */
#define SCR_ERR_PLY_INVALID_HEADER (-11)
#define SCR_ERR_PLY_UNSUPPORTED_FORMAT (-12)
#define SCR_ERR_PLY_MISSING_VERTEX_ELEMENT (-13)
#define SCR_ERR_PLY_MISSING_PROPERTY (-14)
#define SCR_ERR_PLY_UNSUPPORTED_PROPERTY (-15)
#define SCR_ERR_PLY_TRUNCATED_BODY (-16)
#define SCR_ERR_PLY_INVALID_ASCII_VALUE (-17)

/* ---- GPU data layout: Pipelines/ComputeRasteriserTypes.h:11-63 (byte-identical) */
typedef struct scr_raster_batch {
    int32_t state;               /* 0 = slot not resident */
    float minX, minY, minZ;
    float maxX, maxY, maxZ;
    uint32_t numPoints;
    uint32_t firstPoint;         /* absolute pool point index = slot * pointsPerBatch */
    uint32_t fileIndex;
    uint32_t lodCum01, lodCum23, lodCum45, lodCum67; /* cum[L] = points with level <= L, 2 x u16 per word */
    uint32_t padding7, padding8;
} scr_raster_batch; /* 64 B */

typedef struct scr_raster_file {
    float transform[16];         /* viewProjection * world */
    float transformFrustum[16];  /* (frustum ?? viewProjection) * world */
    float world[16];             /* model * file.world */
    float prevTransform[16];     /* previous viewProjection * world */
} scr_raster_file; /* 256 B */

typedef struct scr_visible_batch {
    uint32_t batchIndex;
    int32_t level;               /* position precision: 0 = 30-bit, 1 = 20-bit, >=2 = 10-bit */
    float lodThreshold;
    uint32_t activePoints;       /* LOD survivor prefix */
} scr_visible_batch; /* 16 B */

typedef struct scr_raster_pixel {
    uint32_t depth, red, green, blue, count, weight;
    uint32_t padding[2];
} scr_raster_pixel; /* 32 B, readback form of the HQS accumulators */

/* ---- configuration: ComputeRasteriserTypes.swift:8-142 */
#define SCR_MODE_HIGH_QUALITY_AVERAGE 0
#define SCR_MODE_NEAREST_POINT 1
#define SCR_POINT_SIZE_SCREEN_SPACE 0
#define SCR_POINT_SIZE_WORLD_SPACE 1

typedef struct scr_config {
    int32_t mode;
    float depthTolerance;
    float backgroundColor[4];
    int32_t enableFrustumCulling;
    int32_t lodBias;
    int32_t enableCLOD;
    int32_t enableLODDither;
    int32_t colorizeChunks;
    int32_t colorizeOverdraw;
    int32_t pointSizeMode;
    float minimumPointSize;
    float maximumPointSize;
    float pointSizeScale;
    int32_t applyTint;
    int32_t tintAlphaIsCoverage; /* translucent defocus (with applyTint): weighted-blend passes */
    float motionBlur;            /* > 0: splats swept along their screen velocity, weighted blend */
    int32_t holeFillIterations;  /* hole-fill passes after the resolve (<= 0: none) */
    int32_t motionBlurSamples;
    float motionBlurMaxSpread;
    int32_t applyDisplacement;   /* add each point's displacement before projection (depth, colour, blend passes) */
    int32_t reserved[1];
} scr_config;

/* Reference defaults, ComputeRasteriserTypes.swift:97-119. */
void scr_config_default(scr_config* cfg);

typedef struct scr_camera {
    float viewMatrix[16];
    float projectionMatrix[16];
} scr_camera;

/* ---- opaque handles */
typedef struct scr_ctx scr_ctx;
typedef struct scr_cloud scr_cloud;
typedef struct scr_frame scr_frame;

/* ---- context */
int scr_ctx_create(int device, scr_ctx** out);
/* Releases the caller's reference: the context (its stream and scratch) is
 * freed once every cloud, frame and multi-view renderer made on it has been
 * destroyed as well, so objects may be destroyed in any order. */
int scr_ctx_destroy(scr_ctx* ctx);
const char* scr_last_error(const scr_ctx* ctx);
int scr_sync(scr_ctx* ctx);
/* Version string of the library build (never NULL): version, target and the
 * first 12 hex digits of the SHA-256 of the sources it was built from. */
const char* scr_version(void);

/* ---- per-pass timing (HIP events on the context stream; the reference's
 * equivalent is its labelled encoders under Xcode GPU capture,
 * ComputeRasteriser.swift:389-442).  While enabled, every kernel the render
 * and pack paths launch is bracketed by a pair of events. */
typedef struct scr_pass_time {
    char name[32];
    double totalMs;
    int32_t launches;
} scr_pass_time;
int scr_profile_enable(scr_ctx* ctx, int enable);
/* Synchronise, sum the recorded intervals per pass name, clear the record.
 * Up to `capacity` entries are written; *count receives the number of names. */
int scr_profile_collect(scr_ctx* ctx, scr_pass_time* out, int capacity, int* count);

/* ---- device memory helpers (for callers that keep inputs resident in HBM) */
int scr_device_alloc(scr_ctx* ctx, size_t bytes, void** out);
int scr_device_free(scr_ctx* ctx, void* ptr);
int scr_memcpy_h2d(scr_ctx* ctx, void* dst, const void* src, size_t bytes);
int scr_memcpy_d2h(scr_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy, enqueued on the context stream. */
int scr_memcpy_d2d(scr_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- device measurement (no counterpart in the reference): the global
 * atomic rate the draw passes' tile flushes can reach — waves updating 64
 * consecutive u32 words at pseudo-random spots of an 8 MB plane, atomicMax
 * (add = 0) or atomicAdd (add = 1) — in atomics per second, timed with HIP
 * events on the context stream.  Synchronous; allocates its own buffer. */
int scr_measure_atomic_rate(scr_ctx* ctx, int add, double* atomicsPerSecond);

/* ---- device self-check (no counterpart in the reference): compares the
 * three-instruction reciprocal the draw passes use for batches whose clip w
 * is proven to lie in [2^-125, 2^125] with the correctly rounded 1.0f / w,
 * bit for bit, on every float in [2^-126, 2^126).  Writes the number of
 * mismatches and the bit pattern of the smallest mismatching float
 * (0xffffffff when none).  Synchronous; allocates its own counters. */
int scr_check_fast_reciprocal(scr_ctx* ctx, uint64_t* mismatches, uint32_t* firstMismatch);

/* ---- point cloud slot pool: ComputeRasteriserPointCloud.swift */

/* init(context:capacity:files:originShift:) — ComputeRasteriserPointCloud.swift:201-217.
 * files == NULL / nFiles == 0 gives one identity RasterFile. */
int scr_cloud_create(scr_ctx* ctx, int maxResidentBatches, int pointsPerBatch,
                     const scr_raster_file* files, int nFiles, scr_cloud** out);
int scr_cloud_destroy(scr_cloud* cloud);

typedef struct scr_cloud_info {
    int32_t maxResidentBatches;
    int32_t pointsPerBatch;
    int32_t residentBatchCount;
    int32_t freeSlotCount;
    int64_t pointCount;
    int32_t nFiles;
    int32_t hasTints;
} scr_cloud_info;
int scr_cloud_get_info(const scr_cloud* cloud, scr_cloud_info* out);

/* addBatches(positionsXYZLow:…:batches:commit:) — ComputeRasteriserPointCloud.swift:410-452.
 * Each batch i copies numPoints points starting at batches[i].firstPoint of the
 * source arrays (nSourcePoints long) into a free slot (LIFO), rebases firstPoint
 * to slot * pointsPerBatch and sets state = 1.  outSlots[i] receives the slot. */
int scr_cloud_add_batches(scr_cloud* cloud, const uint32_t* xyzLow, const uint32_t* xyzMed,
                          const uint32_t* xyzHigh, const uint32_t* colors, const uint8_t* levels,
                          size_t nSourcePoints, const scr_raster_batch* batches, int nBatches,
                          int commit, int32_t* outSlots);
/* removeBatches(slots:commit:) — ComputeRasteriserPointCloud.swift:460-473 */
int scr_cloud_remove_batches(scr_cloud* cloud, const int32_t* slots, int n, int commit);
/* commitBatchUpdates() — ComputeRasteriserPointCloud.swift:479-481 */
int scr_cloud_commit(scr_cloud* cloud);
/* clearAllBatches() — ComputeRasteriserPointCloud.swift:367-377 */
int scr_cloud_clear(scr_cloud* cloud);
/* Object.worldMatrix of the cloud node (composed into every file's world). */
int scr_cloud_set_model(scr_cloud* cloud, const float model[16]);
/* updateFiles(viewProjection:modelMatrix:frustumTransform:prevViewProjection:)
 * — ComputeRasteriserPointCloud.swift:487-515.  frustum / prevVP may be NULL. */
int scr_cloud_update_files(scr_cloud* cloud, const float viewProjection[16], const float model[16],
                           const float* frustum16, const float* prevViewProjection16);
/* TintPass output (TintPass.swift:130-184): one float4 per pool point,
 * nPoints == maxResidentBatches * pointsPerBatch.  onDevice != 0: rgba is a
 * device pointer.  rgba == NULL drops the tint buffer. */
int scr_cloud_set_tints(scr_cloud* cloud, const float* rgba, size_t nPoints, int onDevice);
/* displacementBuffer (ComputeRasteriserPointCloud.swift:91-94): one float3
 * per pool point in the 16-byte stride of a Metal float3 array (x, y, z,
 * unused), nPoints == maxResidentBatches * pointsPerBatch, copied in.  A
 * render with applyDisplacement draws point + displacement in the depth,
 * colour and weighted-blend passes (DepthPass/Shaders.metal:69-74,
 * ColorPass/Shaders.metal:72-78; nearest mode does not displace); a NaN
 * component removes the point from all of them.  Clouds without a buffer read
 * zeros (ComputeRasteriser.swift:333-342).  xyz4 == NULL drops the buffer
 * (and the previous-frame copies kept for motion blur). */
int scr_cloud_set_displacements(scr_cloud* cloud, const float* xyz4, size_t nPoints, int onDevice);

/* Device view of a cloud's pool for user kernels — what the TintPass /
 * DisplacementPass preamble binds for user Metal kernels (TintPass.swift:215-338,
 * buffers SCR_TINT_BUF_*).  Device pointers; tints is one float4 per pool
 * point (NULL unless requested); stream is the context's hipStream_t, on which
 * a user kernel runs in order with renders.  The helpers that read it are in
 * scr_device.h. */
typedef struct scr_cloud_view {
    const scr_raster_batch* batches; /* `slots` entries */
    const uint32_t* xyzLow;
    const uint32_t* xyzMed;
    const uint32_t* xyzHigh;
    const uint32_t* colors;          /* packed RGBA8 */
    const uint8_t* levels;           /* LOD level per point */
    float* tints;                    /* float4 per pool point: TintPass output */
    uint32_t pointsPerBatch;
    uint32_t slots;                  /* maxResidentBatches */
    void* stream;
    float* displacements;            /* float4 per pool point (xyz used): DisplacementPass output */
} scr_cloud_view;
/* Fills *out after committing pending batch updates (commitBatchUpdates).
 * flags: SCR_VIEW_TINTS makes sure the cloud has a tint buffer (a new one
 * holds zeros; one installed with scr_cloud_set_tints is returned as is) for
 * a user kernel to write and a render with applyTint to read (TintPass.swift:
 * 130-184); SCR_VIEW_DISPLACEMENTS does the same for the displacement buffer,
 * which a render with applyDisplacement reads (DisplacementPass.swift:120-172
 * allocates it and turns applyDisplacement on).  The pointers stay valid until
 * the pool is resized or destroyed. */
#define SCR_VIEW_TINTS 1
#define SCR_VIEW_DISPLACEMENTS 2
int scr_cloud_device_view(scr_cloud* cloud, int flags, scr_cloud_view* out);

#define SCR_POOL_XYZ_LOW 0
#define SCR_POOL_XYZ_MED 1
#define SCR_POOL_XYZ_HIGH 2
#define SCR_POOL_COLORS 3
#define SCR_POOL_LEVELS 4
#define SCR_POOL_BATCHES 5
/* Read back a pool buffer (whole capacity) into host memory. */
int scr_cloud_read_pool(scr_cloud* cloud, int which, void* host, size_t bytes);
/* Read back the visible-batch list of the last render of this cloud
 * (VisibleBatch records, cull order).  *count receives the visible count. */
int scr_cloud_read_visible(scr_cloud* cloud, scr_visible_batch* host, int capacity, int* count);

/* ---- GPU packer: GPUPacker.swift (pack :140-160, packChunk :179-197, addRawChunksGPU :545-577)
 * and PackedPointCloudFixtures.pack (PackedPointCloudFixtures.swift:30-167).
 * Inputs are float4 positions (xyz + pad, the 16-byte stride of an MTLBuffer of
 * float3) and float4 RGBA colours in [0, 1]. */
typedef struct scr_pack_options {
    int32_t keyBits;              /* 30: 10 bits/axis (the reference); 63: 21 bits/axis */
    int32_t lodLevels;            /* clamped to 1..8 (default 4) */
    int32_t coarseVoxelDivisions; /* >= 1 (default 64) */
    int32_t inputOnDevice;        /* positions/colors are device pointers */
} scr_pack_options;
void scr_pack_options_default(scr_pack_options* opt);

/* Wholesale pack: GPUPacker.pack + prepareForGPUPack (ComputeRasteriserPointCloud.swift:278-307).
 * Grows the pool when needed (keeping pointsPerBatch; an installed tint buffer
 * is dropped then), makes slots [0, nb) resident, the rest empty.  count == 0
 * clears the cloud.  sourceIndicesOut (host, count entries, may be NULL)
 * receives the pack permutation (PackedPointCloud.sourceIndices). */
int scr_pack(scr_cloud* cloud, const float* positions4, const float* colors4, size_t count,
             const scr_pack_options* opt, uint32_t* sourceIndicesOut);
/* Slot-targeted chunk pack (addRawChunksGPU for one chunk): reserves the first
 * run of ceil(count / pointsPerBatch) free contiguous slots, packs into it and
 * adds the batches to the resident set.  *outFirstSlot receives the run start. */
int scr_pack_chunk(scr_cloud* cloud, const float* positions4, const float* colors4, size_t count,
                   const scr_pack_options* opt, int32_t* outFirstSlot);

/* ---- PLY input: PLYPointCloudLoader.parse (PLYPointCloudLoader.swift:38-333)
 * up to the point where the reference hands its arrays to the CPU pack.
 * Parses an ASCII or binary little-endian PLY image in host memory into
 * float4 positions (x, y, z, 0) and float4 colours (r, g, b, 1 — channels the
 * file lacks stay 1; integer channels divided by their type's maximum, all
 * clamped to [0, 1]), the arrays scr_pack takes.  With positions4 and colors4
 * NULL it validates the whole image and returns the vertex count only; with
 * both set (capacity >= the count, 4 floats per vertex each) it also writes
 * the vertices.  Errors are the SCR_ERR_PLY_* codes above, with the
 * reference's message (property name, vertex line) in msg when given.  Host
 * only: needs no context or device. */
int scr_ply_parse(const void* data, size_t bytes, size_t capacity, float* positions4, float* colors4,
                  size_t* vertexCount, char* msg, size_t msgBytes);

/* ---- synthetic scene generator (benchmark and test input; no reference
 * counterpart).  A LiDAR-like scene: terrainFraction of the points on a
 * value-noise heightfield over [-halfExtent, halfExtent]^2, the rest on the
 * side and top faces of numBoxes axis-aligned boxes standing on the terrain;
 * colours from a height colour map plus uniform noise.  Deterministic in
 * (seed, point index).  Outputs float4 positions and float4 colours. */
typedef struct scr_scene_params {
    uint64_t seed;
    float halfExtent;       /* metres: 1000 for a 2 km square */
    int32_t numBoxes;
    float terrainFraction;  /* 0.7 */
    float amplitude;        /* heightfield amplitude, metres (120) */
    float baseWavelength;   /* first octave wavelength, metres (500) */
    int32_t octaves;        /* 6 */
    float boxMinEdge;       /* 10 */
    float boxMaxEdge;       /* 60 */
    float colorNoise;       /* +/- noise per channel in 0..255 units (10) */
    int32_t quantizeColors; /* 1: colours are RGB8 values / 255; 0: continuous colour map */
} scr_scene_params;
void scr_scene_params_default(scr_scene_params* p);
int scr_gen_scene(scr_ctx* ctx, const scr_scene_params* params, size_t count, float* positions4,
                  float* colors4, int outputOnDevice);
/* Terrain height of the scene's heightfield at (x, y) (host computation,
 * the same function the generator evaluates), for placing cameras. */
float scr_scene_height(const scr_scene_params* params, float x, float y);

/* ---- frames (per-viewport resources: ComputeRasteriser.swift CachedResources) */
int scr_frame_create(scr_ctx* ctx, int width, int height, scr_frame** out);
int scr_frame_destroy(scr_frame* frame);

#define SCR_PLANE_DEPTH 0          /* u32 per pixel, HQS depth */
#define SCR_PLANE_RED 1            /* u32 sums */
#define SCR_PLANE_GREEN 2
#define SCR_PLANE_BLUE 3
#define SCR_PLANE_COUNT 4
#define SCR_PLANE_NEAREST_DEPTH 5  /* u32 */
#define SCR_PLANE_NEAREST_INDEX 6  /* u32, 0xffffffff = none */
#define SCR_PLANE_RGBA8 7          /* u32 = r | g<<8 | b<<16 | a<<24 (rgba8Unorm texel) */
#define SCR_PLANE_DEPTH_F32 8      /* float reversed-Z depth texture */
#define SCR_PLANE_DIFF_SUM_R 9     /* int64 per pixel: differentiable-mode fixed-point sums (x 2^16) */
#define SCR_PLANE_DIFF_SUM_G 10
#define SCR_PLANE_DIFF_SUM_B 11
/* Copy one plane (width*height elements of 4 B; 8 B for the DIFF_SUM planes)
 * to host memory. */
int scr_frame_read(scr_frame* frame, int which, void* host, size_t bytes);
/* HoleFillProcessor.update(_:iterations:) (HoleFillProcessor.swift:22-45,
 * HoleFill/Shaders.metal:10-48): `iterations` passes of the 3 x 3
 * neighbour-average expansion over the frame's RGBA8 plane, in place (through
 * a scratch plane the frame allocates on first use), enqueued on the context
 * stream after the frame's last render.  A texel with alpha >= 0.5 is kept;
 * an empty one takes the mean colour of its valid neighbours with alpha 1. */
int scr_frame_hole_fill(scr_frame* frame, int iterations);

/* ---- composite: ComputeRasteriser.draw (ComputeRasteriser.swift:672-683)
 * and its fragment shaders (Pipelines/Post.metal:23-50).  Composites the
 * frame's final RGBA8 plane (after any hole fill) onto a caller's colour
 * target of the same size, enqueued on the context stream after the frame's
 * last render:
 *   SCR_COMPOSITE_OVER  (computeRasteriserPostFragment): every pixel, the cloud
 *                       colour blended over the target, no depth;
 *   SCR_COMPOSITE_DEPTH (computeRasteriserPostDepthFragment, writesSceneDepth):
 *                       pixels where the cloud's reverse-Z depth d is <= 0 are
 *                       left alone; elsewhere the fragment carries depth d, is
 *                       kept where depthCompare(d, target depth) holds, writes
 *                       d to the depth target when writeDepth, and blends.
 * Blending and the depth compare are render-pipeline state the reference takes
 * from its post material; here they are options: SCR_BLEND_REPLACE, or
 * SCR_BLEND_OVER (non-premultiplied source-over: c = s.rgb * s.a + d.rgb * (1 -
 * s.a), a = s.a + d.a * (1 - s.a), in binary32, bytes rounded to nearest-even).
 * targetRGBA8: device, r | g<<8 | b<<16 | a<<24 per pixel; targetDepth: device
 * floats (SCR_COMPOSITE_DEPTH only); pitch in pixels (>= width), row 0 = top. */
#define SCR_COMPOSITE_OVER 0
#define SCR_COMPOSITE_DEPTH 1
#define SCR_BLEND_REPLACE 0
#define SCR_BLEND_OVER 1
#define SCR_COMPARE_ALWAYS 0
#define SCR_COMPARE_GREATER 1
#define SCR_COMPARE_GREATER_EQUAL 2
#define SCR_COMPARE_LESS 3
#define SCR_COMPARE_LESS_EQUAL 4
typedef struct scr_composite_options {
    int32_t mode;          /* SCR_COMPOSITE_* */
    int32_t blend;         /* SCR_BLEND_* */
    int32_t depthCompare;  /* SCR_COMPARE_*: fragment depth against target depth (reverse-Z: GREATER is nearer) */
    int32_t writeDepth;    /* SCR_COMPOSITE_DEPTH: store the cloud depth where the fragment is kept */
} scr_composite_options;
void scr_composite_options_default(scr_composite_options* opt); /* DEPTH, OVER, GREATER, no depth write */
int scr_frame_composite(scr_frame* frame, const scr_composite_options* opt, uint32_t* targetRGBA8, float* targetDepth,
                        int targetPitch);

/* ---- rendering: ComputeRasteriser.update (:201-268) + encode (:277-296).
 * HQS: clear → cull(all clouds) → depth(all) → colour(all) → resolve
 * (encodeHighQualityAverage :322-456).  Nearest: first cloud with resident
 * batches only (encodeNearestPoint :483-497).  Computes VP = P·V on the host
 * and calls scr_cloud_update_files for every cloud with its model matrix. */
int scr_render(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam,
               scr_cloud* const* clouds, int nClouds, scr_frame* frame);
/* The same render for the camera `cameraKey` names.  Motion blur sweeps from
 * the previous frame's pose of that camera: the frame keeps the last
 * view-projection per key (ComputeRasteriser.swift:255-266 keys it by the
 * camera object so stereo eyes do not collide; the current one is used the
 * first time a key is seen), and each displaced cloud keeps per key a copy of
 * the displacements its last motion-blurred render used, the previous
 * displacements of the next one (ComputeRasteriser.swift:434-480).
 * scr_render is scr_render_camera with cameraKey 0. */
int scr_render_camera(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, uint64_t cameraKey,
                      scr_cloud* const* clouds, int nClouds, scr_frame* frame);

/* pickPointIndex(atNDC:in:camera:searchRadius:) — ComputeRasteriser.swift:609-670.
 * Runs the nearest-index pass for one cloud into `frame`, reads back a band of
 * rows and returns the nearest non-sentinel pixel's point index within
 * searchRadius.  *found = 0 when there is none. */
int scr_pick(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, scr_cloud* cloud,
             scr_frame* frame, float ndcX, float ndcY, int searchRadius,
             uint32_t* outIndex, int* found);

/* ---- differentiable HQS.
 * No reference counterpart: it extends the HQS passes (encodeHighQualityAverage,
 * ComputeRasteriser.swift:322-456) for fitting per-point colours.  Visibility
 * is exactly HQS's (cull, depth pass, the colour pass's depth-tolerance test,
 * ColorPass/Shaders.metal:201-209), so the frame's DEPTH and COUNT planes equal
 * scr_render's.  The colour of pool point p is params[4p .. 4p+2] (device
 * float4 per pool point, .w unused), accumulated per pixel as int64
 * rint(param * 2^16) (SCR_PLANE_DIFF_SUM_*), and
 *     image[3 * pixel + k] = (float)(sum_k / (2^16 * count))   (FP64 divide)
 * or backgroundColor[k] where count == 0; image is a device pointer to
 * 3 * width * height floats, row 0 = top of the viewport.  Integer sums make
 * the forward deterministic.  mode must be HQS; applyTint / colorize* must be
 * 0 (SCR_ERR_UNSUPPORTED otherwise). */
int scr_diff_render(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, scr_cloud* cloud,
                    const float* params, scr_frame* frame, float* image);
/* Backward of scr_diff_render for the same cfg / cam / cloud / frame (the
 * frame's depth and count planes from that forward are read, each point's
 * fragments are recomputed; nothing per fragment is stored).  For every
 * point p and channel k:
 *     grad[4p + k] += Σ over p's kept fragments of dLdImage[3 * pixel + k] · (1 / count[pixel])
 * in FP64 (the fixed-point rounding passes the gradient straight through).
 * dLdImage: device, 3 floats per pixel; grad: device, 4 doubles per pool
 * point, accumulated (views add up in call order; no atomics). */
int scr_diff_backward(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, scr_cloud* cloud,
                      scr_frame* frame, const float* dLdImage, double* grad);

/* ---- batched multi-view rendering.
 * The reference renders one camera per encode (ComputeRasteriser.update
 * :201-268 + encode :277-296); a caller with many cameras repeats that call.
 * A multi-view renderer does the same work for nViews cameras at once: view v
 * renders cloud clouds[viewCloud[v]] into frames[v], exactly as
 *     scr_render(ctx, cfg, &cameras[v], &clouds[viewCloud[v]], 1, frames[v])
 * would (bit-identical planes), but every phase is one launch for all views:
 * clear, cull over every (view, slot) pair, then persistent draw kernels that
 * take (visible batch, view) work items from a device queue, then resolve.
 * The launch sequence can be replayed from a captured HIP graph.  Frames must
 * be distinct; clouds may be shared by several views. */
typedef struct scr_multiview scr_multiview;

/* Draw-kernel shape for one phase: threads per workgroup (128, 256 or 512),
 * points per thread (4, 8 or 16; threads * points <= 4096), the number of
 * work items each visible batch is split into (1, 2 or 4), and whether each
 * wave draws its own chunk with a private LDS tile (1) or the workgroup
 * shares one chunk and tile (0). */
typedef struct scr_mv_shape {
    int32_t threads;
    int32_t pointsPerThread;
    int32_t splits;
    int32_t waveTiles;
} scr_mv_shape;

typedef struct scr_mv_tuning {
    scr_mv_shape depth;
    scr_mv_shape color;
    scr_mv_shape nearest;
} scr_mv_tuning;

int scr_multiview_create(scr_ctx* ctx, scr_cloud* const* clouds, int nClouds, scr_frame* const* frames,
                         const int32_t* viewCloud, int nViews, scr_multiview** out);
int scr_multiview_destroy(scr_multiview* mv);
/* Renders all views.  useGraph != 0 replays a HIP graph of the launch
 * sequence (captured on first use and again whenever the config's mode or
 * footprint class, the tuning or the clouds' capacities change). */
int scr_multiview_render(scr_multiview* mv, const scr_config* cfg, const scr_camera* cameras, int useGraph);
int scr_multiview_get_tuning(const scr_multiview* mv, scr_mv_tuning* out);
int scr_multiview_set_tuning(scr_multiview* mv, const scr_mv_tuning* tuning);
/* Times every supported shape of each draw phase on these cameras (reps
 * renders each, best time kept), installs the fastest per phase and returns
 * it.  Results do not depend on the shape, only the time does. */
int scr_multiview_autotune(scr_multiview* mv, const scr_config* cfg, const scr_camera* cameras, int reps,
                           scr_mv_tuning* out);
/* Differentiable HQS (see scr_diff_render) for every view at once; the
 * multi-view renderer must have exactly one cloud.  images[v]: device, 3
 * floats per pixel of frame v.  Results equal per-view scr_diff_render calls
 * bit for bit. */
int scr_multiview_diff_render(scr_multiview* mv, const scr_config* cfg, const scr_camera* cameras,
                              const float* params, float* const* images);
/* Backward of scr_multiview_diff_render for the same cameras: grad (device,
 * 4 doubles per pool point) += the gradients of all views.  Each point's
 * views are added in view order by one thread, so the result equals per-view
 * scr_diff_backward calls made in view order, bit for bit. */
int scr_multiview_diff_backward(scr_multiview* mv, const scr_config* cfg, const scr_camera* cameras,
                                const float* const* dLdImages, double* grad);

#ifdef __cplusplus
}
#endif
#endif /* SCR_H */
