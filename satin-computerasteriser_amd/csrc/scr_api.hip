// scr_api.hip — C ABI host runtime: context, point-cloud slot pool, frames,
// render orchestration and picking.  Reference behaviour is cited per entry
// point in include/scr.h; the orchestration follows ComputeRasteriser.swift.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "scr_internal.hpp"

namespace scr {

int set_error(scr_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->lastError = msg;
    return code;
}

int hip_error(scr_ctx* ctx, hipError_t e, const char* what) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return set_error(ctx, e == hipErrorOutOfMemory ? SCR_ERR_OUT_OF_MEMORY : SCR_ERR_HIP, buf);
}

void ctx_retain(scr_ctx* ctx) { ctx->refs++; }

void ctx_release(scr_ctx* ctx) {
    if (--ctx->refs > 0) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& iv : ctx->intervals) {
        (void)hipEventDestroy(iv.begin);
        (void)hipEventDestroy(iv.end);
    }
    for (auto e : ctx->eventPool) (void)hipEventDestroy(e);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->packBounds) (void)hipFree(ctx->packBounds);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

hipEvent_t take_event(scr_ctx* ctx) {
    if (!ctx->eventPool.empty()) {
        hipEvent_t e = ctx->eventPool.back();
        ctx->eventPool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

PassScope::PassScope(scr_ctx* ctx, const char* name) : ctx_(ctx) {
    if (!ctx->profiling) return;
    scr_ctx::Interval iv{name, take_event(ctx), take_event(ctx)};
    if (!iv.begin || !iv.end) return;
    (void)hipEventRecord(iv.begin, ctx->stream);
    ctx->intervals.push_back(iv);
    index_ = (int)ctx->intervals.size() - 1;
}

PassScope::~PassScope() {
    if (index_ >= 0) (void)hipEventRecord(ctx_->intervals[(size_t)index_].end, ctx_->stream);
}

}  // namespace scr

using scr::hip_error;
using scr::set_error;

#define SCR_HIP(ctx, call)                                   \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_error(ctx, e_, #call); \
    } while (0)

namespace scr {

// flushBatchMirror — ComputeRasteriserPointCloud.swift:629-644 (dirty range only)
int flush_mirror(scr_cloud* c) {
    if (c->dirtyLo > c->dirtyHi) return SCR_OK;
    const int lo = std::max(0, c->dirtyLo);
    const int hi = std::min(c->dirtyHi, c->maxResidentBatches - 1);
    c->dirtyLo = 0x7fffffff;
    c->dirtyHi = -1;
    if (lo > hi) return SCR_OK;
    SCR_HIP(c->ctx, hipStreamSynchronize(c->ctx->stream));
    SCR_HIP(c->ctx, hipMemcpy(c->dBatches + lo, c->mirror.data() + lo, (size_t)(hi - lo + 1) * sizeof(scr_raster_batch),
                              hipMemcpyHostToDevice));
    return SCR_OK;
}

// updateFiles composition (ComputeRasteriserPointCloud.swift:487-515):
// world = model·file.world, transform = VP·world, transformFrustum =
// (frustum ?? VP)·world, prevTransform = (prevVP ?? VP)·world.
void compose_files(const scr_cloud* c, const float* model, const float* vp, const float* frustum,
                   const float* prevVP, scr_raster_file* out) {
    for (size_t i = 0; i < c->files.size(); i++) {
        scr_raster_file f;
        float world[16];
        mat_mul(model, c->files[i].world, world);
        mat_mul(vp, world, f.transform);
        mat_mul(frustum ? frustum : vp, world, f.transformFrustum);
        std::memcpy(f.world, world, sizeof world);
        mat_mul(prevVP ? prevVP : vp, world, f.prevTransform);
        out[i] = f;
    }
}

int check_config(scr_ctx* ctx, const scr_config* cfg, bool allowCoverage) {
    if (!cfg) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "config is NULL");
    if (cfg->mode != SCR_MODE_HIGH_QUALITY_AVERAGE && cfg->mode != SCR_MODE_NEAREST_POINT)
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "unknown render mode");
    if (!allowCoverage && cfg->tintAlphaIsCoverage != 0 && cfg->applyTint != 0)
        return set_error(ctx, SCR_ERR_UNSUPPORTED, "translucent defocus (tintAlphaIsCoverage) needs scr_render");
    if (!allowCoverage && cfg->motionBlur > 0.0f)
        return set_error(ctx, SCR_ERR_UNSUPPORTED, "motion blur needs scr_render");
    if (!allowCoverage && cfg->holeFillIterations > 0)
        return set_error(ctx, SCR_ERR_UNSUPPORTED, "hole fill runs in scr_render or scr_frame_hole_fill");
    if (!allowCoverage && cfg->applyDisplacement != 0)
        return set_error(ctx, SCR_ERR_UNSUPPORTED, "displacement needs scr_render");
    if (cfg->pointSizeMode != 0 && cfg->pointSizeMode != 1)
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "unknown point size mode");
    return SCR_OK;
}

PassParams make_params(const scr_config* cfg, const scr_camera* cam, int width, int height) {
    PassParams p;
    std::memset(&p, 0, sizeof p);
    std::memcpy(p.view, cam->viewMatrix, sizeof p.view);
    std::memcpy(p.proj, cam->projectionMatrix, sizeof p.proj);
    p.width = width;
    p.height = height;
    p.pointSizeMode = cfg->pointSizeMode;
    p.minPointSize = cfg->minimumPointSize;
    p.maxPointSize = cfg->maximumPointSize;
    p.pointSizeScale = cfg->pointSizeScale;
    p.lodDither = cfg->enableLODDither;
    p.enableFrustumCulling = cfg->enableFrustumCulling;
    p.lodBias = cfg->lodBias;
    p.enableCLOD = cfg->enableCLOD;
    p.depthKeep = 1.0f - cfg->depthTolerance;
    p.colorizeChunks = cfg->colorizeChunks;
    p.colorizeOverdraw = cfg->colorizeOverdraw;
    p.applyTint = cfg->applyTint;
    // the nearest passes do not displace (ComputeRasteriser.swift:727-731)
    p.applyDisplacement = (cfg->applyDisplacement != 0 && cfg->mode == SCR_MODE_HIGH_QUALITY_AVERAGE) ? 1 : 0;
    const float mn = cfg->minimumPointSize, mx = cfg->maximumPointSize;
    p.lo = std::fmax(std::fmin(mn, mx), 1.0f);
    p.hi = std::fmax(std::fmax(mn, mx), p.lo);
    p.constRadius = (p.lo == p.hi) ? const_radius_for(p.lo) : -1;
    p.ortho = cam->projectionMatrix[15] == 1.0f;
    p.focal = ((float)height * 0.5f) * cam->projectionMatrix[5];
    if (p.constRadius < 0 && p.ortho) {
        // orthographic sizes do not depend on the point (Common.metal:42-43, 49-50)
        const float ps = p.pointSizeMode == 1 ? (2.0f * p.pointSizeScale) * p.focal : p.pointSizeScale;
        p.constRadius = radius_of_size(ps, p.lo, p.hi);
    }
    p.maxRadius = const_radius_for(p.hi);
    for (int i = 0; i < 4; i++) p.bg[i] = cfg->backgroundColor[i];
    return p;
}

}  // namespace scr

using scr::check_config;
using scr::flush_mirror;
using scr::make_params;

namespace {

const float kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

template <typename T>
hipError_t dalloc(T*& p, size_t count) {
    p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)&p, count * sizeof(T));
}

void free_pool(scr_cloud* c);

// The displacement buffer and its per-camera previous-frame copies.
void drop_displacements(scr_cloud* c) {
    dfree(c->dDisp);
    for (auto& kv : c->prevDisp) dfree(kv.second);
    c->prevDisp.clear();
}

void release_cloud(scr_cloud* c) {
    if (!c) return;
    free_pool(c);
    dfree(c->dFiles);
    dfree(c->dCounter);
    if (c->hFiles) (void)hipHostFree(c->hFiles);
    c->hFiles = nullptr;
    for (hipEvent_t& e : c->filesCopied) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
}

void mark_dirty(scr_cloud* c, int slot) {
    c->dirtyLo = std::min(c->dirtyLo, slot);
    c->dirtyHi = std::max(c->dirtyHi, slot);
}

// Upload of the composed table through the pinned ring; the slot's event keeps
// its staging untouched until the copy kFilesRing uploads ago has consumed it.
int upload_files(scr_cloud* c, const float* model, const float* vp, const float* frustum, const float* prevVP) {
    const int k = c->filesNext;
    c->filesNext = (k + 1) % scr_cloud::kFilesRing;
    const size_t n = c->files.size();
    scr_raster_file* stage = c->hFiles + (size_t)k * n;
    SCR_HIP(c->ctx, hipEventSynchronize(c->filesCopied[k]));
    scr::compose_files(c, model, vp, frustum, prevVP, stage);
    SCR_HIP(c->ctx, hipMemcpyAsync(c->dFiles, stage, n * sizeof(scr_raster_file), hipMemcpyHostToDevice,
                                   c->ctx->stream));
    SCR_HIP(c->ctx, hipEventRecord(c->filesCopied[k], c->ctx->stream));
    return SCR_OK;
}

// (Re)allocate every per-slot / per-point device buffer for the current
// capacity and reset the pool to empty (reallocate + allocateBuffers,
// ComputeRasteriserPointCloud.swift:519-547).
hipError_t alloc_pool(scr_cloud* c) {
    scr_ctx* ctx = c->ctx;
    const size_t pts = c->poolPoints();
    const int n = c->maxResidentBatches;
    scr_raster_batch empty;
    std::memset(&empty, 0, sizeof empty);
    c->mirror.assign((size_t)n, empty);
    c->freeSlots.resize((size_t)n);
    for (int i = 0; i < n; i++) c->freeSlots[(size_t)i] = n - 1 - i;
    c->residentBatchCount = 0;
    c->pointCount = 0;
    c->dirtyLo = 0x7fffffff;
    c->dirtyHi = -1;
    hipError_t e = dalloc(c->dLow, pts);
    if (e == hipSuccess) e = dalloc(c->dMed, pts);
    if (e == hipSuccess) e = dalloc(c->dHigh, pts);
    if (e == hipSuccess) e = dalloc(c->dColors, pts);
    if (e == hipSuccess) e = dalloc(c->dLevels, pts);
    if (e == hipSuccess) e = dalloc(c->dBatches, (size_t)n);
    if (e == hipSuccess) e = dalloc(c->dVisible, (size_t)n);
    if (e == hipSuccess) e = dalloc(c->dBounds, (size_t)n);
    if (e == hipSuccess) e = hipMemsetAsync(c->dLevels, 0, pts, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpy(c->dBatches, c->mirror.data(), c->mirror.size() * sizeof(scr_raster_batch), hipMemcpyHostToDevice);
    return e;
}

void free_pool(scr_cloud* c) {
    dfree(c->dLow);
    dfree(c->dMed);
    dfree(c->dHigh);
    dfree(c->dColors);
    dfree(c->dLevels);
    dfree(c->dBatches);
    dfree(c->dVisible);
    dfree(c->dBounds);
    drop_displacements(c);
    if (c->tintsOwned) dfree(c->dTints);
    c->dTints = nullptr;
    c->tintsOwned = false;
}

}  // namespace

extern "C" {

#ifndef SCR_SOURCE_HASH
#define SCR_SOURCE_HASH "unknown"
#endif
const char* scr_version(void) { return "scr 0.1 (gfx950, source " SCR_SOURCE_HASH ")"; }

void scr_config_default(scr_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->mode = SCR_MODE_HIGH_QUALITY_AVERAGE;
    c->depthTolerance = 0.01f;
    c->enableFrustumCulling = 1;
    c->lodBias = 0;
    c->enableCLOD = 1;
    c->enableLODDither = 1;
    c->pointSizeMode = SCR_POINT_SIZE_SCREEN_SPACE;
    c->minimumPointSize = 1.0f;
    c->maximumPointSize = 1.0f;
    c->pointSizeScale = 1.0f;
    c->holeFillIterations = 0;
    c->motionBlurSamples = 8;
    c->motionBlurMaxSpread = 64.0f;
}

// ------------------------------------------------------------------ context

int scr_ctx_create(int device, scr_ctx** out) {
    if (!out) return SCR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SCR_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = new (std::nothrow) scr_ctx();
    if (!ctx) return SCR_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return SCR_ERR_HIP;
    }
    ctx->numCUs = prop.multiProcessorCount;
    // the pack's bounds words
    if (hipMalloc((void**)&ctx->packBounds, 256) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return SCR_ERR_HIP;
    }
    *out = ctx;
    return SCR_OK;
}

int scr_profile_enable(scr_ctx* ctx, int enable) {
    if (!ctx) return SCR_ERR_INVALID_ARGUMENT;
    ctx->profiling = enable != 0;
    return SCR_OK;
}

int scr_profile_collect(scr_ctx* ctx, scr_pass_time* out, int capacity, int* count) {
    if (!ctx || !count) return SCR_ERR_INVALID_ARGUMENT;
    SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<scr_pass_time> acc;
    for (auto& iv : ctx->intervals) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, iv.begin, iv.end) != hipSuccess) ms = 0.0f;
        size_t k = 0;
        while (k < acc.size() && std::strncmp(acc[k].name, iv.name, sizeof acc[k].name) != 0) k++;
        if (k == acc.size()) {
            scr_pass_time t;
            std::memset(&t, 0, sizeof t);
            std::snprintf(t.name, sizeof t.name, "%s", iv.name);
            acc.push_back(t);
        }
        acc[k].totalMs += ms;
        acc[k].launches += 1;
        ctx->eventPool.push_back(iv.begin);
        ctx->eventPool.push_back(iv.end);
    }
    ctx->intervals.clear();
    *count = (int)acc.size();
    for (int i = 0; i < capacity && i < (int)acc.size(); i++) out[i] = acc[(size_t)i];
    return SCR_OK;
}

int scr_ctx_destroy(scr_ctx* ctx) {
    if (!ctx) return SCR_OK;
    scr::ctx_release(ctx);
    return SCR_OK;
}

const char* scr_last_error(const scr_ctx* ctx) { return ctx ? ctx->lastError.c_str() : "no context"; }

int scr_sync(scr_ctx* ctx) {
    if (!ctx) return SCR_ERR_INVALID_ARGUMENT;
    SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SCR_OK;
}

int scr_device_alloc(scr_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return SCR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (bytes == 0) return SCR_OK;
    SCR_HIP(ctx, hipMalloc(out, bytes));
    return SCR_OK;
}

int scr_device_free(scr_ctx* ctx, void* ptr) {
    if (!ctx) return SCR_ERR_INVALID_ARGUMENT;
    if (ptr) {
        SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        SCR_HIP(ctx, hipFree(ptr));
    }
    return SCR_OK;
}

int scr_memcpy_h2d(scr_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return SCR_ERR_INVALID_ARGUMENT;
    if (!bytes) return SCR_OK;
    SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SCR_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SCR_OK;
}

int scr_memcpy_d2h(scr_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return SCR_ERR_INVALID_ARGUMENT;
    if (!bytes) return SCR_OK;
    SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SCR_HIP(ctx, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SCR_OK;
}

int scr_memcpy_d2d(scr_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return SCR_ERR_INVALID_ARGUMENT;
    if (!bytes) return SCR_OK;
    SCR_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return SCR_OK;
}

int scr_measure_atomic_rate(scr_ctx* ctx, int add, double* atomicsPerSecond) {
    if (!ctx || !atomicsPerSecond) return SCR_ERR_INVALID_ARGUMENT;
    *atomicsPerSecond = 0.0;
    constexpr uint32_t kSpots = 1u << 15;  // 32 K spots x 64 words: an 8 MB plane
    uint32_t* buf = nullptr;
    SCR_HIP(ctx, hipMalloc(&buf, (size_t)kSpots * 64u * sizeof(uint32_t)));
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipError_t e = scr::launch_fill_u32(ctx->stream, buf, (size_t)kSpots * 64u, 0u);
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    uint64_t n = 0;
    if (e == hipSuccess) e = scr::launch_atomic_rate(ctx->stream, ctx->numCUs, buf, kSpots, add, 64, &n);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(t0, ctx->stream);
    if (e == hipSuccess) e = scr::launch_atomic_rate(ctx->stream, ctx->numCUs, buf, kSpots, add, 512, &n);
    if (e == hipSuccess) e = hipEventRecord(t1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_error(ctx, e, "atomic rate");
    *atomicsPerSecond = ms > 0.0f ? (double)n / ((double)ms * 1e-3) : 0.0;
    return SCR_OK;
}

int scr_check_fast_reciprocal(scr_ctx* ctx, uint64_t* mismatches, uint32_t* firstMismatch) {
    if (!ctx || !mismatches || !firstMismatch) return SCR_ERR_INVALID_ARGUMENT;
    *mismatches = 0;
    *firstMismatch = 0xffffffffu;
    unsigned long long* d = nullptr;
    SCR_HIP(ctx, hipMalloc(&d, 2 * sizeof(unsigned long long)));
    uint32_t* first = reinterpret_cast<uint32_t*>(d + 1);
    hipError_t e = scr::launch_fill_u64(ctx->stream, d, 1, 0ull);
    if (e == hipSuccess) e = scr::launch_fill_u32(ctx->stream, first, 1, 0xffffffffu);
    if (e == hipSuccess) e = scr::launch_check_fast_rcp(ctx->stream, ctx->numCUs, d, first);
    unsigned long long h[2] = {0ull, 0ull};
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_error(ctx, e, "fast reciprocal check");
    *mismatches = (uint64_t)h[0];
    uint32_t f;
    memcpy(&f, &h[1], sizeof(f));
    *firstMismatch = f;
    return SCR_OK;
}

// ------------------------------------------------------------------ clouds

int scr_cloud_create(scr_ctx* ctx, int maxResidentBatches, int pointsPerBatch, const scr_raster_file* files,
                     int nFiles, scr_cloud** out) {
    if (!ctx || !out) return SCR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (maxResidentBatches <= 0) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "maxResidentBatches must be positive");
    if (pointsPerBatch <= 0 || pointsPerBatch > SCR_MAX_POINTS_PER_BATCH)
        return set_error(ctx, SCR_ERR_POINTS_PER_BATCH, "pointsPerBatch must be in 1..65535");
    if (nFiles < 0 || (nFiles > 0 && !files)) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "bad files table");
    scr_cloud* c = new (std::nothrow) scr_cloud();
    if (!c) return SCR_ERR_OUT_OF_MEMORY;
    c->ctx = ctx;
    c->maxResidentBatches = maxResidentBatches;
    c->pointsPerBatch = pointsPerBatch;
    std::memcpy(c->model, kIdentity, sizeof c->model);
    if (nFiles == 0) {
        scr_raster_file f;
        for (int i = 0; i < 16; i++) f.transform[i] = f.transformFrustum[i] = f.world[i] = f.prevTransform[i] = kIdentity[i];
        c->files.push_back(f);
    } else {
        c->files.assign(files, files + nFiles);
    }
    (void)hipSetDevice(ctx->device);
    hipError_t e = alloc_pool(c);
    if (e == hipSuccess) e = dalloc(c->dFiles, c->files.size());
    if (e == hipSuccess) e = dalloc(c->dCounter, 2);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&c->hFiles, scr_cloud::kFilesRing * c->files.size() * sizeof(scr_raster_file));
    for (int k = 0; k < scr_cloud::kFilesRing && e == hipSuccess; k++)
        e = hipEventCreateWithFlags(&c->filesCopied[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemsetAsync(c->dCounter, 0, 2 * sizeof(uint32_t), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpy(c->dFiles, c->files.data(), c->files.size() * sizeof(scr_raster_file), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        release_cloud(c);
        delete c;
        return hip_error(ctx, e, "scr_cloud_create");
    }
    scr::ctx_retain(ctx);
    *out = c;
    return SCR_OK;
}

int scr_cloud_destroy(scr_cloud* c) {
    if (!c) return SCR_OK;
    scr_ctx* ctx = c->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    release_cloud(c);
    delete c;
    scr::ctx_release(ctx);
    return SCR_OK;
}

int scr_cloud_get_info(const scr_cloud* c, scr_cloud_info* out) {
    if (!c || !out) return SCR_ERR_INVALID_ARGUMENT;
    out->maxResidentBatches = c->maxResidentBatches;
    out->pointsPerBatch = c->pointsPerBatch;
    out->residentBatchCount = c->residentBatchCount;
    out->freeSlotCount = (int32_t)c->freeSlots.size();
    out->pointCount = c->pointCount;
    out->nFiles = (int32_t)c->files.size();
    out->hasTints = c->dTints != nullptr;
    return SCR_OK;
}

int scr_cloud_add_batches(scr_cloud* c, const uint32_t* low, const uint32_t* med, const uint32_t* high,
                          const uint32_t* colors, const uint8_t* levels, size_t nSourcePoints,
                          const scr_raster_batch* batches, int nBatches, int commit, int32_t* outSlots) {
    if (!c) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = c->ctx;
    if (nBatches < 0 || (nBatches > 0 && !batches)) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "bad batches");
    if ((size_t)nBatches > c->freeSlots.size()) {
        char buf[128];
        std::snprintf(buf, sizeof buf, "addBatches: not enough free slots (have %zu, need %d)", c->freeSlots.size(),
                      nBatches);
        return set_error(ctx, SCR_ERR_NO_FREE_SLOTS, buf);
    }
    if (nBatches == 0) return SCR_OK;
    // validate everything before mutating (the reference traps part-way)
    for (int i = 0; i < nBatches; i++) {
        const scr_raster_batch& b = batches[i];
        if ((int64_t)b.numPoints > c->pointsPerBatch)
            return set_error(ctx, SCR_ERR_BATCH_TOO_LARGE, "batch numPoints exceeds slot capacity");
        if ((uint64_t)b.firstPoint + b.numPoints > nSourcePoints)
            return set_error(ctx, SCR_ERR_SOURCE_TOO_SMALL, "addBatches: source data too small for batch");
        if (b.numPoints > 0 && (!low || !med || !high || !colors || !levels))
            return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "addBatches: NULL source array");
    }
    SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int> slots((size_t)nBatches);
    for (int i = 0; i < nBatches; i++) {
        slots[(size_t)i] = c->freeSlots.back();
        c->freeSlots.pop_back();
    }
    const size_t ppb = (size_t)c->pointsPerBatch;
    // copy runs of adjacent slots fed by adjacent full source batches in one go
    int i = 0;
    while (i < nBatches) {
        int j = i;
        while (j + 1 < nBatches && slots[(size_t)j + 1] == slots[(size_t)j] + 1 &&
               batches[j].numPoints == ppb &&
               (uint64_t)batches[j + 1].firstPoint == (uint64_t)batches[j].firstPoint + ppb)
            j++;
        const size_t src = batches[i].firstPoint;
        const size_t dst = (size_t)slots[(size_t)i] * ppb;
        const size_t n = (size_t)(j - i) * ppb + batches[j].numPoints;
        if (n > 0) {
            SCR_HIP(ctx, hipMemcpy(c->dLow + dst, low + src, n * 4, hipMemcpyHostToDevice));
            SCR_HIP(ctx, hipMemcpy(c->dMed + dst, med + src, n * 4, hipMemcpyHostToDevice));
            SCR_HIP(ctx, hipMemcpy(c->dHigh + dst, high + src, n * 4, hipMemcpyHostToDevice));
            SCR_HIP(ctx, hipMemcpy(c->dColors + dst, colors + src, n * 4, hipMemcpyHostToDevice));
            SCR_HIP(ctx, hipMemcpy(c->dLevels + dst, levels + src, n, hipMemcpyHostToDevice));
        }
        i = j + 1;
    }
    for (int k = 0; k < nBatches; k++) {
        const int slot = slots[(size_t)k];
        scr_raster_batch b = batches[k];
        b.firstPoint = (uint32_t)((size_t)slot * ppb);
        b.state = 1;
        c->mirror[(size_t)slot] = b;
        mark_dirty(c, slot);
        c->residentBatchCount += 1;
        c->pointCount += b.numPoints;
        if (outSlots) outSlots[k] = slot;
    }
    if (commit) return flush_mirror(c);
    return SCR_OK;
}

int scr_cloud_remove_batches(scr_cloud* c, const int32_t* slots, int n, int commit) {
    if (!c) return SCR_ERR_INVALID_ARGUMENT;
    if (n <= 0) return SCR_OK;
    if (!slots) return set_error(c->ctx, SCR_ERR_INVALID_ARGUMENT, "slots is NULL");
    for (int i = 0; i < n; i++)
        if (slots[i] < 0 || slots[i] >= c->maxResidentBatches)
            return set_error(c->ctx, SCR_ERR_SLOT_RANGE, "removeBatches: slot out of range");
    for (int i = 0; i < n; i++) {
        const int s = slots[i];
        scr_raster_batch& b = c->mirror[(size_t)s];
        if (b.state == 0) continue;
        c->pointCount -= b.numPoints;
        c->residentBatchCount -= 1;
        b.state = 0;
        b.numPoints = 0;
        mark_dirty(c, s);
        c->freeSlots.push_back(s);
    }
    if (commit) return flush_mirror(c);
    return SCR_OK;
}

int scr_cloud_commit(scr_cloud* c) {
    if (!c) return SCR_ERR_INVALID_ARGUMENT;
    return flush_mirror(c);
}

int scr_cloud_clear(scr_cloud* c) {
    if (!c) return SCR_ERR_INVALID_ARGUMENT;
    for (auto& b : c->mirror) {
        b.state = 0;
        b.numPoints = 0;
    }
    c->freeSlots.resize((size_t)c->maxResidentBatches);
    for (int i = 0; i < c->maxResidentBatches; i++) c->freeSlots[(size_t)i] = c->maxResidentBatches - 1 - i;
    c->residentBatchCount = 0;
    c->pointCount = 0;
    c->dirtyLo = 0;
    c->dirtyHi = c->maxResidentBatches - 1;
    return flush_mirror(c);
}

int scr_cloud_set_model(scr_cloud* c, const float model[16]) {
    if (!c || !model) return SCR_ERR_INVALID_ARGUMENT;
    std::memcpy(c->model, model, sizeof c->model);
    return SCR_OK;
}

int scr_cloud_update_files(scr_cloud* c, const float vp[16], const float model[16], const float* frustum16,
                           const float* prevVP16) {
    if (!c || !vp || !model) return SCR_ERR_INVALID_ARGUMENT;
    return upload_files(c, model, vp, frustum16, prevVP16);
}

int scr_cloud_set_tints(scr_cloud* c, const float* rgba, size_t nPoints, int onDevice) {
    if (!c) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = c->ctx;
    SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!rgba) {
        if (c->tintsOwned) dfree(c->dTints);
        c->dTints = nullptr;
        c->tintsOwned = false;
        return SCR_OK;
    }
    if (nPoints != c->poolPoints())
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "tint buffer must hold maxResidentBatches * pointsPerBatch");
    if (!c->dTints || !c->tintsOwned) {
        c->dTints = nullptr;
        SCR_HIP(ctx, dalloc(c->dTints, nPoints));
        c->tintsOwned = true;
    }
    SCR_HIP(ctx, hipMemcpy(c->dTints, rgba, nPoints * sizeof(float4),
                           onDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
    return SCR_OK;
}

int scr_cloud_set_displacements(scr_cloud* c, const float* xyz4, size_t nPoints, int onDevice) {
    if (!c) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = c->ctx;
    SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!xyz4) {
        drop_displacements(c);
        return SCR_OK;
    }
    if (nPoints != c->poolPoints())
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT,
                         "displacement buffer must hold maxResidentBatches * pointsPerBatch");
    if (!c->dDisp) SCR_HIP(ctx, dalloc(c->dDisp, nPoints));
    SCR_HIP(ctx, hipMemcpy(c->dDisp, xyz4, nPoints * sizeof(float4),
                           onDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
    return SCR_OK;
}

int scr_cloud_device_view(scr_cloud* c, int flags, scr_cloud_view* out) {
    if (!c || !out) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = c->ctx;
    const int rc = flush_mirror(c);
    if (rc != SCR_OK) return rc;
    if ((flags & SCR_VIEW_DISPLACEMENTS) && !c->dDisp) {
        SCR_HIP(ctx, dalloc(c->dDisp, c->poolPoints()));
        SCR_HIP(ctx, scr::launch_fill_u32(ctx->stream, reinterpret_cast<uint32_t*>(c->dDisp), c->poolPoints() * 4,
                                          0u));
    }
    if ((flags & SCR_VIEW_TINTS) && !c->dTints) {
        SCR_HIP(ctx, dalloc(c->dTints, c->poolPoints()));
        c->tintsOwned = true;
        SCR_HIP(ctx, scr::launch_fill_u32(ctx->stream, reinterpret_cast<uint32_t*>(c->dTints), c->poolPoints() * 4,
                                          0u));
    }
    out->batches = c->dBatches;
    out->xyzLow = c->dLow;
    out->xyzMed = c->dMed;
    out->xyzHigh = c->dHigh;
    out->colors = c->dColors;
    out->levels = c->dLevels;
    out->tints = reinterpret_cast<float*>(c->dTints);
    out->displacements = reinterpret_cast<float*>(c->dDisp);
    out->pointsPerBatch = (uint32_t)c->pointsPerBatch;
    out->slots = (uint32_t)c->maxResidentBatches;
    out->stream = ctx->stream;
    return SCR_OK;
}

int scr_cloud_read_pool(scr_cloud* c, int which, void* host, size_t bytes) {
    if (!c || !host) return SCR_ERR_INVALID_ARGUMENT;
    const void* src = nullptr;
    size_t need = 0;
    const size_t pts = c->poolPoints();
    switch (which) {
        case SCR_POOL_XYZ_LOW: src = c->dLow; need = pts * 4; break;
        case SCR_POOL_XYZ_MED: src = c->dMed; need = pts * 4; break;
        case SCR_POOL_XYZ_HIGH: src = c->dHigh; need = pts * 4; break;
        case SCR_POOL_COLORS: src = c->dColors; need = pts * 4; break;
        case SCR_POOL_LEVELS: src = c->dLevels; need = pts; break;
        case SCR_POOL_BATCHES: src = c->dBatches; need = (size_t)c->maxResidentBatches * sizeof(scr_raster_batch); break;
        default: return set_error(c->ctx, SCR_ERR_INVALID_ARGUMENT, "unknown pool buffer");
    }
    if (bytes < need) return set_error(c->ctx, SCR_ERR_INVALID_ARGUMENT, "host buffer too small");
    SCR_HIP(c->ctx, hipStreamSynchronize(c->ctx->stream));
    SCR_HIP(c->ctx, hipMemcpy(host, src, need, hipMemcpyDeviceToHost));
    return SCR_OK;
}

int scr_cloud_read_visible(scr_cloud* c, scr_visible_batch* host, int capacity, int* count) {
    if (!c || !count) return SCR_ERR_INVALID_ARGUMENT;
    uint32_t n = 0;
    SCR_HIP(c->ctx, hipStreamSynchronize(c->ctx->stream));
    SCR_HIP(c->ctx, hipMemcpy(&n, c->dCounter + c->counterWord, 4, hipMemcpyDeviceToHost));
    *count = (int)n;
    const int k = std::min((int)n, capacity);
    if (k > 0 && host) SCR_HIP(c->ctx, hipMemcpy(host, c->dVisible, (size_t)k * sizeof(scr_visible_batch), hipMemcpyDeviceToHost));
    return SCR_OK;
}

// ------------------------------------------------------------------ GPU packer

void scr_pack_options_default(scr_pack_options* o) {
    if (!o) return;
    o->keyBits = 30;
    o->lodLevels = 4;
    o->coarseVoxelDivisions = 64;
    o->inputOnDevice = 0;
}

namespace {

// Device copies of host inputs (freed by the destructor).
struct PackInputs {
    const float4* pos = nullptr;
    const float4* col = nullptr;
    float4* ownPos = nullptr;
    float4* ownCol = nullptr;
    ~PackInputs() {
        if (ownPos) (void)hipFree(ownPos);
        if (ownCol) (void)hipFree(ownCol);
    }
};

int stage_inputs(scr_ctx* ctx, const float* p, const float* c, size_t n, bool onDevice, PackInputs& in) {
    if (onDevice) {
        in.pos = (const float4*)p;
        in.col = (const float4*)c;
        return SCR_OK;
    }
    SCR_HIP(ctx, hipMalloc((void**)&in.ownPos, n * sizeof(float4)));
    SCR_HIP(ctx, hipMalloc((void**)&in.ownCol, n * sizeof(float4)));
    SCR_HIP(ctx, hipMemcpy(in.ownPos, p, n * sizeof(float4), hipMemcpyHostToDevice));
    SCR_HIP(ctx, hipMemcpy(in.ownCol, c, n * sizeof(float4), hipMemcpyHostToDevice));
    in.pos = in.ownPos;
    in.col = in.ownCol;
    return SCR_OK;
}

int check_pack(scr_cloud* c, const float* p, const float* col, size_t count, const scr_pack_options* opt) {
    if (!c) return SCR_ERR_INVALID_ARGUMENT;
    if (!opt) return set_error(c->ctx, SCR_ERR_INVALID_ARGUMENT, "pack options are NULL");
    if (opt->keyBits != 30 && opt->keyBits != 63) return set_error(c->ctx, SCR_ERR_INVALID_ARGUMENT, "keyBits must be 30 or 63");
    if (count && (!p || !col)) return set_error(c->ctx, SCR_ERR_INVALID_ARGUMENT, "NULL positions/colors");
    if (count > 0xffffffffull) return set_error(c->ctx, SCR_ERR_INVALID_ARGUMENT, "more than 2^32 - 1 points");
    return SCR_OK;
}

scr::PackDest dest_at(scr_cloud* c, int firstSlot) {
    const size_t base = (size_t)firstSlot * (size_t)c->pointsPerBatch;
    scr::PackDest d;
    d.low = c->dLow + base;
    d.med = c->dMed + base;
    d.high = c->dHigh + base;
    d.colors = c->dColors + base;
    d.levels = c->dLevels + base;
    d.batches = c->dBatches + firstSlot;
    d.slotBase = (uint32_t)firstSlot;
    d.ppb = (uint32_t)c->pointsPerBatch;
    return d;
}

}  // namespace

int scr_pack(scr_cloud* c, const float* positions4, const float* colors4, size_t count, const scr_pack_options* opt,
             uint32_t* sourceIndicesOut) {
    int rc = check_pack(c, positions4, colors4, count, opt);
    if (rc != SCR_OK) return rc;
    scr_ctx* ctx = c->ctx;
    if (count == 0) return scr_cloud_clear(c);
    const size_t ppb = (size_t)c->pointsPerBatch;
    const size_t nb = (count + ppb - 1) / ppb;
    if (nb > 0x7fffffffull) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "too many batches");
    SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if ((int)nb > c->maxResidentBatches) {  // grow the pool, keeping pointsPerBatch
        free_pool(c);
        c->maxResidentBatches = (int)nb;
        SCR_HIP(ctx, alloc_pool(c));
    }
    // prepareForGPUPack: slots [0, nb) resident, the rest empty
    for (int s = 0; s < c->maxResidentBatches; s++) {
        scr_raster_batch& b = c->mirror[(size_t)s];
        std::memset(&b, 0, sizeof b);
        if ((size_t)s < nb) {
            b.state = 1;
            b.firstPoint = (uint32_t)((size_t)s * ppb);
            b.numPoints = (uint32_t)std::min(ppb, count - (size_t)s * ppb);
        }
    }
    c->freeSlots.clear();
    for (int s = c->maxResidentBatches - 1; s >= (int)nb; s--) c->freeSlots.push_back(s);
    c->residentBatchCount = (int)nb;
    c->pointCount = (int64_t)count;
    c->dirtyLo = 0;
    c->dirtyHi = c->maxResidentBatches - 1;
    rc = flush_mirror(c);
    if (rc != SCR_OK) return rc;
    PackInputs in;
    rc = stage_inputs(ctx, positions4, colors4, count, opt->inputOnDevice != 0, in);
    if (rc != SCR_OK) return rc;
    rc = scr::pack_points(ctx, in.pos, in.col, count, *opt, dest_at(c, 0), sourceIndicesOut);
    if (rc != SCR_OK) return rc;
    // adoptGPUBatchBounds (ComputeRasteriserPointCloud.swift:313-320)
    SCR_HIP(ctx, hipMemcpy(c->mirror.data(), c->dBatches, nb * sizeof(scr_raster_batch), hipMemcpyDeviceToHost));
    return SCR_OK;
}

int scr_pack_chunk(scr_cloud* c, const float* positions4, const float* colors4, size_t count,
                   const scr_pack_options* opt, int32_t* outFirstSlot) {
    int rc = check_pack(c, positions4, colors4, count, opt);
    if (rc != SCR_OK) return rc;
    scr_ctx* ctx = c->ctx;
    if (outFirstSlot) *outFirstSlot = -1;
    if (count == 0) return SCR_OK;
    const size_t ppb = (size_t)c->pointsPerBatch;
    const int nb = (int)((count + ppb - 1) / ppb);
    // reserveContiguousSlots — ComputeRasteriserPointCloud.swift:330-347 (first fit)
    int first = -1;
    for (int s = 0; s + nb <= c->maxResidentBatches;) {
        int k = 0;
        while (k < nb && c->mirror[(size_t)(s + k)].state == 0) k++;
        if (k == nb) {
            first = s;
            break;
        }
        s += k + 1;
    }
    if (first < 0) return set_error(ctx, SCR_ERR_NO_FREE_SLOTS, "no run of free contiguous slots for the chunk");
    c->freeSlots.erase(std::remove_if(c->freeSlots.begin(), c->freeSlots.end(),
                                      [&](int s) { return s >= first && s < first + nb; }),
                       c->freeSlots.end());
    PackInputs in;
    rc = stage_inputs(ctx, positions4, colors4, count, opt->inputOnDevice != 0, in);
    if (rc == SCR_OK) rc = scr::pack_points(ctx, in.pos, in.col, count, *opt, dest_at(c, first), nullptr);
    if (rc != SCR_OK) {
        for (int s = first + nb - 1; s >= first; s--) c->freeSlots.push_back(s);
        return rc;
    }
    // finalizeGPUChunk — ComputeRasteriserPointCloud.swift:351-364
    SCR_HIP(ctx, hipMemcpy(c->mirror.data() + first, c->dBatches + first, (size_t)nb * sizeof(scr_raster_batch),
                           hipMemcpyDeviceToHost));
    c->residentBatchCount += nb;
    c->pointCount += (int64_t)count;
    if (outFirstSlot) *outFirstSlot = first;
    return SCR_OK;
}

// ------------------------------------------------------------------ frames

int scr_frame_create(scr_ctx* ctx, int width, int height, scr_frame** out) {
    if (!ctx || !out) return SCR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (width <= 0 || height <= 0 || (int64_t)width * height > (int64_t)1 << 31)
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "bad frame size");
    scr_frame* f = new (std::nothrow) scr_frame();
    if (!f) return SCR_ERR_OUT_OF_MEMORY;
    f->ctx = ctx;
    f->width = width;
    f->height = height;
    const size_t n = f->pixels();
    hipError_t e = dalloc(f->dPlanes, 5 * f->planeStride());
    if (e == hipSuccess) e = dalloc(f->dNearest, n);
    if (e == hipSuccess) e = dalloc(f->dRGBA, n);
    if (e == hipSuccess) e = dalloc(f->dDepthF, n);
    if (e == hipSuccess) e = scr::launch_fill_u32(ctx->stream, f->dPlanes, 5 * f->planeStride(), 0u);
    if (e == hipSuccess) e = scr::launch_fill_u64(ctx->stream, f->dNearest, n, 0ull);
    if (e == hipSuccess) e = scr::launch_fill_u32(ctx->stream, f->dRGBA, n, 0u);
    if (e == hipSuccess) e = scr::launch_fill_u32(ctx->stream, (uint32_t*)f->dDepthF, n, 0u);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        dfree(f->dPlanes);
        dfree(f->dNearest);
        dfree(f->dRGBA);
        dfree(f->dDepthF);
        delete f;
        return hip_error(ctx, e, "scr_frame_create");
    }
    scr::ctx_retain(ctx);
    *out = f;
    return SCR_OK;
}

int scr_frame_destroy(scr_frame* f) {
    if (!f) return SCR_OK;
    scr_ctx* ctx = f->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    dfree(f->dPlanes);
    dfree(f->dNearest);
    dfree(f->dRGBA);
    dfree(f->dRGBAScratch);
    dfree(f->dDepthF);
    dfree(f->dSums);
    delete f;
    scr::ctx_release(ctx);
    return SCR_OK;
}

int scr_frame_hole_fill(scr_frame* f, int iterations) {
    if (!f) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = f->ctx;
    if (iterations < 0) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    if (iterations == 0) return SCR_OK;
    if (!f->dRGBAScratch) SCR_HIP(ctx, dalloc(f->dRGBAScratch, f->pixels()));
    scr::PassScope scope(ctx, "hole_fill");
    // passes alternate plane -> scratch -> plane; an odd count ends with a copy back
    uint32_t* src = f->dRGBA;
    uint32_t* dst = f->dRGBAScratch;
    for (int k = 0; k < iterations; k++) {
        SCR_HIP(ctx, scr::launch_hole_fill(ctx->stream, src, dst, f->width, f->height));
        std::swap(src, dst);
    }
    if (src != f->dRGBA)
        SCR_HIP(ctx, hipMemcpyAsync(f->dRGBA, src, f->pixels() * 4, hipMemcpyDeviceToDevice, ctx->stream));
    return SCR_OK;
}

int scr_frame_read(scr_frame* f, int which, void* host, size_t bytes) {
    if (!f || !host) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = f->ctx;
    const size_t n = f->pixels();
    if (bytes < n * 4) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "host buffer too small");
    SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (which >= SCR_PLANE_DEPTH && which <= SCR_PLANE_COUNT) {
        SCR_HIP(ctx, hipMemcpy(host, f->plane(which), n * 4, hipMemcpyDeviceToHost));
    } else if (which == SCR_PLANE_NEAREST_DEPTH || which == SCR_PLANE_NEAREST_INDEX) {
        std::vector<unsigned long long> tmp(n);
        SCR_HIP(ctx, hipMemcpy(tmp.data(), f->dNearest, n * 8, hipMemcpyDeviceToHost));
        uint32_t* o = (uint32_t*)host;
        for (size_t i = 0; i < n; i++)
            o[i] = which == SCR_PLANE_NEAREST_DEPTH ? (uint32_t)(tmp[i] >> 32) : ~(uint32_t)tmp[i];
    } else if (which == SCR_PLANE_RGBA8) {
        SCR_HIP(ctx, hipMemcpy(host, f->dRGBA, n * 4, hipMemcpyDeviceToHost));
    } else if (which == SCR_PLANE_DEPTH_F32) {
        SCR_HIP(ctx, hipMemcpy(host, f->dDepthF, n * 4, hipMemcpyDeviceToHost));
    } else if (which >= SCR_PLANE_DIFF_SUM_R && which <= SCR_PLANE_DIFF_SUM_B) {
        if (bytes < n * 8) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "host buffer too small (8 B per pixel)");
        if (!f->dSums) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "frame has no differentiable render");
        SCR_HIP(ctx, hipMemcpy(host, f->dSums + (size_t)(which - SCR_PLANE_DIFF_SUM_R) * f->planeStride(), n * 8,
                               hipMemcpyDeviceToHost));
    } else {
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "unknown plane");
    }
    return SCR_OK;
}

// ------------------------------------------------------------------ composite

void scr_composite_options_default(scr_composite_options* o) {
    if (!o) return;
    o->mode = SCR_COMPOSITE_DEPTH;
    o->blend = SCR_BLEND_OVER;
    o->depthCompare = SCR_COMPARE_GREATER;
    o->writeDepth = 0;
}

int scr_frame_composite(scr_frame* frame, const scr_composite_options* opt, uint32_t* targetRGBA8, float* targetDepth,
                        int targetPitch) {
    if (!frame) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = frame->ctx;
    if (!opt || !targetRGBA8) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "NULL options or colour target");
    if (opt->mode != SCR_COMPOSITE_OVER && opt->mode != SCR_COMPOSITE_DEPTH)
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "unknown composite mode");
    if (opt->blend != SCR_BLEND_REPLACE && opt->blend != SCR_BLEND_OVER)
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "unknown blend");
    if (opt->depthCompare < SCR_COMPARE_ALWAYS || opt->depthCompare > SCR_COMPARE_LESS_EQUAL)
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "unknown depth compare");
    if (opt->mode == SCR_COMPOSITE_DEPTH && !targetDepth)
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "the depth composite needs a depth target");
    if (targetPitch < frame->width) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "target pitch below the width");
    scr::PassScope ps(ctx, "composite");
    SCR_HIP(ctx, scr::launch_composite(ctx->stream, frame->dRGBA, frame->dDepthF, frame->width, frame->height, *opt,
                                       targetRGBA8, opt->mode == SCR_COMPOSITE_DEPTH ? targetDepth : nullptr,
                                       targetPitch));
    return SCR_OK;
}

// ------------------------------------------------------------------ render

namespace {

// The frame's front: the clear of `words` u32 at `planes` and, when the
// clouds' composed file tables fit the launch (at most kSetupClouds clouds
// and kSetupFiles files), their tables (in place of upload_files) and their
// culls in the same launch; *fused tells which.
int frame_front(scr_ctx* ctx, uint32_t* planes, size_t words, scr_cloud* const* clouds, int nClouds,
                const float* vp, const float* prevVP, const scr::PassParams& p, const char* clearPass, bool* fused) {
    scr::FrameSetup su{};
    scr::CullSet cs{};
    su.planes = planes;
    su.words = words;
    size_t files = 0;
    for (int i = 0; i < nClouds; i++) files += clouds[i]->files.size();
    *fused = nClouds > 0 && nClouds <= scr::kSetupClouds && files <= (size_t)scr::kSetupFiles;
    if (*fused) {
        int nf = 0, blocks = 0;
        for (int i = 0; i < nClouds; i++) {
            scr_cloud* c = clouds[i];
            CloudDev d = c->dev();
            su.zero[i] = d.counter;                           // the word the draws read until now
            d.counter = c->dCounter + (c->counterWord ^ 1);   // left zero by the previous front (or never used)
            cs.c[i] = d;
            cs.blockStart[i] = blocks;
            blocks += (c->maxResidentBatches + scr::kCullSlotsPerBlock - 1) / scr::kCullSlotsPerBlock;
            su.dst[i] = c->dFiles;
            su.first[i] = nf;
            su.count[i] = (int)c->files.size();
            scr::compose_files(c, c->model, vp, nullptr, prevVP, su.files + nf);
            nf += su.count[i];
        }
        cs.blockStart[nClouds] = blocks;
        su.nClouds = cs.nClouds = nClouds;
    }
    scr::PassScope ps(ctx, *fused ? "clear_cull" : clearPass);
    SCR_HIP(ctx, scr::launch_frame_front(ctx->stream, su, cs, p));
    if (*fused)
        for (int i = 0; i < nClouds; i++) clouds[i]->counterWord ^= 1;
    return SCR_OK;
}

}  // namespace

int scr_render(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, scr_cloud* const* clouds, int nClouds,
               scr_frame* frame) {
    return scr_render_camera(ctx, cfg, cam, 0, clouds, nClouds, frame);
}

int scr_render_camera(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, uint64_t cameraKey,
                      scr_cloud* const* clouds, int nClouds, scr_frame* frame) {
    if (!ctx || !cam || !frame || nClouds < 0 || (nClouds > 0 && !clouds)) return SCR_ERR_INVALID_ARGUMENT;
    int rc = check_config(ctx, cfg, true);
    if (rc != SCR_OK) return rc;
    for (int i = 0; i < nClouds; i++)
        if (!clouds[i] || clouds[i]->ctx != ctx) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "bad cloud handle");
    const scr::PassParams p = make_params(cfg, cam, frame->width, frame->height);
    float vp[16];
    scr::mat_mul(cam->projectionMatrix, cam->viewMatrix, vp);
    hipStream_t s = ctx->stream;
    // the previous view-projection of this camera, for motion-blur velocity
    // (ComputeRasteriser.swift:255-266: the current one the first time)
    float prevVP[16];
    auto seen = frame->prevVP.find(cameraKey);
    std::memcpy(prevVP, seen != frame->prevVP.end() ? seen->second.data() : vp, sizeof prevVP);
    std::array<float, 16> cur;
    std::memcpy(cur.data(), vp, sizeof vp);
    frame->prevVP[cameraKey] = cur;

    if (cfg->mode == SCR_MODE_NEAREST_POINT) {
        // encodeNearestPoint: first cloud with resident batches only (:483-497)
        scr_cloud* c = nullptr;
        for (int i = 0; i < nClouds && !c; i++)
            if (clouds[i]->residentBatchCount > 0) c = clouds[i];
        scr_cloud* const one[1] = {c};
        bool fused = false;
        rc = frame_front(ctx, reinterpret_cast<uint32_t*>(frame->dNearest), 2 * frame->pixels(), one, c ? 1 : 0, vp,
                         prevVP, p, "clear_winner", &fused);
        if (rc != SCR_OK) return rc;
        const uint32_t* colors = nullptr;
        if (c) {
            const CloudDev d = c->dev();
            if (!fused) {
                rc = upload_files(c, c->model, vp, nullptr, prevVP);
                if (rc != SCR_OK) return rc;
                scr::PassScope ps(ctx, "cull");
                SCR_HIP(ctx, scr::launch_cull(s, d, p));
            }
            {
                scr::PassScope ps(ctx, "nearest");
                SCR_HIP(ctx, scr::launch_nearest(s, ctx->numCUs, d, p, frame->dNearest));
            }
            colors = c->dColors;
        }
        {
            scr::PassScope ps(ctx, "nearest_resolve");
            SCR_HIP(ctx, scr::launch_nearest_resolve(s, p, frame->dNearest, colors, frame->dRGBA, frame->dDepthF));
        }
        // encodeHoleFill (ComputeRasteriser.swift:298-320), after either mode
        return cfg->holeFillIterations > 0 ? scr_frame_hole_fill(frame, cfg->holeFillIterations) : SCR_OK;
    }

    // encodeHighQualityAverage (:322-456)
    std::vector<scr_cloud*> resident;
    for (int i = 0; i < nClouds; i++)
        if (clouds[i]->residentBatchCount > 0) resident.push_back(clouds[i]);
    bool fused = false;
    rc = frame_front(ctx, frame->dPlanes, 5 * frame->planeStride(), resident.data(), (int)resident.size(), vp, prevVP,
                     p, "clear", &fused);
    if (rc != SCR_OK) return rc;
    if (!fused) {
        for (scr_cloud* c : resident) {
            rc = upload_files(c, c->model, vp, nullptr, prevVP);
            if (rc != SCR_OK) return rc;
        }
    }
    // each cloud's device view; the previous displacements of this camera
    // (the current ones until a motion-blurred render of it has kept a copy,
    // ColorPassProcessor: prevDisplacementBuffer ?? displacementBuffer)
    std::vector<CloudDev> devs;
    for (scr_cloud* c : resident) {
        CloudDev d = c->dev();
        auto it = c->prevDisp.find(cameraKey);
        if (it != c->prevDisp.end()) d.prevDisp = it->second;
        devs.push_back(d);
    }
    if (!fused) {
        for (const CloudDev& d : devs) {
            scr::PassScope ps(ctx, "cull");
            SCR_HIP(ctx, scr::launch_cull(s, d, p));
        }
    }
    // translucent defocus and motion blur take the weighted-blend passes
    // (ComputeRasteriser.swift:233-246)
    scr::CoverageParams q;
    q.coverageMode = (cfg->applyTint != 0 && cfg->tintAlphaIsCoverage != 0) ? 1 : 0;
    q.motionBlur = cfg->motionBlur;
    q.samples = cfg->motionBlurSamples;
    q.maxSpread = cfg->motionBlurMaxSpread;
    const bool blend = q.coverageMode != 0 || cfg->motionBlur > 0.0f;
    for (const CloudDev& d : devs) {
        scr::PassScope ps(ctx, "depth");
        if (q.coverageMode != 0)
            SCR_HIP(ctx, scr::launch_depth_coverage(s, ctx->numCUs, d, p, frame->plane(0)));
        else
            SCR_HIP(ctx, scr::launch_depth(s, ctx->numCUs, d, p, frame->plane(0)));
    }
    for (const CloudDev& d : devs) {
        scr::PassScope ps(ctx, "color");
        if (blend)
            SCR_HIP(ctx, scr::launch_color_coverage(s, ctx->numCUs, d, p, q, frame->plane(1), frame->plane(2),
                                                    frame->plane(3), frame->plane(4)));
        else
            SCR_HIP(ctx, scr::launch_color(s, ctx->numCUs, d, p, frame->plane(0), frame->plane(1),
                                           frame->plane(2), frame->plane(3), frame->plane(4)));
    }
    // with motion blur, each displaced cloud keeps this frame's displacements
    // as this camera's previous ones (ComputeRasteriser.swift:451-480)
    if (cfg->motionBlur > 0.0f) {
        for (scr_cloud* c : resident) {
            if (!c->dDisp) continue;
            auto it = c->prevDisp.find(cameraKey);
            if (it == c->prevDisp.end()) {
                float4* fresh = nullptr;
                SCR_HIP(ctx, dalloc(fresh, c->poolPoints()));
                it = c->prevDisp.emplace(cameraKey, fresh).first;
            }
            SCR_HIP(ctx, hipMemcpyAsync(it->second, c->dDisp, c->poolPoints() * sizeof(float4),
                                        hipMemcpyDeviceToDevice, s));
        }
    }
    {
        scr::PassScope ps(ctx, "resolve");
        if (blend)
            SCR_HIP(ctx, scr::launch_resolve_coverage(s, p, frame->dPlanes, frame->planeStride(), frame->dRGBA,
                                                      frame->dDepthF));
        else
            SCR_HIP(ctx, scr::launch_resolve(s, p, frame->dPlanes, frame->planeStride(), frame->dRGBA,
                                             frame->dDepthF));
    }
    return cfg->holeFillIterations > 0 ? scr_frame_hole_fill(frame, cfg->holeFillIterations) : SCR_OK;
}

// ------------------------------------------------------------------ differentiable HQS

namespace {

int check_diff(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, scr_cloud* cloud, scr_frame* frame) {
    if (!cam || !cloud || !frame) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "NULL camera, cloud or frame");
    if (cloud->ctx != ctx || frame->ctx != ctx) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "foreign handle");
    int rc = check_config(ctx, cfg);
    if (rc != SCR_OK) return rc;
    if (cfg->mode != SCR_MODE_HIGH_QUALITY_AVERAGE)
        return set_error(ctx, SCR_ERR_UNSUPPORTED, "differentiable rendering is an HQS mode");
    if (cfg->applyTint != 0 || cfg->colorizeChunks != 0 || cfg->colorizeOverdraw != 0)
        return set_error(ctx, SCR_ERR_UNSUPPORTED, "differentiable rendering takes its colours from the parameters");
    return SCR_OK;
}

}  // namespace

int scr_diff_render(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, scr_cloud* cloud,
                    const float* params, scr_frame* frame, float* image) {
    if (!ctx) return SCR_ERR_INVALID_ARGUMENT;
    int rc = check_diff(ctx, cfg, cam, cloud, frame);
    if (rc != SCR_OK) return rc;
    if (!params || !image) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "NULL params or image");
    if (!frame->dSums) {
        SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        SCR_HIP(ctx, hipMalloc((void**)&frame->dSums, 3 * frame->planeStride() * sizeof(unsigned long long)));
    }
    rc = flush_mirror(cloud);
    if (rc != SCR_OK) return rc;
    const scr::PassParams p = make_params(cfg, cam, frame->width, frame->height);
    float vp[16];
    scr::mat_mul(cam->projectionMatrix, cam->viewMatrix, vp);
    hipStream_t s = ctx->stream;
    {
        scr::PassScope ps(ctx, "clear");
        SCR_HIP(ctx, scr::launch_fill_u32(s, frame->dPlanes, 5 * frame->planeStride(), 0u));
        SCR_HIP(ctx, scr::launch_fill_u64(s, frame->dSums, 3 * frame->planeStride(), 0ull));
    }
    if (cloud->residentBatchCount > 0) {
        rc = upload_files(cloud, cloud->model, vp, nullptr, nullptr);
        if (rc != SCR_OK) return rc;
        const CloudDev d = cloud->dev();
        {
            scr::PassScope ps(ctx, "cull");
            SCR_HIP(ctx, scr::launch_cull(s, d, p));
        }
        {
            scr::PassScope ps(ctx, "depth");
            SCR_HIP(ctx, scr::launch_depth(s, ctx->numCUs, d, p, frame->plane(0)));
        }
        scr::PassScope ps(ctx, "diff_color");
        SCR_HIP(ctx, scr::launch_diff_color(s, ctx->numCUs, d, p, reinterpret_cast<const float4*>(params),
                                            frame->plane(0), frame->dSums, frame->planeStride(), frame->plane(4)));
    }
    scr::PassScope ps(ctx, "diff_resolve");
    SCR_HIP(ctx, scr::launch_diff_resolve(s, p, frame->dSums, frame->planeStride(), frame->plane(4), image));
    return SCR_OK;
}

int scr_diff_backward(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, scr_cloud* cloud,
                      scr_frame* frame, const float* dLdImage, double* grad) {
    if (!ctx) return SCR_ERR_INVALID_ARGUMENT;
    int rc = check_diff(ctx, cfg, cam, cloud, frame);
    if (rc != SCR_OK) return rc;
    if (!dLdImage || !grad) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "NULL dLdImage or grad");
    if (!frame->dSums) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "frame has no differentiable forward");
    if (cloud->residentBatchCount == 0) return SCR_OK;
    rc = flush_mirror(cloud);
    if (rc != SCR_OK) return rc;
    const scr::PassParams p = make_params(cfg, cam, frame->width, frame->height);
    float vp[16];
    scr::mat_mul(cam->projectionMatrix, cam->viewMatrix, vp);
    hipStream_t s = ctx->stream;
    rc = upload_files(cloud, cloud->model, vp, nullptr, nullptr);
    if (rc != SCR_OK) return rc;
    const CloudDev d = cloud->dev();
    {
        scr::PassScope ps(ctx, "cull");
        SCR_HIP(ctx, scr::launch_cull(s, d, p));
    }
    scr::PassScope ps(ctx, "diff_backward");
    SCR_HIP(ctx, scr::launch_diff_backward(s, ctx->numCUs, d, p, frame->plane(0), frame->plane(4), dLdImage, grad));
    return SCR_OK;
}

int scr_pick(scr_ctx* ctx, const scr_config* cfg, const scr_camera* cam, scr_cloud* cloud, scr_frame* frame,
             float ndcX, float ndcY, int searchRadius, uint32_t* outIndex, int* found) {
    if (!ctx || !cfg || !cam || !cloud || !frame || !outIndex || !found) return SCR_ERR_INVALID_ARGUMENT;
    *found = 0;
    const int W = frame->width, H = frame->height;
    if (cloud->residentBatchCount <= 0) return SCR_OK;
    // the nearest index pass alone (ComputeRasteriser.swift:629-651): the
    // blend, hole-fill and displacement settings of the frame do not apply
    scr_config c2 = *cfg;
    c2.mode = SCR_MODE_NEAREST_POINT;
    c2.motionBlur = 0.0f;
    c2.tintAlphaIsCoverage = 0;
    c2.holeFillIterations = 0;
    c2.applyDisplacement = 0;
    int rc = check_config(ctx, &c2);
    if (rc != SCR_OK) return rc;
    // NDC (y up) → buffer pixel (row 0 at the top), ComputeRasteriser.swift:617-621
    int px = scr::f2i((ndcX * 0.5f + 0.5f) * (float)W);
    px = std::min(std::max(px, 0), W - 1);
    const int pyb = scr::f2i((ndcY * 0.5f + 0.5f) * (float)H);
    const int py = std::min(std::max(H - 1 - pyb, 0), H - 1);
    const scr::PassParams p = make_params(&c2, cam, W, H);
    float vp[16];
    scr::mat_mul(cam->projectionMatrix, cam->viewMatrix, vp);
    hipStream_t s = ctx->stream;
    rc = upload_files(cloud, cloud->model, vp, nullptr, nullptr);
    if (rc != SCR_OK) return rc;
    const CloudDev d = cloud->dev();
    SCR_HIP(ctx, scr::launch_fill_u64(s, frame->dNearest, frame->pixels(), 0ull));
    SCR_HIP(ctx, scr::launch_cull(s, d, p));
    SCR_HIP(ctx, scr::launch_nearest(s, ctx->numCUs, d, p, frame->dNearest));
    const int r = std::max(0, searchRadius);
    const int y0 = std::max(0, py - r), y1 = std::min(H - 1, py + r);
    const size_t rows = (size_t)(y1 - y0 + 1);
    std::vector<unsigned long long> band(rows * (size_t)W);
    SCR_HIP(ctx, hipStreamSynchronize(s));
    SCR_HIP(ctx, hipMemcpy(band.data(), frame->dNearest + (size_t)y0 * W, band.size() * 8, hipMemcpyDeviceToHost));
    long long best = 0x7fffffffffffffffLL;
    const long long r2 = (long long)r * r;
    for (int y = y0; y <= y1; y++) {
        const int dy = y - py;
        for (int x = std::max(0, px - r); x <= std::min(W - 1, px + r); x++) {
            const int dx = x - px;
            const long long dist = (long long)dx * dx + (long long)dy * dy;
            if (dist > r2 || dist >= best) continue;
            const uint32_t idx = ~(uint32_t)band[(size_t)(y - y0) * W + x];
            if (idx != 0xffffffffu) {
                best = dist;
                *outIndex = idx;
                *found = 1;
            }
        }
    }
    return SCR_OK;
}

}  // extern "C"
