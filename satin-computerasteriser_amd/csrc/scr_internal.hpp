// scr_internal.hpp — host-side state behind the C ABI (context, point-cloud
// slot pool, per-viewport frame buffers) and the kernel launch interface.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/scr.h"
#include "scr_math.hpp"

struct scr_ctx {
    // the context lives until scr_ctx_destroy and the destroy call of every
    // cloud, frame and multi-view renderer made on it (each holds a reference)
    int refs = 1;
    int device = 0;
    hipStream_t stream = nullptr;
    int numCUs = 0;
    std::string lastError;
    // per-pass timing
    bool profiling = false;
    struct Interval {
        const char* name;
        hipEvent_t begin, end;
    };
    std::vector<Interval> intervals;
    std::vector<hipEvent_t> eventPool;
    // scratch owned by the context (grown on demand, freed at destroy)
    void* scratch = nullptr;
    size_t scratchBytes = 0;
    uint32_t* packBounds = nullptr;  // the pack's six bounds words (256 B, allocated with the context)
};

// Device view of one cloud, passed by value to the point kernels.
struct CloudDev {
    const scr_raster_batch* batches;
    const scr_raster_file* files;
    const uint32_t* xyzLow;
    const uint32_t* xyzMed;
    const uint32_t* xyzHigh;
    const uint32_t* colors;
    const uint8_t* levels;
    const float4* tints;           // nullptr: zero stand-in
    const float4* disp;            // displacement (xyz), nullptr: zero stand-in
    const float4* prevDisp;        // previous frame's displacement (motion blur), nullptr: zero stand-in
    scr_visible_batch* visible;    // cull output
    scr::DrawRecord* bounds;       // cull output, one per visible entry (what its box proves, its draw context)
    uint32_t* counter;             // visible count
    int nSlots;
    int nFiles;
};

struct scr_cloud {
    scr_ctx* ctx = nullptr;
    int maxResidentBatches = 0;
    int pointsPerBatch = 0;
    std::vector<scr_raster_file> files;  // base table (world per file)
    float model[16];
    std::vector<scr_raster_batch> mirror;
    int dirtyLo = 0x7fffffff, dirtyHi = -1;
    std::vector<int> freeSlots;  // LIFO: back() is taken first
    int residentBatchCount = 0;
    int64_t pointCount = 0;
    // device buffers
    uint32_t* dLow = nullptr;
    uint32_t* dMed = nullptr;
    uint32_t* dHigh = nullptr;
    uint32_t* dColors = nullptr;
    uint8_t* dLevels = nullptr;
    scr_raster_batch* dBatches = nullptr;
    scr_raster_file* dFiles = nullptr;
    // pinned staging for the composed table: a ring of kFilesRing tables (the
    // reference keeps 6 file buffers in flight, ComputeRasteriserPointCloud.swift:141-146),
    // so the host composes frame k+6 while earlier copies are still queued
    static constexpr int kFilesRing = 6;
    scr_raster_file* hFiles = nullptr;
    hipEvent_t filesCopied[kFilesRing] = {};
    int filesNext = 0;
    float4* dTints = nullptr;
    bool tintsOwned = false;
    float4* dDisp = nullptr;       // displacement buffer (owned), float4 per pool point
    // per camera key: the displacements the last motion-blurred render of that
    // camera drew (ComputeRasteriserPointCloud.swift:96-103), owned
    std::map<uint64_t, float4*> prevDisp;
    scr_visible_batch* dVisible = nullptr;
    scr::DrawRecord* dBounds = nullptr;
    // two visible-count words: the draws read dCounter[counterWord]; a cull
    // fused into the frame's front counts into the other word (left zero)
    // and zeroes this one for the next
    uint32_t* dCounter = nullptr;
    int counterWord = 0;

    size_t poolPoints() const { return (size_t)maxResidentBatches * (size_t)pointsPerBatch; }
    CloudDev dev() const {
        CloudDev c;
        c.batches = dBatches;
        c.files = dFiles;
        c.xyzLow = dLow;
        c.xyzMed = dMed;
        c.xyzHigh = dHigh;
        c.colors = dColors;
        c.levels = dLevels;
        c.tints = dTints;
        c.disp = dDisp;
        c.prevDisp = dDisp;
        c.visible = dVisible;
        c.bounds = dBounds;
        c.counter = dCounter + counterWord;
        c.nSlots = maxResidentBatches;
        c.nFiles = (int)files.size();
        return c;
    }
};

struct scr_frame {
    scr_ctx* ctx = nullptr;
    int width = 0, height = 0;
    uint32_t* dPlanes = nullptr;  // depth, red, green, blue, count (SoA, 5 planes of planeStride)
    unsigned long long* dNearest = nullptr;  // depth << 32 | ~index per pixel
    unsigned long long* dSums = nullptr;     // differentiable mode: 3 int64 planes of planeStride (on first use)
    uint32_t* dRGBA = nullptr;
    uint32_t* dRGBAScratch = nullptr;  // hole fill ping-pong plane (on first use)
    // last view-projection rendered per camera key (motion-blur velocity)
    std::map<uint64_t, std::array<float, 16>> prevVP;
    float* dDepthF = nullptr;
    size_t pixels() const { return (size_t)width * (size_t)height; }
    // planes start on 256-byte boundaries so vector stores stay aligned
    size_t planeStride() const { return (pixels() + 63) & ~(size_t)63; }
    uint32_t* plane(int i) const { return dPlanes + (size_t)i * planeStride(); }
};

namespace scr {

// Brackets the launches issued during its lifetime with events when the
// context is profiling.
class PassScope {
   public:
    PassScope(scr_ctx* ctx, const char* name);
    ~PassScope();

   private:
    scr_ctx* ctx_;
    int index_ = -1;
};

int set_error(scr_ctx* ctx, int code, const std::string& msg);
int hip_error(scr_ctx* ctx, hipError_t e, const char* what);
void ctx_retain(scr_ctx* ctx);
void ctx_release(scr_ctx* ctx);   // frees the context with its last reference

// Shared by the single-view and multi-view render paths (scr_api.hip).
// allowCoverage: translucent defocus, motion blur and holeFillIterations are
// accepted (the single-view render implements them; multi-view and
// differentiable mode do not)
int check_config(scr_ctx* ctx, const scr_config* cfg, bool allowCoverage = false);
PassParams make_params(const scr_config* cfg, const scr_camera* cam, int width, int height);
int flush_mirror(scr_cloud* c);
// updateFiles composition (ComputeRasteriserPointCloud.swift:487-515) of every
// file of `c` for model matrix `model` and view-projection vp into
// out[0 .. files.size()).
void compose_files(const scr_cloud* c, const float* model, const float* vp, const float* frustum,
                   const float* prevVP, scr_raster_file* out);

// Kernel launchers (scr_render.hip).  All enqueue on ctx->stream.
hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v);
hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, size_t n, unsigned long long v);
hipError_t launch_cull(hipStream_t s, const CloudDev& c, const PassParams& p);
// A frame's front in one launch: the clouds' culls and the frame's set-up.
// The set-up zeroes `words` u32 of the accumulation planes (16-byte aligned)
// and, for each of nClouds clouds, zeroes the counter word `zero` (the one
// the cloud's next such cull counts into) and writes its composed file table,
// which the launch carries in its arguments (files[first .. first + count) to
// dst).  Cloud k's cull counts into cull.c[k].counter and reads its file
// table from the arguments, in blocks [blockStart[k], blockStart[k + 1]).
constexpr int kSetupClouds = 4;
constexpr int kSetupFiles = 8;
constexpr int kCullSlotsPerBlock = 128;   // a cull block of 256 threads: two waves per slot group
struct FrameSetup {
    uint32_t* planes = nullptr;
    size_t words = 0;
    int nClouds = 0;
    uint32_t* zero[kSetupClouds] = {};
    scr_raster_file* dst[kSetupClouds] = {};
    int first[kSetupClouds] = {};
    int count[kSetupClouds] = {};
    scr_raster_file files[kSetupFiles];
};
struct CullSet {
    int nClouds = 0;
    int blockStart[kSetupClouds + 1] = {};
    CloudDev c[kSetupClouds] = {};
};
hipError_t launch_frame_front(hipStream_t s, const FrameSetup& su, const CullSet& cs, const PassParams& p);
// Translucent-defocus / motion-blur settings of the coverage colour pass.
struct CoverageParams {
    int coverageMode;   // applyTint && tintAlphaIsCoverage
    float motionBlur;   // > 0: sweep each splat along its screen velocity
    int samples;        // motionBlurSamples
    float maxSpread;    // motionBlurMaxSpread
};
hipError_t launch_depth_coverage(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, uint32_t* depth);
hipError_t launch_color_coverage(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p,
                                 const CoverageParams& q, uint32_t* red, uint32_t* green, uint32_t* blue,
                                 uint32_t* count);
hipError_t launch_resolve_coverage(hipStream_t s, const PassParams& p, const uint32_t* planes, size_t stride,
                                   uint32_t* rgba, float* depthf);
hipError_t launch_hole_fill(hipStream_t s, const uint32_t* in, uint32_t* out, int w, int h);
hipError_t launch_composite(hipStream_t s, const uint32_t* rgba, const float* depthf, int w, int h,
                            const scr_composite_options& o, uint32_t* target, float* targetDepth, int pitch);
hipError_t launch_check_fast_rcp(hipStream_t s, int numCUs, unsigned long long* mismatches, uint32_t* first);
hipError_t launch_atomic_rate(hipStream_t s, int numCUs, uint32_t* buf, uint32_t spots, int add, int iters,
                              uint64_t* atomics);
hipError_t launch_depth(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, uint32_t* depth);
hipError_t launch_color(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, const uint32_t* depth,
                        uint32_t* red, uint32_t* green, uint32_t* blue, uint32_t* count);
hipError_t launch_resolve(hipStream_t s, const PassParams& p, const uint32_t* planes, size_t stride, uint32_t* rgba,
                          float* depthf);
hipError_t launch_nearest(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p,
                          unsigned long long* packed);
hipError_t launch_nearest_resolve(hipStream_t s, const PassParams& p, const unsigned long long* packed,
                                  const uint32_t* colors, uint32_t* rgba, float* depthf);

// Differentiable HQS (scr_diff.hip).
hipError_t launch_diff_color(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p, const float4* params,
                             const uint32_t* depth, unsigned long long* sums, size_t stride, uint32_t* count);
hipError_t launch_diff_resolve(hipStream_t s, const PassParams& p, const unsigned long long* sums, size_t stride,
                               const uint32_t* count, float* image);
hipError_t launch_diff_backward(hipStream_t s, int numCUs, const CloudDev& c, const PassParams& p,
                                const uint32_t* depth, const uint32_t* count, const float* dLdImage, double* grad);

// Pool destination of one pack (device pointers already offset to the slot run).
struct PackDest {
    uint32_t* low;
    uint32_t* med;
    uint32_t* high;
    uint32_t* colors;
    uint8_t* levels;
    scr_raster_batch* batches;
    uint32_t slotBase;
    uint32_t ppb;
};

// GPU pack of n points (device float4 positions / colours) into dst
// (scr_pack.hip), in the context's scratch, grown on demand.
int pack_points(scr_ctx* ctx, const float4* pos, const float4* col, size_t n, const scr_pack_options& opt,
                const PackDest& dst, uint32_t* hostSourceIndices);

}  // namespace scr
