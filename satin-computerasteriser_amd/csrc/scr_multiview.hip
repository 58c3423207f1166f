// scr_multiview.hip — batched rendering of many views (cameras of their own
// resolution, each over one cloud) with one launch per phase for all views.
//
// Per render:
//   k_mv_clear     every view's planes (blockIdx.y = view) and the queue heads
//   k_mv_cull      cullUpdate (Cull/Shaders.metal:19-83) for every (view, slot)
//                  pair, written to a dense item table (no append order: item
//                  (v, s) sits at itemBase[v] + s * itemStride[v], so the views
//                  of one cloud that see a batch sit next to each other and
//                  read its points while they are in the cache)
//   k_mv_depth / k_mv_color / k_mv_nearest
//                  persistent draw kernels; each workgroup takes the next work
//                  unit (item, split part) from a device counter
//   k_mv_resolve / k_mv_nearest_resolve
//
// The per-batch draw bodies are the single-view ones (scr_draw.hpp), so a view
// gets exactly the planes scr_render gives it: integer max / add do not depend
// on the order in which items or fragments arrive.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "scr_mv.hpp"

namespace scr {
using mv::DrawKernel;
using mv::kDepth;
using mv::kColor;
using mv::kNearest;
using mv::kQueues;
using mv::MVItem;
using mv::MVView;
using mv::Phase;

namespace {

using namespace draw;
using namespace mv;


// kClearHQSDepth / kClearHQSSums: the depth plane (and the queue heads) and
// the four sum planes of an HQS render separately, so that the sums are
// cleared on a side stream while the cull and depth passes run.
enum ClearMode { kClearHQS = 0, kClearNearest = 1, kClearDiff = 2, kClearHQSDepth = 3, kClearHQSSums = 4 };

__global__ __launch_bounds__(256) void k_mv_clear(const MVView* __restrict__ views, int mode,
                                                  uint32_t* __restrict__ heads) {
    const MVView& V = views[blockIdx.y];
    if (mode != kClearHQSSums && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < kQueues) heads[threadIdx.x] = 0u;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == kClearHQSDepth || mode == kClearHQSSums) {
        // planes are `stride` words apart, stride a multiple of 64: uint4 stores
        const size_t n = (mode == kClearHQSDepth ? 1 : 4) * (size_t)V.stride;
        uint32_t* base = V.planes + (mode == kClearHQSDepth ? 0 : (size_t)V.stride);
        const size_t j = i * 4;
        if (j < n) *reinterpret_cast<uint4*>(base + j) = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    if (mode == kClearNearest) {
        const size_t n = (size_t)V.p.width * (size_t)V.p.height;
        const size_t j = i * 2;
        if (j + 2 <= n) {
            *reinterpret_cast<ulonglong2*>(V.nearest + j) = make_ulonglong2(0ull, 0ull);
        } else if (j < n) {
            V.nearest[j] = 0ull;
        }
    } else {
        // 5 planes of `stride` words, stride a multiple of 64: uint4 stores
        const size_t n = 5 * (size_t)V.stride;
        const size_t j = i * 4;
        if (j < n) *reinterpret_cast<uint4*>(V.planes + j) = make_uint4(0u, 0u, 0u, 0u);
        // differentiable mode: 3 int64 planes, two per thread
        if (mode == kClearDiff && i * 2 < 3 * (size_t)V.stride)
            *reinterpret_cast<ulonglong2*>(V.sums + i * 2) = make_ulonglong2(0ull, 0ull);
    }
}

__global__ __launch_bounds__(256) void k_mv_cull(const MVView* __restrict__ views, MVItem* __restrict__ items) {
    const MVView& V = views[blockIdx.y];
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= V.c.nSlots) return;
    MVItem it;
    it.view = blockIdx.y;
    it.pad[0] = it.pad[1] = it.pad[2] = 0u;
    MVItem* dst = items + (V.itemBase + (uint32_t)slot * V.itemStride);
    if (cull_slot(V.c, V.p, slot, it.vb, it.bounds)) {
        *dst = it;
        return;
    }
    // culled: the header only (the draws skip the item without reading its record)
    it.vb.batchIndex = (uint32_t)slot;
    it.vb.level = 0;
    it.vb.lodThreshold = 0.0f;
    it.vb.activePoints = 0u;
    dst->view = it.view;
    dst->pad[0] = dst->pad[1] = dst->pad[2] = 0u;
    dst->vb = it.vb;
}

__global__ __launch_bounds__(256) void k_mv_resolve(const MVView* __restrict__ views) {
    const MVView& V = views[blockIdx.y];
    const size_t n = (size_t)V.p.width * (size_t)V.p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    resolve_pixel(V.p, V.planes, (size_t)V.stride, i, V.rgba, V.depthf);
}

__global__ __launch_bounds__(256) void k_mv_nearest_resolve(const MVView* __restrict__ views) {
    const MVView& V = views[blockIdx.y];
    const size_t n = (size_t)V.p.width * (size_t)V.p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    nearest_resolve_pixel(V.p, V.nearest, V.c.colors, i, V.rgba, V.depthf);
}

// ---------------------------------------------------------------- differentiable views

template <bool kR0>
__global__ __launch_bounds__(kDiffThreads) void k_mv_diff_color(const MVView* __restrict__ views,
                                                                const MVItem* __restrict__ items, uint32_t nUnits,
                                                                const float4* __restrict__ params, uint32_t* heads) {
    __shared__ DiffLds L;
    drain<kDiffThreads, kDiffPPT>(views, items, nUnits, 1, 1u, heads + 3,
                                  [&](const MVView& V, const scr_visible_batch& vb, const DrawRecord* bb, uint32_t lo,
                                      uint32_t hi) __attribute__((always_inline)) {
                                      const BatchCtx bc = batch_ctx(V.c, V.p, vb, bb, hi);
                                      const size_t s = (size_t)V.stride;
                                      if (!kR0 && bc.rconst == 0)
                                          diff_color_chunks<true>(V.c, V.p, bc, lo, params, V.planes, V.sums, s,
                                                                  V.planes + 4 * s, L);
                                      else
                                          diff_color_chunks<kR0>(V.c, V.p, bc, lo, params, V.planes, V.sums, s,
                                                                 V.planes + 4 * s, L);
                                  });
}

__global__ __launch_bounds__(256) void k_mv_diff_resolve(const MVView* __restrict__ views) {
    const MVView& V = views[blockIdx.y];
    const size_t n = (size_t)V.p.width * (size_t)V.p.height;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    diff_resolve_pixel(V.p, V.sums, (size_t)V.stride, V.planes + 4 * (size_t)V.stride, i, V.image);
}

// Backward over the views of one cloud.  A work unit is one chunk of one
// slot (perSlot chunks cover a slot's capacity); it visits that slot's item
// of every view in view order (the items of a slot are adjacent in the
// table), each thread adding its points' per-view sums to running sums in
// registers, loaded from and stored to the gradient once: the additions
// happen in the order separate per-view calls make them, so the result is
// the same bit for bit, with one gradient read and write per point.
template <bool kR0>
__global__ __launch_bounds__(kDiffThreads) void k_mv_diff_backward(const MVView* __restrict__ views,
                                                                   const MVItem* __restrict__ items, int nSlots,
                                                                   int perSlot, int nViews, double* __restrict__ grad) {
    constexpr uint32_t kChunk = kDiffThreads * kDiffPPT;
    const int nUnits = nSlots * perSlot;
    for (int u = blockIdx.x; u < nUnits; u += gridDim.x) {
        const int slot = u / perSlot, chunk = u - slot * perSlot;
        const uint32_t lo = (uint32_t)chunk * kChunk;
        const uint32_t li0 = lo + threadIdx.x * kDiffPPT;
        double acc[kDiffPPT][3];
        bool touched[kDiffPPT];
#pragma unroll
        for (int j = 0; j < kDiffPPT; j++) touched[j] = false;
        uint32_t first = 0u;
        for (int k = 0; k < nViews; k++) {
            const MVItem* ip = items + ((size_t)slot * (size_t)nViews + (size_t)k);
            const scr_visible_batch vb = ip->vb;
            if (vb.activePoints <= lo) continue;
            const MVView& V = views[__builtin_amdgcn_readfirstlane(ip->view)];
            const BatchCtx bc = batch_ctx(V.c, V.p, vb, &ip->bounds, lo + kChunk);
            first = bc.first;
            const size_t s = (size_t)V.stride;
            const bool vec = vec_ok<kDiffPPT>(bc);
            if (!kR0 && bc.rconst == 0)
                diff_backward_acc<true>(V.c, V.p, bc, li0, vec, V.planes, V.planes + 4 * s, V.dLdImage, grad, acc,
                                        touched);
            else
                diff_backward_acc<kR0>(V.c, V.p, bc, li0, vec, V.planes, V.planes + 4 * s, V.dLdImage, grad, acc,
                                       touched);
        }
        diff_backward_store(first + li0, acc, touched, grad);
    }
}

// ---------------------------------------------------------------- draw variants

// The kernel set of one (phase, shape), from its own object file.
const ShapeKernels& shape_kernels(Phase ph, int shape) {
#define SCR_MV_ROW(P) \
    {&SCR_MV_ACCESSOR(P, 0), &SCR_MV_ACCESSOR(P, 1), &SCR_MV_ACCESSOR(P, 2), &SCR_MV_ACCESSOR(P, 3), \
     &SCR_MV_ACCESSOR(P, 4), &SCR_MV_ACCESSOR(P, 5)}
    using Accessor = const ShapeKernels& (*)();
    static const Accessor table[3][kNumShapes] = {SCR_MV_ROW(0), SCR_MV_ROW(1), SCR_MV_ROW(2)};
#undef SCR_MV_ROW
    return table[(int)ph][shape]();
}

struct Variant {
    int threads, ppt;
    DrawKernel depth[2][2], color[2][2], nearest[2][2];  // [kR0][kWave]
};

// The shapes with their kernels from the phase translation units.
const Variant* variants() {
    static const Variant* table = [] {
        static Variant t[kNumShapes];
        for (int i = 0; i < kNumShapes; i++) {
            t[i].threads = kShapes[i].threads;
            t[i].ppt = kShapes[i].ppt;
            for (int r = 0; r < 2; r++)
                for (int w = 0; w < 2; w++) {
                    t[i].depth[r][w] = shape_kernels(kDepth, i)[r][w];
                    t[i].color[r][w] = shape_kernels(kColor, i)[r][w];
                    t[i].nearest[r][w] = shape_kernels(kNearest, i)[r][w];
                }
        }
        return t;
    }();
    return table;
}
constexpr int kNumVariants = kNumShapes;
const int kSplits[] = {1, 2, 4};

const Variant* find_variant(const scr_mv_shape& s) {
    for (int i = 0; i < kNumVariants; i++)
        if (variants()[i].threads == s.threads && variants()[i].ppt == s.pointsPerThread) return variants() + i;
    return nullptr;
}

bool valid_shape(const scr_mv_shape& s) {
    return find_variant(s) && (s.splits == 1 || s.splits == 2 || s.splits == 4) &&
           (s.waveTiles == 0 || s.waveTiles == 1);
}


DrawKernel kernel_of(const Variant& v, Phase ph, bool r0, bool wave) {
    const int k = r0 ? 1 : 0, w = wave ? 1 : 0;
    return ph == kDepth ? v.depth[k][w] : (ph == kColor ? v.color[k][w] : v.nearest[k][w]);
}

unsigned grid_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace
}  // namespace scr

using scr::hip_error;
using scr::set_error;

#define SCR_HIP(ctx, call)                                      \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return hip_error(ctx, e_, #call); \
    } while (0)

struct scr_multiview {
    scr_ctx* ctx = nullptr;
    std::vector<scr_cloud*> clouds;
    std::vector<scr_frame*> frames;
    std::vector<int> viewCloud;
    int nViews = 0;
    size_t maxWords = 0;     // largest 5 * planeStride over the views
    size_t maxPixels = 0;
    scr_mv_tuning tuning{};
    // device: [views][files], item table, queue heads
    unsigned char* dParams = nullptr;
    size_t paramsBytes = 0;
    scr::MVItem* dItems = nullptr;
    size_t itemsCap = 0;
    uint32_t* dHeads = nullptr;
    // pinned staging of dParams, two in flight
    unsigned char* hStage[2] = {nullptr, nullptr};
    hipEvent_t staged[2] = {nullptr, nullptr};
    int stageNext = 0;
    // per-kernel persistent grid (resident workgroups) cache
    std::vector<std::pair<scr::DrawKernel, int>> residency;
    // side stream (default priority) for the sum planes' clear, and the events
    // that fork it from and join it to the context stream
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // captured launch sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    struct Key {
        int mode = -1, r0 = -1, maxSlots = 0;
        uint32_t nItems = 0;
        scr_mv_tuning tuning{};
        bool operator==(const Key& o) const {
            return mode == o.mode && r0 == o.r0 && maxSlots == o.maxSlots && nItems == o.nItems &&
                   std::memcmp(&tuning, &o.tuning, sizeof tuning) == 0;
        }
    } graphKey;
};

namespace {

using scr::MVItem;
using scr::MVView;

void drop_graph(scr_multiview* mv) {
    if (mv->graphExec) (void)hipGraphExecDestroy(mv->graphExec);
    if (mv->graph) (void)hipGraphDestroy(mv->graph);
    mv->graphExec = nullptr;
    mv->graph = nullptr;
    mv->graphKey = scr_multiview::Key{};
}

void release(scr_multiview* mv) {
    drop_graph(mv);
    if (mv->dParams) (void)hipFree(mv->dParams);
    if (mv->dItems) (void)hipFree(mv->dItems);
    if (mv->dHeads) (void)hipFree(mv->dHeads);
    for (int k = 0; k < 2; k++) {
        if (mv->hStage[k]) (void)hipHostFree(mv->hStage[k]);
        if (mv->staged[k]) (void)hipEventDestroy(mv->staged[k]);
        mv->hStage[k] = nullptr;
        mv->staged[k] = nullptr;
    }
    if (mv->fork) (void)hipEventDestroy(mv->fork);
    if (mv->join) (void)hipEventDestroy(mv->join);
    if (mv->side) (void)hipStreamDestroy(mv->side);
    mv->fork = mv->join = nullptr;
    mv->side = nullptr;
    mv->dParams = nullptr;
    mv->dItems = nullptr;
    mv->dHeads = nullptr;
}

size_t files_total(const scr_multiview* mv) {
    size_t n = 0;
    for (int v = 0; v < mv->nViews; v++) n += mv->clouds[(size_t)mv->viewCloud[(size_t)v]]->files.size();
    return n;
}

int resident_grid(scr_multiview* mv, scr::DrawKernel k, int threads, uint32_t nUnits, unsigned* grid) {
    int perCU = -1;
    for (const auto& e : mv->residency)
        if (e.first == k) perCU = e.second;
    if (perCU < 0) {
        SCR_HIP(mv->ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(k),
                                                                      threads, 0));
        if (perCU < 1) perCU = 1;
        mv->residency.emplace_back(k, perCU);
    }
    const uint32_t g = (uint32_t)perCU * (uint32_t)mv->ctx->numCUs;
    *grid = std::max(1u, std::min(g, nUnits));
    return SCR_OK;
}

// The launch sequence of one render; eager, or recorded under stream capture.
// Events t0 / t1 (eager only) bracket the draw kernel of phase `timedPhase`.
int enqueue(scr_multiview* mv, int mode, bool r0, uint32_t nItems, int timedPhase = -1, hipEvent_t t0 = nullptr,
            hipEvent_t t1 = nullptr) {
    scr_ctx* ctx = mv->ctx;
    hipStream_t s = ctx->stream;
    const MVView* views = reinterpret_cast<const MVView*>(mv->dParams);
    const bool nearest = mode == SCR_MODE_NEAREST_POINT;
    int maxSlots = 0;
    for (scr_cloud* c : mv->clouds) maxSlots = std::max(maxSlots, c->maxResidentBatches);
    const unsigned nv = (unsigned)mv->nViews;
    // HQS: the sum planes are cleared on the side stream, overlapping the
    // cull and depth passes (joined before the colour pass); with pass
    // profiling on, in order on the context stream
    const bool overlap = !nearest && !ctx->profiling;
    bool joined = !overlap;
    if (overlap) {
        const size_t stride = mv->maxWords / 5;
        scr::PassScope ps(ctx, "mv_clear");
        hipLaunchKernelGGL(scr::k_mv_clear, dim3(std::max(1u, scr::grid_for(stride, 256 * 4u)), nv), dim3(256), 0, s,
                           views, (int)scr::kClearHQSDepth, mv->dHeads);
        SCR_HIP(ctx, hipGetLastError());
        SCR_HIP(ctx, hipEventRecord(mv->fork, s));
        SCR_HIP(ctx, hipStreamWaitEvent(mv->side, mv->fork, 0));
        hipLaunchKernelGGL(scr::k_mv_clear, dim3(std::max(1u, scr::grid_for(4 * stride, 256 * 4u)), nv), dim3(256), 0,
                           mv->side, views, (int)scr::kClearHQSSums, mv->dHeads);
        SCR_HIP(ctx, hipGetLastError());
        SCR_HIP(ctx, hipEventRecord(mv->join, mv->side));
    } else {
        scr::PassScope ps(ctx, "mv_clear");
        const size_t perThread = nearest ? 2 : 4;
        const size_t n = nearest ? mv->maxPixels : mv->maxWords;
        hipLaunchKernelGGL(scr::k_mv_clear, dim3(std::max(1u, scr::grid_for(n, 256 * (unsigned)perThread)), nv),
                           dim3(256), 0, s, views, nearest ? (int)scr::kClearNearest : (int)scr::kClearHQS,
                           mv->dHeads);
        SCR_HIP(ctx, hipGetLastError());
    }
    if (maxSlots > 0) {
        scr::PassScope ps(ctx, "mv_cull");
        hipLaunchKernelGGL(scr::k_mv_cull, dim3(scr::grid_for((size_t)maxSlots, 256), nv), dim3(256), 0, s, views,
                           mv->dItems);
        SCR_HIP(ctx, hipGetLastError());
    }
    const scr::Phase phases[2] = {nearest ? scr::kNearest : scr::kDepth, scr::kColor};
    const int nPhases = nearest ? 1 : 2;
    for (int k = 0; k < nPhases && nItems > 0; k++) {
        const scr::Phase ph = phases[k];
        const scr_mv_shape& sh = ph == scr::kDepth ? mv->tuning.depth : (ph == scr::kColor ? mv->tuning.color
                                                                                            : mv->tuning.nearest);
        const scr::Variant* var = scr::find_variant(sh);
        if (!var) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "unsupported draw shape");
        const scr::DrawKernel kern = scr::kernel_of(*var, ph, r0, sh.waveTiles != 0);
        const uint32_t nUnits = nItems * (uint32_t)sh.splits;
        unsigned grid = 1;
        int rc = resident_grid(mv, kern, var->threads, nUnits, &grid);
        if (rc != SCR_OK) return rc;
        if (ph == scr::kColor && !joined) {
            SCR_HIP(ctx, hipStreamWaitEvent(s, mv->join, 0));
            joined = true;
        }
        if (timedPhase == (int)ph) SCR_HIP(ctx, hipEventRecord(t0, s));
        {
            scr::PassScope ps(ctx, ph == scr::kDepth ? "mv_depth" : (ph == scr::kColor ? "mv_color" : "mv_nearest"));
            hipLaunchKernelGGL(kern, dim3(grid), dim3((unsigned)var->threads), 0, s, views, mv->dItems, nUnits,
                               sh.splits, scr::mv::claim_for(nUnits, grid), mv->dHeads);
            SCR_HIP(ctx, hipGetLastError());
        }
        if (timedPhase == (int)ph) SCR_HIP(ctx, hipEventRecord(t1, s));
    }
    if (!joined) SCR_HIP(ctx, hipStreamWaitEvent(s, mv->join, 0));
    {
        scr::PassScope ps(ctx, nearest ? "mv_nearest_resolve" : "mv_resolve");
        if (nearest)
            hipLaunchKernelGGL(scr::k_mv_nearest_resolve, dim3(std::max(1u, scr::grid_for(mv->maxPixels, 256)), nv),
                               dim3(256), 0, s, views);
        else
            hipLaunchKernelGGL(scr::k_mv_resolve, dim3(std::max(1u, scr::grid_for(mv->maxPixels, 256)), nv),
                               dim3(256), 0, s, views);
        SCR_HIP(ctx, hipGetLastError());
    }
    return SCR_OK;
}

// Composes the per-view parameter table on the host and queues its upload;
// returns the item count.
int stage_params(scr_multiview* mv, const scr_config* cfg, const scr_camera* cams, uint32_t* nItemsOut, bool* r0,
                 float* const* images = nullptr, const float* const* dLdImages = nullptr) {
    scr_ctx* ctx = mv->ctx;
    for (scr_cloud* c : mv->clouds) {
        int rc = scr::flush_mirror(c);
        if (rc != SCR_OK) return rc;
    }
    // group the views of each cloud: item (v, s) = base(cloud) + s * nv(cloud) + rank(v)
    const size_t nc = mv->clouds.size();
    std::vector<uint32_t> nvOf(nc, 0), base(nc, 0);
    for (int v = 0; v < mv->nViews; v++) nvOf[(size_t)mv->viewCloud[(size_t)v]]++;
    uint64_t total = 0;
    for (size_t c = 0; c < nc; c++) {
        base[c] = (uint32_t)total;
        total += (uint64_t)nvOf[c] * (uint64_t)mv->clouds[c]->maxResidentBatches;
    }
    if (total > 0x0fffffffull) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "too many (view, slot) pairs");
    const size_t nItems = (size_t)total;
    if (nItems > mv->itemsCap) {
        SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (mv->dItems) SCR_HIP(ctx, hipFree(mv->dItems));
        mv->dItems = nullptr;
        mv->itemsCap = 0;
        drop_graph(mv);
        SCR_HIP(ctx, hipMalloc((void**)&mv->dItems, nItems * sizeof(MVItem)));
        mv->itemsCap = nItems;
    }
    const int k = mv->stageNext;
    mv->stageNext ^= 1;
    SCR_HIP(ctx, hipEventSynchronize(mv->staged[k]));
    unsigned char* stage = mv->hStage[k];
    MVView* views = reinterpret_cast<MVView*>(stage);
    scr_raster_file* files = reinterpret_cast<scr_raster_file*>(stage + sizeof(MVView) * (size_t)mv->nViews);
    const scr_raster_file* dFiles =
        reinterpret_cast<const scr_raster_file*>(mv->dParams + sizeof(MVView) * (size_t)mv->nViews);
    std::vector<uint32_t> rank(nc, 0);
    size_t fileOff = 0;
    for (int v = 0; v < mv->nViews; v++) {
        const size_t ci = (size_t)mv->viewCloud[(size_t)v];
        scr_cloud* c = mv->clouds[ci];
        scr_frame* f = mv->frames[(size_t)v];
        MVView& V = views[v];
        std::memset(&V, 0, sizeof V);
        V.p = scr::make_params(cfg, &cams[v], f->width, f->height);
        float vp[16];
        scr::mat_mul(cams[v].projectionMatrix, cams[v].viewMatrix, vp);
        scr::compose_files(c, c->model, vp, nullptr, nullptr, files + fileOff);
        V.c = c->dev();
        V.c.files = dFiles + fileOff;
        V.c.visible = nullptr;
        V.c.bounds = nullptr;
        V.c.counter = nullptr;
        fileOff += c->files.size();
        V.planes = f->dPlanes;
        V.nearest = f->dNearest;
        V.rgba = f->dRGBA;
        V.depthf = f->dDepthF;
        V.stride = f->planeStride();
        V.itemBase = base[ci] + rank[ci]++;
        V.itemStride = nvOf[ci];
        V.sums = f->dSums;
        V.image = images ? images[v] : nullptr;
        V.dLdImage = dLdImages ? dLdImages[v] : nullptr;
    }
    SCR_HIP(ctx, hipMemcpyAsync(mv->dParams, stage, mv->paramsBytes, hipMemcpyHostToDevice, ctx->stream));
    SCR_HIP(ctx, hipEventRecord(mv->staged[k], ctx->stream));
    *nItemsOut = (uint32_t)nItems;
    *r0 = views[0].p.constRadius == 0;
    return SCR_OK;
}

int check_tuning(scr_ctx* ctx, const scr_mv_tuning& t) {
    if (!scr::valid_shape(t.depth) || !scr::valid_shape(t.color) || !scr::valid_shape(t.nearest))
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT,
                         "draw shape must be threads/pointsPerThread in {128/8, 256/4, 256/8, 256/16, 512/4, 512/8}, "
                         "splits in {1, 2, 4} and waveTiles in {0, 1}");
    return SCR_OK;
}

int render_impl(scr_multiview* mv, const scr_config* cfg, const scr_camera* cams, bool useGraph) {
    scr_ctx* ctx = mv->ctx;
    int rc = scr::check_config(ctx, cfg);
    if (rc != SCR_OK) return rc;
    uint32_t nItems = 0;
    bool r0 = false;
    rc = stage_params(mv, cfg, cams, &nItems, &r0);
    if (rc != SCR_OK) return rc;
    if (!useGraph || ctx->profiling) return enqueue(mv, cfg->mode, r0, nItems);
    scr_multiview::Key key;
    key.mode = cfg->mode;
    key.r0 = r0 ? 1 : 0;
    key.nItems = nItems;
    for (scr_cloud* c : mv->clouds) key.maxSlots = std::max(key.maxSlots, c->maxResidentBatches);
    key.tuning = mv->tuning;
    if (!mv->graphExec || !(key == mv->graphKey)) {
        drop_graph(mv);
        SCR_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue(mv, cfg->mode, r0, nItems);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc != SCR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_error(ctx, e, "hipStreamEndCapture");
        mv->graph = g;
        SCR_HIP(ctx, hipGraphInstantiate(&mv->graphExec, g, nullptr, nullptr, 0));
        mv->graphKey = key;
    }
    SCR_HIP(ctx, hipGraphLaunch(mv->graphExec, ctx->stream));
    return SCR_OK;
}

// Differentiable views: every view must render the multi-view's one cloud.
int check_diff_views(scr_multiview* mv, const scr_config* cfg) {
    scr_ctx* ctx = mv->ctx;
    int rc = scr::check_config(ctx, cfg);
    if (rc != SCR_OK) return rc;
    if (cfg->mode != SCR_MODE_HIGH_QUALITY_AVERAGE || cfg->applyTint != 0 || cfg->colorizeChunks != 0 ||
        cfg->colorizeOverdraw != 0)
        return set_error(ctx, SCR_ERR_UNSUPPORTED, "differentiable rendering is HQS with colours from the parameters");
    if (mv->clouds.size() != 1)
        return set_error(ctx, SCR_ERR_UNSUPPORTED, "differentiable views render one cloud");
    for (scr_frame* f : mv->frames)
        if (!f->dSums) {
            SCR_HIP(ctx, hipStreamSynchronize(ctx->stream));
            SCR_HIP(ctx, hipMalloc((void**)&f->dSums, 3 * f->planeStride() * sizeof(unsigned long long)));
        }
    return SCR_OK;
}

unsigned diff_grid(scr_multiview* mv, scr::DrawKernel k, uint32_t nUnits) {
    unsigned grid = 1;
    if (resident_grid(mv, k, scr::kDiffThreads, nUnits, &grid) != SCR_OK) grid = (unsigned)mv->ctx->numCUs * 4u;
    return std::max(1u, grid);
}

}  // namespace

extern "C" {

int scr_multiview_create(scr_ctx* ctx, scr_cloud* const* clouds, int nClouds, scr_frame* const* frames,
                         const int32_t* viewCloud, int nViews, scr_multiview** out) {
    if (!ctx || !out) return SCR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (nClouds <= 0 || !clouds || nViews <= 0 || !frames || !viewCloud)
        return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "need at least one cloud and one view");
    if (nViews > 65535) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "at most 65535 views");
    for (int i = 0; i < nClouds; i++)
        if (!clouds[i] || clouds[i]->ctx != ctx) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "bad cloud handle");
    for (int v = 0; v < nViews; v++) {
        if (!frames[v] || frames[v]->ctx != ctx) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "bad frame handle");
        if (viewCloud[v] < 0 || viewCloud[v] >= nClouds)
            return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "view cloud index out of range");
        for (int u = 0; u < v; u++)
            if (frames[u] == frames[v])
                return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "views must render into distinct frames");
    }
    scr_multiview* mv = new (std::nothrow) scr_multiview();
    if (!mv) return SCR_ERR_OUT_OF_MEMORY;
    mv->ctx = ctx;
    mv->clouds.assign(clouds, clouds + nClouds);
    mv->frames.assign(frames, frames + nViews);
    mv->viewCloud.assign(viewCloud, viewCloud + nViews);
    mv->nViews = nViews;
    for (scr_frame* f : mv->frames) {
        mv->maxWords = std::max(mv->maxWords, 5 * f->planeStride());
        mv->maxPixels = std::max(mv->maxPixels, f->pixels());
    }
    const scr_mv_shape def{256, 8, 1, 0};
    mv->tuning.depth = mv->tuning.color = mv->tuning.nearest = def;
    mv->paramsBytes = sizeof(MVView) * (size_t)nViews + sizeof(scr_raster_file) * files_total(mv);
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc((void**)&mv->dParams, mv->paramsBytes);
    if (e == hipSuccess) e = hipMalloc((void**)&mv->dHeads, scr::kQueues * sizeof(uint32_t));
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        e = hipHostMalloc((void**)&mv->hStage[k], mv->paramsBytes);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&mv->staged[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&mv->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&mv->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&mv->join, hipEventDisableTiming);
    if (e != hipSuccess) {
        release(mv);
        delete mv;
        return hip_error(ctx, e, "scr_multiview_create");
    }
    scr::ctx_retain(ctx);
    *out = mv;
    return SCR_OK;
}

int scr_multiview_destroy(scr_multiview* mv) {
    if (!mv) return SCR_OK;
    scr_ctx* ctx = mv->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    release(mv);
    delete mv;
    scr::ctx_release(ctx);
    return SCR_OK;
}

int scr_multiview_render(scr_multiview* mv, const scr_config* cfg, const scr_camera* cameras, int useGraph) {
    if (!mv || !cameras) return SCR_ERR_INVALID_ARGUMENT;
    return render_impl(mv, cfg, cameras, useGraph != 0);
}

int scr_multiview_get_tuning(const scr_multiview* mv, scr_mv_tuning* out) {
    if (!mv || !out) return SCR_ERR_INVALID_ARGUMENT;
    *out = mv->tuning;
    return SCR_OK;
}

int scr_multiview_set_tuning(scr_multiview* mv, const scr_mv_tuning* t) {
    if (!mv || !t) return SCR_ERR_INVALID_ARGUMENT;
    int rc = check_tuning(mv->ctx, *t);
    if (rc != SCR_OK) return rc;
    mv->tuning = *t;
    return SCR_OK;
}

int scr_multiview_autotune(scr_multiview* mv, const scr_config* cfg, const scr_camera* cameras, int reps,
                           scr_mv_tuning* out) {
    if (!mv || !cameras) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = mv->ctx;
    int rc = scr::check_config(ctx, cfg);
    if (rc != SCR_OK) return rc;
    if (reps < 1) reps = 1;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    SCR_HIP(ctx, hipEventCreate(&t0));
    SCR_HIP(ctx, hipEventCreate(&t1));
    const bool nearest = cfg->mode == SCR_MODE_NEAREST_POINT;
    const scr::Phase phases[2] = {nearest ? scr::kNearest : scr::kDepth, scr::kColor};
    const int nPhases = nearest ? 1 : 2;
    const bool wasProfiling = ctx->profiling;
    ctx->profiling = false;
    scr_mv_tuning best = mv->tuning;
    for (int k = 0; k < nPhases && rc == SCR_OK; k++) {
        const scr::Phase ph = phases[k];
        scr_mv_shape* slot = ph == scr::kDepth ? &mv->tuning.depth
                                               : (ph == scr::kColor ? &mv->tuning.color : &mv->tuning.nearest);
        scr_mv_shape bestShape = *slot;
        float bestMs = 1e30f;
        for (int vi = 0; vi < scr::kNumVariants && rc == SCR_OK; vi++)
            for (int wave = 0; wave < 2; wave++)
            for (int split : scr::kSplits) {
                *slot = scr_mv_shape{scr::variants()[vi].threads, scr::variants()[vi].ppt, split, wave};
                float ms = 1e30f;
                for (int r = 0; r < reps && rc == SCR_OK; r++) {
                    uint32_t nItems = 0;
                    bool r0 = false;
                    rc = stage_params(mv, cfg, cameras, &nItems, &r0);
                    if (rc == SCR_OK) rc = enqueue(mv, cfg->mode, r0, nItems, (int)ph, t0, t1);
                    if (rc != SCR_OK) break;
                    hipError_t e = hipEventSynchronize(t1);
                    float m = 0.0f;
                    if (e == hipSuccess) e = hipEventElapsedTime(&m, t0, t1);
                    if (e != hipSuccess) {
                        rc = hip_error(ctx, e, "autotune timing");
                        break;
                    }
                    ms = std::min(ms, m);
                }
                if (rc == SCR_OK && ms < bestMs) {
                    bestMs = ms;
                    bestShape = *slot;
                }
            }
        *slot = bestShape;
        best = mv->tuning;
    }
    ctx->profiling = wasProfiling;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc != SCR_OK) return rc;
    mv->tuning = best;
    drop_graph(mv);
    if (out) *out = best;
    return SCR_OK;
}

int scr_multiview_diff_render(scr_multiview* mv, const scr_config* cfg, const scr_camera* cameras,
                              const float* params, float* const* images) {
    if (!mv || !cameras || !params || !images) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = mv->ctx;
    for (int v = 0; v < mv->nViews; v++)
        if (!images[v]) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "NULL image pointer");
    int rc = check_diff_views(mv, cfg);
    if (rc != SCR_OK) return rc;
    uint32_t nItems = 0;
    bool r0 = false;
    rc = stage_params(mv, cfg, cameras, &nItems, &r0, images, nullptr);
    if (rc != SCR_OK) return rc;
    hipStream_t s = ctx->stream;
    const MVView* views = reinterpret_cast<const MVView*>(mv->dParams);
    const unsigned nv = (unsigned)mv->nViews;
    const int maxSlots = mv->clouds[0]->maxResidentBatches;
    {
        scr::PassScope ps(ctx, "mv_clear");
        // threads: 4 plane words or 2 sums each; the 3 sum planes need the most
        const size_t maxStride = mv->maxWords / 5;
        const size_t threads = std::max(mv->maxWords / 4, 3 * maxStride / 2);
        hipLaunchKernelGGL(scr::k_mv_clear, dim3(std::max(1u, scr::grid_for(threads, 256)), nv), dim3(256), 0, s,
                           views, (int)scr::kClearDiff, mv->dHeads);
        SCR_HIP(ctx, hipGetLastError());
    }
    if (nItems == 0 || maxSlots == 0) return SCR_OK;
    {
        scr::PassScope ps(ctx, "mv_cull");
        hipLaunchKernelGGL(scr::k_mv_cull, dim3(scr::grid_for((size_t)maxSlots, 256), nv), dim3(256), 0, s, views,
                           mv->dItems);
        SCR_HIP(ctx, hipGetLastError());
    }
    {
        const scr_mv_shape& sh = mv->tuning.depth;
        const scr::Variant* var = scr::find_variant(sh);
        const scr::DrawKernel kern = scr::kernel_of(*var, scr::kDepth, r0, sh.waveTiles != 0);
        const uint32_t nUnits = nItems * (uint32_t)sh.splits;
        unsigned grid = 1;
        rc = resident_grid(mv, kern, var->threads, nUnits, &grid);
        if (rc != SCR_OK) return rc;
        scr::PassScope ps(ctx, "mv_depth");
        hipLaunchKernelGGL(kern, dim3(grid), dim3((unsigned)var->threads), 0, s, views, mv->dItems, nUnits, sh.splits,
                           scr::mv::claim_for(nUnits, grid), mv->dHeads);
        SCR_HIP(ctx, hipGetLastError());
    }
    {
        const scr::DrawKernel kern = reinterpret_cast<scr::DrawKernel>(
            r0 ? (void*)scr::k_mv_diff_color<true> : (void*)scr::k_mv_diff_color<false>);
        const unsigned grid = diff_grid(mv, kern, nItems);
        scr::PassScope ps(ctx, "mv_diff_color");
        if (r0)
            hipLaunchKernelGGL(scr::k_mv_diff_color<true>, dim3(grid), dim3(scr::kDiffThreads), 0, s, views,
                               mv->dItems, nItems, reinterpret_cast<const float4*>(params), mv->dHeads);
        else
            hipLaunchKernelGGL(scr::k_mv_diff_color<false>, dim3(grid), dim3(scr::kDiffThreads), 0, s, views,
                               mv->dItems, nItems, reinterpret_cast<const float4*>(params), mv->dHeads);
        SCR_HIP(ctx, hipGetLastError());
    }
    scr::PassScope ps(ctx, "mv_diff_resolve");
    hipLaunchKernelGGL(scr::k_mv_diff_resolve, dim3(std::max(1u, scr::grid_for(mv->maxPixels, 256)), nv), dim3(256),
                       0, s, views);
    SCR_HIP(ctx, hipGetLastError());
    return SCR_OK;
}

int scr_multiview_diff_backward(scr_multiview* mv, const scr_config* cfg, const scr_camera* cameras,
                                const float* const* dLdImages, double* grad) {
    if (!mv || !cameras || !dLdImages || !grad) return SCR_ERR_INVALID_ARGUMENT;
    scr_ctx* ctx = mv->ctx;
    for (int v = 0; v < mv->nViews; v++)
        if (!dLdImages[v]) return set_error(ctx, SCR_ERR_INVALID_ARGUMENT, "NULL dLdImage pointer");
    int rc = check_diff_views(mv, cfg);
    if (rc != SCR_OK) return rc;
    uint32_t nItems = 0;
    bool r0 = false;
    rc = stage_params(mv, cfg, cameras, &nItems, &r0, nullptr, dLdImages);
    if (rc != SCR_OK) return rc;
    scr_cloud* cloud = mv->clouds[0];
    const int maxSlots = cloud->maxResidentBatches;
    if (nItems == 0 || maxSlots == 0 || cloud->residentBatchCount == 0) return SCR_OK;
    hipStream_t s = ctx->stream;
    const MVView* views = reinterpret_cast<const MVView*>(mv->dParams);
    {
        scr::PassScope ps(ctx, "mv_cull");
        hipLaunchKernelGGL(scr::k_mv_cull, dim3(scr::grid_for((size_t)maxSlots, 256), (unsigned)mv->nViews),
                           dim3(256), 0, s, views, mv->dItems);
        SCR_HIP(ctx, hipGetLastError());
    }
    constexpr int kChunk = scr::kDiffThreads * scr::kDiffPPT;
    const int perSlot = (cloud->pointsPerBatch + kChunk - 1) / kChunk;
    const uint32_t nUnits = (uint32_t)maxSlots * (uint32_t)perSlot;
    const scr::DrawKernel kern = reinterpret_cast<scr::DrawKernel>(
        r0 ? (void*)scr::k_mv_diff_backward<true> : (void*)scr::k_mv_diff_backward<false>);
    const unsigned grid = diff_grid(mv, kern, nUnits);
    scr::PassScope ps(ctx, "mv_diff_backward");
    if (r0)
        hipLaunchKernelGGL(scr::k_mv_diff_backward<true>, dim3(grid), dim3(scr::kDiffThreads), 0, s, views,
                           mv->dItems, maxSlots, perSlot, mv->nViews, grad);
    else
        hipLaunchKernelGGL(scr::k_mv_diff_backward<false>, dim3(grid), dim3(scr::kDiffThreads), 0, s, views,
                           mv->dItems, maxSlots, perSlot, mv->nViews, grad);
    SCR_HIP(ctx, hipGetLastError());
    return SCR_OK;
}

}  // extern "C"
