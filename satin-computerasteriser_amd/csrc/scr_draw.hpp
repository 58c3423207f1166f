// scr_draw.hpp — device code shared by the single-view (scr_render.hip) and
// multi-view (scr_mv.hpp, scr_mv_*.hip) draw kernels: the cull decision, the point
// front end, and the per-batch bodies of the depth, colour and nearest passes.
//
// A visible batch is drawn in chunks of NT * PPT points: every
// thread owns PPT consecutive points (vector loads of the position words) and
// keeps their fragments in registers.  When the batch's projected box fits an
// LDS tile, all of its chunks meet in that one tile (LDS atomics), flushed
// once with one global update per touched pixel.  Otherwise the group reduces
// each chunk's actual screen rectangle and, when that is small next to the
// fragment count, uses a tile per chunk; the rest go straight to the global
// planes.  Integer max / add are order independent, so every route gives the
// same planes.
#pragma once

#include <type_traits>

#include "scr_internal.hpp"

namespace scr {
namespace draw {

// Draw workgroups hold NT threads (a multiple of 64) that own PPT points each.
constexpr int kMaxWaves = 8;        // NT <= 512
constexpr int kDepthTile = 9984;    // u32 per pixel: 39 KB; with the spare and reduction words four workgroups per CU
constexpr int kColorTile = 2528;    // threshold u32 + (r | g << 24 | count << 48) u64 + b u32: 39.5 KB, four per CU
constexpr int kBlendTile = 2016;    // two u64 sums per pixel (weighted-blend colour): 31.5 KB
constexpr int kNearestTile = 4992;  // u64 per pixel: 39 KB, likewise four per CU
// The multi-view draw kernels keep smaller tiles: their 128-thread shapes
// stay at 4 waves per SIMD within a CU's LDS (measured faster there).
constexpr int kMVDepthTile = 4096;
constexpr int kMVNearestTile = 2048;
constexpr int kSpare = 64;          // words past a tile: one per lane, the target of no-op updates

// LDS tiles are read and updated through pointers typed as LDS: code that
// picks between an LDS and a global pointer would otherwise compile to flat
// instructions, which are slower than the ds_* forms.
#define SCR_LDS __attribute__((address_space(3)))
template <typename T>
__device__ __forceinline__ SCR_LDS T* lds_ptr(T* p) { return (SCR_LDS T*)p; }
template <typename T>
__device__ __forceinline__ T lds_load(const T* p) { return *lds_ptr(const_cast<T*>(p)); }
template <typename T>
__device__ __forceinline__ void lds_store(T* p, T v) { *lds_ptr(p) = v; }
template <typename T>
__device__ __forceinline__ void lds_max(T* p, T v) {
    __hip_atomic_fetch_max(lds_ptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void lds_add(T* p, T v) {
    __hip_atomic_fetch_add(lds_ptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The same for device memory: pointers that reach the draw bodies through
// memory (the multi-view tables) are not known to be global, and would
// otherwise be read and updated with flat instructions.
#define SCR_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ SCR_GLOBAL T* gptr(T* p) { return (SCR_GLOBAL T*)p; }
template <typename T>
__device__ __forceinline__ const SCR_GLOBAL T* gptr(const T* p) { return (const SCR_GLOBAL T*)p; }
template <typename T>
__device__ __forceinline__ T gload(const T* p) { return *gptr(p); }
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// A 16-byte-aligned record of whole 16-byte words (vector types and structs
// have no constructor from a global reference, so they go through words).
template <typename T>
__device__ __forceinline__ T gload_rec(const T* p) {
    static_assert(sizeof(T) % 16 == 0, "record of 16-byte words");
    u32x4 w[sizeof(T) / 16];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 16); i++) w[i] = ((const SCR_GLOBAL u32x4*)p)[i];
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}
template <typename T>
__device__ __forceinline__ void gmax(T* p, T v) {
    __hip_atomic_fetch_max(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void gadd(T* p, T v) {
    __hip_atomic_fetch_add(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waves per SIMD the depth and nearest kernels are compiled for: with up to
// 8 points per thread, 4 (at most 128 VGPRs) hide memory latency better than
// the 3 an unconstrained allocation leaves (measured); 2 for 16 points.
constexpr int draw_waves(int ppt) { return ppt <= 8 ? 4 : 2; }

// ---------------------------------------------------------------- cull

// makePlane + intersectsFrustum — Common.metal:119-153
__device__ __forceinline__ bool intersects_frustum(const float* m, const scr_raster_batch& b) {
#define M(c, r) m[(c) * 4 + (r)]
    const float pl[6][4] = {
        {M(0, 3) - M(0, 0), M(1, 3) - M(1, 0), M(2, 3) - M(2, 0), M(3, 3) - M(3, 0)},
        {M(0, 3) + M(0, 0), M(1, 3) + M(1, 0), M(2, 3) + M(2, 0), M(3, 3) + M(3, 0)},
        {M(0, 3) + M(0, 1), M(1, 3) + M(1, 1), M(2, 3) + M(2, 1), M(3, 3) + M(3, 1)},
        {M(0, 3) - M(0, 1), M(1, 3) - M(1, 1), M(2, 3) - M(2, 1), M(3, 3) - M(3, 1)},
        {M(0, 3) - M(0, 2), M(1, 3) - M(1, 2), M(2, 3) - M(2, 2), M(3, 3) - M(3, 2)},
        {M(0, 2), M(1, 2), M(2, 2), M(3, 2)},
    };
#undef M
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const float len = fmaxf(len3(pl[i][0], pl[i][1], pl[i][2]), 0.000001f);
        const float r = rcp(len);
        const float nx = pl[i][0] * r, ny = pl[i][1] * r, nz = pl[i][2] * r;
        const float off = pl[i][3] * r;
        const float px = nx > 0.0f ? b.maxX : b.minX;
        const float py = ny > 0.0f ? b.maxY : b.minY;
        const float pz = nz > 0.0f ? b.maxZ : b.minZ;
        const float d = (nx * px + ny * py) + nz * pz;
        if (d + off < 0.0f) return false;
    }
    return true;
}

// pixelSizeOnScreen — Common.metal:155-172
__device__ __forceinline__ float pixel_size_on_screen(const scr_raster_batch& b, const float* world, const PassParams& p) {
    const float cx = (b.minX + b.maxX) * 0.5f;
    const float cy = (b.minY + b.maxY) * 0.5f;
    const float cz = (b.minZ + b.maxZ) * 0.5f;
    const float radius = len3(b.minX - b.maxX, b.minY - b.maxY, b.minZ - b.maxZ);
    float vw[16];
    mat_mul(p.view, world, vw);
    const f4 vc = mul(vw, f4{cx, cy, cz, 1.0f});
    const f4 ve{vc.x + radius, vc.y + 0.0f, vc.z + 0.0f, vc.w + 0.0f};
    f4 pc = mul(p.proj, vc);
    f4 pe = mul(p.proj, ve);
    const float rc = rcp(fmaxf(fabsf(pc.w), 0.000001f));
    const float re = rcp(fmaxf(fabsf(pe.w), 0.000001f));
    pc.x = pc.x * rc;
    pc.y = pc.y * rc;
    pe.x = pe.x * re;
    pe.y = pe.y * re;
    const float sw = (float)p.width, sh = (float)p.height;
    const float scx = (sw * (pc.x + 1.0f)) * 0.5f, scy = (sh * (pc.y + 1.0f)) * 0.5f;
    const float sex = (sw * (pe.x + 1.0f)) * 0.5f, sey = (sh * (pe.y + 1.0f)) * 0.5f;
    const float dx = sex - scx, dy = sey - scy;
    return sqrtf(dx * dx + dy * dy);
}

// The decode step per axis of a batch at a drawn precision level
// (Common.metal:73-117): max(size, 1e-6) * 2^-bits, 30 bits at level 0, 20 at
// level 1, 10 below.
__device__ __forceinline__ void decode_steps(const scr_raster_batch& b, int level, float step[3]) {
    const float inv = level == 0 ? (1.0f / 1073741824.0f) : (level == 1 ? (1.0f / 1048576.0f) : (1.0f / 1024.0f));
    step[0] = fmaxf(b.maxX - b.minX, 0.000001f) * inv;
    step[1] = fmaxf(b.maxY - b.minY, 0.000001f) * inv;
    step[2] = fmaxf(b.maxZ - b.minZ, 0.000001f) * inv;
}

// The decision of cullUpdate (Cull/Shaders.metal:19-83) for slot `gid`;
// true when the batch is drawn, with its VisibleBatch record in `vb` and its
// draw record (what its box proves about its points, batch_bounds, and its
// draw context) in `dr`.
// The bounds part of the decision, which depends on nothing else of it: what
// the box of slot gid's batch proves under its file transform (for a resident
// batch with a valid file; undisplaced draws only read it).
__device__ __forceinline__ BatchBounds cull_bounds(const CloudDev& c, const PassParams& p, int gid) {
    const scr_raster_batch b = c.batches[gid];
    if (b.state == 0 || (int)b.fileIndex >= c.nFiles || p.applyDisplacement != 0) return BatchBounds{};
    const scr_raster_file* f = c.files + b.fileIndex;
    return batch_bounds(f->transform, b, p.width, p.height, batch_const_radius(p, f->world, b));
}

// given: the slot's cull_bounds, computed elsewhere (by another wave of the
// workgroup), or nullptr to compute them here.
__device__ __forceinline__ bool cull_slot(const CloudDev& c, const PassParams& p, int gid, scr_visible_batch& vb,
                                          DrawRecord& dr, const BatchBounds* given = nullptr) {
    const scr_raster_batch b = c.batches[gid];
    if (b.state == 0) return false;  // residency gate, Cull/Shaders.metal:34-36
    if ((int)b.fileIndex >= c.nFiles) return false;
    const scr_raster_file* f = c.files + b.fileIndex;
    if (p.enableFrustumCulling != 0 && !intersects_frustum(f->transformFrustum, b)) return false;
    const float px = pixel_size_on_screen(b, f->world, p);
    int level;
    if (px < 100.0f) level = 4;
    else if (px < 200.0f) level = 3;
    else if (px < 500.0f) level = 2;
    else if (px < 10000.0f) level = 1;
    else level = 0;
    const float thr = (p.enableCLOD != 0)
                          ? fmaxf(0.0f, (scr_log2f(fmaxf(px, 1.0f) * (1.0f / 25.0f)) + 3.0f) + (float)p.lodBias)
                          : 99.0f;
    uint32_t active = b.numPoints;
    if (b.lodCum67 != 0u) {
        int lmax = f2i(floorf(thr + 0.5f));
        lmax = lmax < 0 ? 0 : (lmax > 7 ? 7 : lmax);
        const uint32_t w = lmax < 2 ? b.lodCum01 : (lmax < 4 ? b.lodCum23 : (lmax < 6 ? b.lodCum45 : b.lodCum67));
        active = (lmax & 1) ? (w >> 16) : (w & 0xffffu);
    }
    if (active == 0u) return false;
    vb.batchIndex = (uint32_t)gid;
    vb.level = level;
    vb.lodThreshold = thr;
    vb.activePoints = active;
    if (p.applyDisplacement != 0) {
        // displaced points leave the box: nothing about them follows from it
        dr.bb = BatchBounds{};
        dr.bb.rconst = p.constRadius;
        dr.bb.zmax = __builtin_inff();
    } else if (given) {
        dr.bb = *given;
    } else {
        dr.bb = batch_bounds(f->transform, b, p.width, p.height, batch_const_radius(p, f->world, b));
    }
    dr.first = b.firstPoint;
    dr.fileIndex = b.fileIndex;
    dr.mn[0] = b.minX;
    dr.mn[1] = b.minY;
    dr.mn[2] = b.minZ;
    decode_steps(b, level, dr.step);
    dr.aff = batch_affine(f->transform, dr.step[0], dr.step[1], dr.step[2], b.minX, b.minY, b.minZ);
    return true;
}

// ---------------------------------------------------------------- batch context

struct BatchCtx {
    float mnx, mny, mnz;
    float sx, sy, sz;   // batch size * 2^-bits
    int level;
    float lodThreshold;
    bool keepAll;       // no point of the batch can fail the LOD/dither test
    uint32_t hashFrom;  // the lowest level whose points take the dither test (8: none)
    int rconst;         // footprint radius shared by the batch's points, or -1
    bool fastRcp;       // every point's clip w lies in [2^-125, 2^125] (rcp_fast is exact)
    bool proven;        // every fragment centre lies in [bx0, bx1] x [by0, by1] (column, unflipped row), z <= zmax
    bool inside;        // no point fails the clip-w, NDC or viewport tests
    int bx0, bx1, by0, by1;
    float zmax;
    uint32_t first, active, batchIndex;   // active: end of the drawn range
    BatchAffine aff;    // file.transform folded with the batch decode (project_q)
    float T[16];        // file.transform, displaced points only (held in
                        // registers: the atomics below would otherwise force
                        // reloads through memory)
    const float* W;     // file.world (only read for per-point sizes)
};

// `hi` bounds the local point indices drawn (at most the active prefix).
// `drp` points at the batch's draw record from the cull pass: the context
// comes from it and the visible record in independent loads (the batch
// record and its file are not read; the file transform only for displaced
// points).
__device__ __forceinline__ BatchCtx batch_ctx(const CloudDev& c, const PassParams& p, const scr_visible_batch& vb,
                                              const DrawRecord* drp, uint32_t hi) {
    const DrawRecord dr = gload_rec(drp);
    BatchCtx bc;
    bc.level = vb.level;
    bc.lodThreshold = vb.lodThreshold;
    // dither is in [0, 1] and t(L) = (thr - L) + 0.5 does not increase with
    // L (each rounding is monotone): points of levels L with t(L) > 1 are kept
    // without a dither value, and those are the levels below hashFrom
    bc.keepAll = ((vb.lodThreshold - 7.0f) + 0.5f) > 1.0f;
    bc.hashFrom = 8u;
#pragma unroll
    for (int l = 7; l >= 0; l--)
        if (!(((vb.lodThreshold - (float)l) + 0.5f) > 1.0f)) bc.hashFrom = (uint32_t)l;
    bc.active = min(vb.activePoints, hi);
    bc.batchIndex = vb.batchIndex;
    bc.first = dr.first;
    bc.mnx = dr.mn[0];
    bc.mny = dr.mn[1];
    bc.mnz = dr.mn[2];
    bc.sx = dr.step[0];
    bc.sy = dr.step[1];
    bc.sz = dr.step[2];
    bc.aff = dr.aff;
    const scr_raster_file* f = c.files + dr.fileIndex;
    if (p.applyDisplacement != 0) {
#pragma unroll
        for (int i = 0; i < 16; i++) bc.T[i] = gload(f->transform + i);
    }
    bc.W = f->world;
    const BatchBounds& bb = dr.bb;
    bc.rconst = bb.rconst;
    bc.fastRcp = (bb.flags & kBoundsFastRcp) != 0u;
    bc.proven = (bb.flags & kBoundsProven) != 0u;
    bc.inside = (bb.flags & kBoundsInside) != 0u;
    bc.bx0 = bb.x0;
    bc.bx1 = bb.x1;
    bc.by0 = bb.y0;
    bc.by1 = bb.y1;
    bc.zmax = bb.zmax;
    return bc;
}

template <int PPT>
__device__ __forceinline__ bool vec_ok(const BatchCtx& bc) {
    return (bc.first & (uint32_t)(PPT - 1)) == 0u;
}

// ---------------------------------------------------------------- point front end

template <int PPT>
struct ChunkFrags {
    bool ok[PPT];     // the point has a fragment (px, py, z, r are defined only then)
    int px[PPT];
    int py[PPT];      // y-flipped row
    int r[PPT];       // footprint radius
    float z[PPT];     // ndc z
    float wx[PPT], wy[PPT], wz[PPT];  // decoded position (only the motion-blur sweep reads it)
};

template <int PPT>
__device__ __forceinline__ void load_words(const uint32_t* src, uint32_t base, uint32_t* out) {
    const SCR_GLOBAL u32x4* p = (const SCR_GLOBAL u32x4*)(src + base);
#pragma unroll
    for (int q = 0; q < PPT / 4; q++) {
        const u32x4 v = p[q];
        out[4 * q + 0] = v.x;
        out[4 * q + 1] = v.y;
        out[4 * q + 2] = v.z;
        out[4 * q + 3] = v.w;
    }
}

// The raw words of a thread's PPT points: positions at the batch precision,
// levels packed four to a word.
template <int PPT>
struct ChunkRaw {
    uint32_t lo[PPT], me[PPT], hi[PPT], lv4[(PPT + 3) / 4];
};

// Issues the loads of the thread's PPT points starting at local index li0.
template <int kLevel, int PPT>
__device__ __forceinline__ void chunk_fetch(const CloudDev& c, const BatchCtx& bc, uint32_t li0, bool vec, ChunkRaw<PPT>& r) {
    const uint32_t base = bc.first + li0;
    const uint32_t n = li0 < bc.active ? min((uint32_t)PPT, bc.active - li0) : 0u;
    if (vec && n == (uint32_t)PPT) {
        // base is a multiple of PPT: 16-byte aligned words, PPT-byte aligned levels
        load_words<PPT>(c.xyzLow, base, r.lo);
        if (kLevel <= 1) load_words<PPT>(c.xyzMed, base, r.me);
        if (kLevel == 0) load_words<PPT>(c.xyzHigh, base, r.hi);
        if (!bc.keepAll) {
#pragma unroll
            for (int q = 0; q < PPT / 4; q++) r.lv4[q] = gload(reinterpret_cast<const uint32_t*>(c.levels + base) + q);
        }
    } else {
#pragma unroll
        for (int q = 0; q < (PPT + 3) / 4; q++) r.lv4[q] = 0u;
#pragma unroll
        for (int j = 0; j < PPT; j++) {
            const uint32_t pi = base + (uint32_t)j;
            const bool ok = (uint32_t)j < n;
            r.lo[j] = ok ? gload(c.xyzLow + pi) : 0u;
            r.me[j] = (ok && kLevel <= 1) ? gload(c.xyzMed + pi) : 0u;
            r.hi[j] = (ok && kLevel == 0) ? gload(c.xyzHigh + pi) : 0u;
            if (ok && !bc.keepAll) r.lv4[j / 4] |= (uint32_t)gload(c.levels + pi) << (8 * (j % 4));
        }
    }
}

// The clip-w, NDC and viewport tests (skipped when kInside: the batch box
// proves every point passes them), the y flip and the footprint radius (of
// the drawn position x, y, z; bx, by, bz: the decoded one, kept for the
// motion-blur sweep).
template <bool kNearestZ, bool kR0, bool kInside, int PPT>
__device__ __forceinline__ void chunk_tests(const PassParams& p, const BatchCtx& bc, const f4* clip, const float* rw,
                                            const bool* keep, const float* wx, const float* wy, const float* wz,
                                            const float* bx, const float* by, const float* bz, ChunkFrags<PPT>& f) {
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        bool ok = keep[j];
        const float nx = clip[j].x * rw[j], ny = clip[j].y * rw[j], nz = clip[j].z * rw[j];
        const int ix = f2i_hw(fmaf(nx, 0.5f, 0.5f) * (float)p.width);
        const int iy = f2i_hw(fmaf(ny, 0.5f, 0.5f) * (float)p.height);
        if (!kInside) {
            ok = ok && !(clip[j].w <= 0.0f);
            ok = ok && !(nx < -1.0f || ny < -1.0f || nx > 1.0f || ny > 1.0f);
            if (kNearestZ) ok = ok && !(nz <= 0.0f || nz > 1.0f);
            else ok = ok && !(nz < 0.0f || nz > 1.0f);
            // ix, iy >= 0 whenever the tests above pass: fma(n, 0.5, 0.5) of
            // n in [-1, 1] lies in [0, 1], and a NaN converts to 0
            ok = ok && !(ix >= p.width || iy >= p.height);
        }
        f.ok[j] = ok;
        f.px[j] = ix;
        f.py[j] = p.height - 1 - iy;
        f.z[j] = nz;
        f.r[j] = (kR0 || !ok) ? 0 : (bc.rconst >= 0 ? bc.rconst : footprint_radius(p, bc.W, wx[j], wy[j], wz[j]));
        f.wx[j] = bx[j];
        f.wy[j] = by[j];
        f.wz[j] = bz[j];
    }
}

// What a draw loop knows of its batch before the chunks, chosen once per
// batch so that no per-chunk branch merges the points' flags (the compiler
// packs merged flag arrays into bytes and unpacks them again, several VALU
// instructions per point).  kFixRuntime: the LOD/dither and inside tests are
// chosen per chunk.  Otherwise bits: kFixKeep, bc.keepAll (no LOD/dither
// test; without it the batch needs the test); kFixInside, bc.inside for
// single-pixel undisplaced points (the clip tests are skipped; without it
// they run).
constexpr int kFixRuntime = -1;
constexpr int kFixKeep = 1;
constexpr int kFixInside = 2;

// Runs the front end of every point pass on the fetched points: dither/LOD
// skip, decodePoint (Common.metal:73-117) and projection (project_q for
// undisplaced points), displacement (kDisp: position + displacement, a NaN
// component drops the point, then the transform; DepthPass/Shaders.metal:
// 69-74, ColorPass/Shaders.metal:72-78), clip and viewport tests, y flip,
// footprint radius.
template <bool kNearestZ, bool kR0, int kLevel, int PPT, bool kUseInside = true, bool kDisp = false,
          int kFix = kFixRuntime>
__device__ __forceinline__ void chunk_decode(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t li0,
                                    const ChunkRaw<PPT>& r, ChunkFrags<PPT>& f) {
    constexpr bool kRuntime = kFix == kFixRuntime;
    static_assert(kRuntime || !(kFix & kFixInside) || (kUseInside && kR0 && !kDisp),
                  "inside needs single-pixel undisplaced points");
    const uint32_t base = bc.first + li0;
    const uint32_t n = li0 < bc.active ? min((uint32_t)PPT, bc.active - li0) : 0u;
    // Branch-free per point: every test of the reference folds into `ok`
    // (each written as the negation of its reject condition, so NaNs behave
    // as in the Metal source) and the fragment is kept or dropped once.
    f4 clip[PPT];
    float wx[PPT], wy[PPT], wz[PPT];      // decoded
    float dx[PPT], dy[PPT], dz[PPT];      // drawn (decoded + displacement)
    bool keep[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) keep[j] = (uint32_t)j < n;
    if (kRuntime ? !bc.keepAll : !(kFix & kFixKeep)) {
        // LOD / dither test: a point is kept when its dither value lies below
        // t = (lodThreshold - level) + 0.5.  Dither values lie in [0, 1], so
        // a wave whose points all have t > 1 keeps them all without hashing
        // (one uniform branch per chunk; the active prefix ends at the
        // threshold's level, so most waves hold only lower levels).  The
        // dropped points leave the branch as one word of bits.
        // Whether any of the thread's levels (bytes & 7) reaches hashFrom:
        // four at once, level + 8 - hashFrom sets bit 3 of its byte exactly
        // when level >= hashFrom (no carry leaves a byte).
        const uint32_t bias = (8u - bc.hashFrom) * 0x01010101u;
        uint32_t reach = 0u;
#pragma unroll
        for (int q = 0; q < (PPT + 3) / 4; q++) reach |= ((r.lv4[q] & 0x07070707u) + bias) & 0x08080808u;
        uint32_t drop = 0u;
        if (__ballot(reach != 0u) != 0ull) {
#pragma unroll
            for (int j = 0; j < PPT; j++) {
                const float t = (bc.lodThreshold - (float)((r.lv4[j / 4] >> (8 * (j % 4))) & 7u)) + 0.5f;
                const float dither = (p.lodDither != 0) ? hash_unit(base + (uint32_t)j) : 0.5f;
                drop |= (dither >= t) ? 1u << j : 0u;
            }
        }
#pragma unroll
        for (int j = 0; j < PPT; j++) keep[j] = keep[j] && ((drop >> j) & 1u) == 0u;
    }
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        const uint32_t pi = base + (uint32_t)j;
        uint32_t qx, qy, qz;
        const uint32_t l = r.lo[j];
        if (kLevel == 0) {
            const uint32_t m = r.me[j], h = r.hi[j];
            qx = ((l & 1023u) << 20) | ((m & 1023u) << 10) | (h & 1023u);
            qy = (((l >> 10) & 1023u) << 20) | (((m >> 10) & 1023u) << 10) | ((h >> 10) & 1023u);
            qz = (((l >> 20) & 1023u) << 20) | (((m >> 20) & 1023u) << 10) | ((h >> 20) & 1023u);
        } else if (kLevel == 1) {
            const uint32_t m = r.me[j];
            qx = ((l & 1023u) << 10) | (m & 1023u);
            qy = (((l >> 10) & 1023u) << 10) | ((m >> 10) & 1023u);
            qz = (((l >> 20) & 1023u) << 10) | ((m >> 20) & 1023u);
        } else {
            qx = l & 1023u;
            qy = (l >> 10) & 1023u;
            qz = (l >> 20) & 1023u;
        }
        wx[j] = fmaf((float)qx, bc.sx, bc.mnx);
        wy[j] = fmaf((float)qy, bc.sy, bc.mny);
        wz[j] = fmaf((float)qz, bc.sz, bc.mnz);
        if (kDisp) {
            // a cloud without a displacement buffer reads zeros; points past
            // the drawn range read nothing
            const float4 d = (c.disp && (uint32_t)j < n) ? gload_rec(c.disp + pi) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            dx[j] = wx[j] + d.x;
            dy[j] = wy[j] + d.y;
            dz[j] = wz[j] + d.z;
            keep[j] = keep[j] && !(__builtin_isnan(dx[j]) || __builtin_isnan(dy[j]) || __builtin_isnan(dz[j]));
        } else {
            dx[j] = wx[j];
            dy[j] = wy[j];
            dz[j] = wz[j];
        }
        // undisplaced points: the batch projection of the quantised
        // coordinates; displaced ones: decode + displacement, then the
        // transform
        clip[j] = kDisp ? mul_point(bc.T, dx[j], dy[j], dz[j]) : project_q(bc.aff, (float)qx, (float)qy, (float)qz);
    }
    // 1 / w, one wave-uniform branch per chunk
    float rw[PPT];
    if (bc.fastRcp) {
#pragma unroll
        for (int j = 0; j < PPT; j++) rw[j] = rcp_fast(clip[j].w);
    } else {
#pragma unroll
        for (int j = 0; j < PPT; j++) rw[j] = rcp(clip[j].w);
    }
    // a batch whose box proves that every point passes the clip-w, NDC and
    // viewport tests (kBoundsInside) skips them: one uniform branch per chunk
    if (!kRuntime) {
        if constexpr (kFix > 0 && (kFix & kFixInside) != 0)
            chunk_tests<kNearestZ, kR0, true, PPT>(p, bc, clip, rw, keep, dx, dy, dz, wx, wy, wz, f);
        else
            chunk_tests<kNearestZ, kR0, false, PPT>(p, bc, clip, rw, keep, dx, dy, dz, wx, wy, wz, f);
    } else if (kUseInside && !kDisp && kR0 && bc.inside) {
        chunk_tests<kNearestZ, kR0, true, PPT>(p, bc, clip, rw, keep, dx, dy, dz, wx, wy, wz, f);
    } else {
        chunk_tests<kNearestZ, kR0, false, PPT>(p, bc, clip, rw, keep, dx, dy, dz, wx, wy, wz, f);
    }
}

template <bool kNearestZ, bool kR0, int kLevel, int PPT, bool kDisp = false, int kFix = kFixRuntime>
__device__ __forceinline__ void chunk_front(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t li0,
                                   bool vec, ChunkFrags<PPT>& f) {
    ChunkRaw<PPT> r;
    chunk_fetch<kLevel, PPT>(c, bc, li0, vec, r);
    chunk_decode<kNearestZ, kR0, kLevel, PPT, true, kDisp, kFix>(c, p, bc, li0, r, f);
}

// The fix a batch's draw loop takes (single-pixel undisplaced points may skip
// the clip tests; kUseInside false: never), and the loop run with it as a
// compile-time constant: loop(std::integral_constant<int, fix>).
template <bool kR0, bool kDisp, bool kUseInside = true>
__device__ __forceinline__ int front_fix(const BatchCtx& bc) {
    return (bc.keepAll ? kFixKeep : 0) | ((kUseInside && kR0 && !kDisp && bc.inside) ? kFixInside : 0);
}
// (The loops are lambdas marked always_inline: a draw body called from
// several places is otherwise outlined into a function call with a stack
// frame, measured up to twice as slow in the multi-view kernels.)
template <bool kInsideOk, typename Loop>
__device__ __forceinline__ void with_fix(int fix, const Loop& loop) {
    if (fix & kFixInside) {
        if constexpr (kInsideOk) {
            if (fix & kFixKeep) loop(std::integral_constant<int, kFixKeep | kFixInside>{});
            else loop(std::integral_constant<int, kFixInside>{});
        }
    } else if (fix & kFixKeep) {
        loop(std::integral_constant<int, kFixKeep>{});
    } else {
        loop(std::integral_constant<int, 0>{});
    }
}

// ---------------------------------------------------------------- chunk rectangle

// Screen rectangle of the chunk's fragments (footprints included, clamped to
// the screen) and the number of fragment centres.
struct ChunkBox {
    int x0, y0, w, h, frags;
    // k / w == umulhi(k, magic) for 2 <= w <= 16384 and k < 16384 (checked
    // exhaustively); w == 1 has no 32-bit magic and divides by itself
    uint32_t magic;
    bool lds;
    bool none;   // (segment_box) the batch provably has no fragment
    float zmax;  // (segment_box) above every point's ndc z; +inf when unknown
    // every on-screen footprint pixel of the chunk lies inside the box (it is
    // the clamped union of the footprints), so no bounds test is needed
    // a box holds at most 2^24 pixels per side: one 24-bit multiply-add
    // unsigned arithmetic: defined (and discarded by the callers) for the
    // undefined pixel of a point without a fragment too
    __device__ inline int local(int tx, int ty) const {
        return (int)(__umul24((uint32_t)ty - (uint32_t)y0, (uint32_t)w) + ((uint32_t)tx - (uint32_t)x0));
    }
    __device__ inline size_t pixel(int k, int width) const {
        const int row = w == 1 ? k : (int)__umulhi((uint32_t)k, magic);
        const int col = k - row * w;
        return (size_t)(y0 + row) * (size_t)width + (size_t)(x0 + col);
    }
};

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_sum(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// The threads that share one chunk and its LDS tile: the whole workgroup
// (block barriers), or each wave on its own (a workgroup then works on NT/64
// chunks at once, with a private slice of the tile each and no block
// barriers; a wave's LDS operations complete in issue order, so a fence that
// keeps the compiler from reordering them is all the synchronisation needed).
template <bool kWave, int NT>
struct Group {
    static constexpr int kSize = kWave ? 64 : NT;
    static constexpr int kCount = kWave ? NT / 64 : 1;
    __device__ static int rank() { return kWave ? (int)__lane_id() : (int)threadIdx.x; }
    __device__ static int index() { return kWave ? (int)(threadIdx.x >> 6) : 0; }
    __device__ static void sync() {
        if (kWave) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
        } else {
            __syncthreads();
        }
    }
};

// Rectangle reduction over the group; red holds 5 * kMaxWaves ints (block
// groups only).  Ends with a barrier, so the caller may reuse LDS right after.
template <bool kWave, int NT, int PPT>
__device__ __forceinline__ ChunkBox chunk_box(const ChunkFrags<PPT>& f, const PassParams& p, int cap, int* red) {
    int x0 = 0x7fffffff, y0 = 0x7fffffff, x1 = -1, y1 = -1, n = 0;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        if (!f.ok[j]) continue;
        x0 = min(x0, f.px[j] - f.r[j]);
        x1 = max(x1, f.px[j] + f.r[j]);
        y0 = min(y0, f.py[j] - f.r[j]);
        y1 = max(y1, f.py[j] + f.r[j]);
        n++;
    }
    x0 = wave_min(x0);
    y0 = wave_min(y0);
    x1 = wave_max(x1);
    y1 = wave_max(y1);
    n = wave_sum(n);
    constexpr int kWaves = kWave ? 1 : NT / 64;
    const int w = threadIdx.x >> 6;
    if (!kWave && __lane_id() == 0) {
        lds_store(red + w, x0);
        lds_store(red + kWaves + w, y0);
        lds_store(red + 2 * kWaves + w, x1);
        lds_store(red + 3 * kWaves + w, y1);
        lds_store(red + 4 * kWaves + w, n);
    }
    if (!kWave) {
        __syncthreads();
        x0 = lds_load(red);
        y0 = lds_load(red + kWaves);
        x1 = lds_load(red + 2 * kWaves);
        y1 = lds_load(red + 3 * kWaves);
        n = lds_load(red + 4 * kWaves);
#pragma unroll
        for (int k = 1; k < kWaves; k++) {
            x0 = min(x0, lds_load(red + k));
            y0 = min(y0, lds_load(red + kWaves + k));
            x1 = max(x1, lds_load(red + 2 * kWaves + k));
            y1 = max(y1, lds_load(red + 3 * kWaves + k));
            n += lds_load(red + 4 * kWaves + k);
        }
        __syncthreads();
    }
    ChunkBox b;
    b.zmax = __builtin_inff();
    b.none = false;
    b.frags = n;
    b.x0 = max(x0, 0);
    b.y0 = max(y0, 0);
    x1 = min(x1, p.width - 1);
    y1 = min(y1, p.height - 1);
    b.w = x1 >= b.x0 ? x1 - b.x0 + 1 : 0;
    b.h = y1 >= b.y0 ? y1 - b.y0 + 1 : 0;
    const long long area = (long long)b.w * b.h;
    // LDS pays when the tile is small next to the fragments it absorbs
    b.lds = n > 0 && area > 0 && area <= cap && area <= 8LL * n;
    b.magic = (b.lds && b.w > 1) ? 0xffffffffu / (uint32_t)b.w + 1u : 0u;
    return b;
}

// Screen rectangle of a whole batch: the fragment-centre rectangle the cull
// pass proved from the batch box (batch_bounds), widened by the largest
// footprint radius (and by `extra`, the motion-blur sweep's reach), clamped to
// the screen.  Without that proof the batch has no rectangle (lds false,
// zmax +inf).  `none` is set when the proven rectangle misses the screen: the
// batch has no fragment at all.  `lds` is set when the rectangle fits `cap`
// pixels and is small next to the `points` the group draws from it; then
// every fragment of the batch (footprint pixels on the screen included)
// lies inside it.
__device__ __forceinline__ ChunkBox segment_box(const PassParams& p, const BatchCtx& bc, uint32_t points, int cap,
                                                int extra = 0) {
    ChunkBox b;
    b.frags = 0;
    b.none = false;
    b.zmax = __builtin_inff();
    b.x0 = b.y0 = b.w = b.h = 0;
    b.lds = false;
    b.magic = 0u;
    if (!bc.proven) return b;
    const int r = (bc.rconst >= 0 ? bc.rconst : p.maxRadius) + extra;
    b.zmax = bc.zmax;
    b.x0 = max(bc.bx0 - r, 0);
    const int bx1 = min(bc.bx1 + r, p.width - 1);
    // rows are flipped: py = height - 1 - iy
    b.y0 = max(p.height - 1 - (bc.by1 + r), 0);
    const int by1 = min(p.height - 1 - (bc.by0 - r), p.height - 1);
    b.w = bx1 >= b.x0 ? bx1 - b.x0 + 1 : 0;
    b.h = by1 >= b.y0 ? by1 - b.y0 + 1 : 0;
    const long long area = (long long)b.w * b.h;
    b.none = area == 0;
    b.lds = area > 0 && area <= cap && area <= 8LL * points;
    b.magic = (b.lds && b.w > 1) ? 0xffffffffu / (uint32_t)b.w + 1u : 0u;
    return b;
}

// Whether (x, y) lies in the rectangle.
__device__ __forceinline__ bool box_holds(const ChunkBox& b, int x, int y) {
    // both compares, no short circuit: no branch per fragment
    return ((uint32_t)(x - b.x0) < (uint32_t)b.w) & ((uint32_t)(y - b.y0) < (uint32_t)b.h);
}

// ---------------------------------------------------------------- depth pass body
// depthPassUpdate — DepthPass/Shaders.metal:21-116, for the local points
// [lo, hi) of one visible batch (lo a multiple of NT * PPT).  The group's
// chunks start at lo + group * group size * PPT and step by NT * PPT.
//
// Two ways to meet in LDS: when the batch's screen rectangle fits the tile
// (segment_box), every chunk of the group adds to one tile that is flushed
// once; otherwise each chunk reduces its own rectangle and flushes it.

// A chunk's depth maxima: with box.lds into the LDS tile (the box, the
// chunk's own rectangle or the batch's proven one, holds every fragment),
// else straight to the plane.
template <bool kR0, int PPT>
__device__ __forceinline__ void depth_frags(const ChunkFrags<PPT>& f, const ChunkBox& box, const PassParams& p, uint32_t* lds,
                                   uint32_t* __restrict__ depth) {
    // single-pixel points: a thread's consecutive points often share a
    // pixel (Morton order); a max the thread has already issued there
    // makes a smaller one a no-op, so it is skipped
    int lastK = -1;
    uint32_t lastD = 0u;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        if (!f.ok[j]) continue;
        const uint32_t d = depth_to_uint_hw(f.z[j]);
        if (kR0) {
            const int k = box.lds ? box.local(f.px[j], f.py[j]) : f.py[j] * p.width + f.px[j];
            if (k == lastK && d <= lastD) continue;
            lastD = (k == lastK) ? max(lastD, d) : d;
            lastK = k;
            if (box.lds) lds_max(lds + k, d);
            else gmax(depth + (size_t)k, d);
            continue;
        }
        const int r = f.r[j];
        for (int oy = -r; oy <= r; oy++)
            for (int ox = -r; ox <= r; ox++) {
                if (!inside_footprint(ox, oy, r)) continue;
                const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                if (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height) continue;
                if (box.lds) lds_max(lds + box.local(tx, ty), d);
                else gmax(depth + (size_t)ty * p.width + tx, d);
            }
    }
}

// depth_frags for single-pixel points meeting in a batch tile (the proven
// rectangle holds every fragment), with no per-fragment branches: a fragment
// that is invalid or already exceeded by this thread aims its max at the
// lane's spare word past the tile, which nothing reads.
template <int PPT>
__device__ __forceinline__ void depth_frags_tile(const ChunkFrags<PPT>& f, const ChunkBox& box,
                                                 const PassParams& p, uint32_t* lds, uint32_t* spare,
                                                 uint32_t* __restrict__ depth) {
    int lastK = -1;
    uint32_t lastD = 0u;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        const uint32_t d = depth_to_uint_hw(f.z[j]);
        const bool in = f.ok[j];
        const int k = box.local(f.px[j], f.py[j]);
        const bool fresh = in & !((k == lastK) & (d <= lastD));
        lastD = fresh ? d : lastD;
        lastK = fresh ? k : lastK;
        // the spare words are never read, so any value may land there
        lds_max(fresh ? lds + k : spare, d);
    }
}

// Flushes a tile to the depth plane.  No read-before-max: waiting for the
// read costs more than the atomics it would save (measured on the 100M frame).
template <typename G>
__device__ __forceinline__ void depth_flush(const ChunkBox& box, const PassParams& p, const uint32_t* lds,
                                   uint32_t* __restrict__ depth) {
    const int area = box.w * box.h;
    for (int i = G::rank(); i < area; i += G::kSize) {
        const uint32_t d = lds_load(lds + i);
        if (d == 0u) continue;
        gmax(depth + box.pixel(i, p.width), d);
    }
}

template <bool kR0, bool kWave, int NT, int PPT, int kLevel, int kTile, bool kDisp = false>
__device__ __forceinline__ void depth_chunks_level(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t lo,
                                    uint32_t* __restrict__ depth, uint32_t* ldsAll, int* red,
        const ChunkBox& seg) {
    using G = Group<kWave, NT>;
    constexpr int kCap = kTile / G::kCount;
    uint32_t* lds = ldsAll + G::index() * kCap;
    const bool vec = vec_ok<PPT>(bc);
    const uint32_t first = lo + (uint32_t)(G::index() * G::kSize * PPT);
    if (seg.lds) {
        const int area = seg.w * seg.h;
        for (int i = G::rank(); i < area; i += G::kSize) lds_store(lds + i, 0u);
        G::sync();
        uint32_t* spare = ldsAll + kTile + __lane_id();
        const auto loop = [&](auto fix) __attribute__((always_inline)) {
            for (uint32_t cb = first; cb < bc.active; cb += NT * PPT) {
                ChunkFrags<PPT> f;
                chunk_front<false, kR0, kLevel, PPT, kDisp, decltype(fix)::value>(c, p, bc, cb + G::rank() * PPT, vec,
                                                                                  f);
                if (kR0) depth_frags_tile<PPT>(f, seg, p, lds, spare, depth);
                else depth_frags<kR0, PPT>(f, seg, p, lds, depth);
            }
        };
        with_fix<kR0 && !kDisp>(front_fix<kR0, kDisp>(bc), loop);
        G::sync();
        depth_flush<G>(seg, p, lds, depth);
        G::sync();
        return;
    }
    const auto loop = [&](auto fix) __attribute__((always_inline)) {
        for (uint32_t cb = first; cb < bc.active; cb += NT * PPT) {
            ChunkFrags<PPT> f;
            chunk_front<false, kR0, kLevel, PPT, kDisp, decltype(fix)::value>(c, p, bc, cb + G::rank() * PPT, vec, f);
            const ChunkBox box = chunk_box<kWave, NT, PPT>(f, p, kCap, red);
            if (box.frags == 0) continue;
            if (box.lds) {
                const int area = box.w * box.h;
                for (int i = G::rank(); i < area; i += G::kSize) lds_store(lds + i, 0u);
                G::sync();
                // the chunk's rectangle holds every one of its fragments
                if (kR0) depth_frags_tile<PPT>(f, box, p, lds, ldsAll + kTile + __lane_id(), depth);
                else depth_frags<kR0, PPT>(f, box, p, lds, depth);
                G::sync();
                depth_flush<G>(box, p, lds, depth);
                G::sync();
            } else {
                depth_frags<kR0, PPT>(f, box, p, lds, depth);
            }
        }
    };
    with_fix<kR0 && !kDisp>(front_fix<kR0, kDisp>(bc), loop);
}

template <bool kR0, bool kWave, int NT, int PPT, int kTile = kDepthTile, bool kDisp = false>
__device__ __forceinline__ void depth_chunks(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t lo,
                                    uint32_t* __restrict__ depth, uint32_t* ldsAll, int* red) {
    using G = Group<kWave, NT>;
    constexpr int kCap = kTile / G::kCount;
    const ChunkBox seg = segment_box(p, bc, bc.active > lo ? (bc.active - lo) / G::kCount : 0u, kCap);
    if (seg.none) return;  // no point of the batch reaches the screen
    // one straight-line copy of the body per position precision
    if (bc.level == 0) depth_chunks_level<kR0, kWave, NT, PPT, 0, kTile, kDisp>(c, p, bc, lo, depth, ldsAll, red, seg);
    else if (bc.level == 1) depth_chunks_level<kR0, kWave, NT, PPT, 1, kTile, kDisp>(c, p, bc, lo, depth, ldsAll, red, seg);
    else depth_chunks_level<kR0, kWave, NT, PPT, 2, kTile, kDisp>(c, p, bc, lo, depth, ldsAll, red, seg);
}

// ---------------------------------------------------------------- colour pass body
// Whether the colour pass skips the per-point clip tests of batches proven
// inside the view (the depth and nearest passes do).
constexpr bool kColorInside = false;

// colorPassUpdate (opaque path) — ColorPass/Shaders.metal:30-219, one visible batch.
struct ColorLds {
    uint32_t* depth;   // per tile pixel: color_threshold of its depth, as float bits
    unsigned long long* rgn;  // per tile pixel: red | green << 24 | count << 48
    uint32_t* b;              // per tile pixel: blue
    int* red;
};

// Built from arguments rather than a constant initializer: a constant
// aggregate of LDS addresses would become a global the back end cannot emit.
__device__ __forceinline__ ColorLds color_lds(uint32_t* depth, unsigned long long* rgn, uint32_t* b, int* red) {
    ColorLds L;
    L.depth = depth;
    L.rgn = rgn;
    L.b = b;
    L.red = red;
    return L;
}

// The colour test accepts a fragment at ndc z when z >= this: the pixel's
// depth times depthKeep, or +inf (nothing accepted) where no point reached
// the pixel.  Same value and comparison as `d != 0 && z >= uint_to_depth(d)
// * depthKeep` for every z, NaN included.
__device__ __forceinline__ float color_threshold(uint32_t d, const PassParams& p) {
    return d != 0u ? uint_to_depth(d) * p.depthKeep : __builtin_inff();
}

// A fragment's or a run's sums: red | green << 24 | count << 48 and blue.
// Every tile and run holds points of one batch (at most 65,535, each adding
// at most 255 to a pixel once), so red, green and blue stay below 2^24 and
// the count below 2^16: the packed fields never carry into each other.
__device__ __forceinline__ unsigned long long color_rgn(uint32_t rgb) {
    return (unsigned long long)(rgb & 0xffu) | ((unsigned long long)((rgb >> 8) & 0xffu) << 24) | (1ull << 48);
}

// Adds a run's sums to LDS tile slot k or to global pixel k.
__device__ __forceinline__ void color_flush(bool lds, int k, unsigned long long rgn, uint32_t b, ColorLds L,
                                   uint32_t* red, uint32_t* green, uint32_t* blue, uint32_t* count) {
    if (lds) {
        lds_add(L.rgn + k, rgn);
        lds_add(L.b + k, b);
    } else {
        gadd(red + k, (uint32_t)rgn & 0xffffffu);
        gadd(green + k, (uint32_t)(rgn >> 24) & 0xffffffu);
        gadd(blue + k, b);
        gadd(count + k, (uint32_t)(rgn >> 48));
    }
}

// Issues the colour loads of the thread's PPT points (with the position
// loads, so that both are in flight while the front end waits for the first).
template <int PPT>
__device__ __forceinline__ void color_fetch(const CloudDev& c, const BatchCtx& bc, uint32_t li0, bool vec,
                                            uint32_t* rgb) {
    const uint32_t base = bc.first + li0;
    const uint32_t n = li0 < bc.active ? min((uint32_t)PPT, bc.active - li0) : 0u;
    if (vec && n == (uint32_t)PPT) {
        load_words<PPT>(c.colors, base, rgb);
    } else {
#pragma unroll
        for (int j = 0; j < PPT; j++) rgb[j] = (uint32_t)j < n ? gload(c.colors + base + (uint32_t)j) : 0u;
    }
}

// The colour stage (ColorPass/Shaders.metal:95-133) of a chunk: packed
// r | g << 8 | b << 16 per fragment; a tint discard drops the fragment.
template <int PPT>
__device__ __forceinline__ void chunk_colors(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t li0,
                                             ChunkFrags<PPT>& f, uint32_t* rgb) {
    if (p.colorizeChunks == 0 && p.colorizeOverdraw == 0 && p.applyTint == 0) {
        // plain colours (one uniform test per chunk): the point word's RGB
        // bytes; the words of points without a fragment are never read
#pragma unroll
        for (int j = 0; j < PPT; j++) rgb[j] &= 0x00ffffffu;
        return;
    }
    const uint32_t base = bc.first + li0;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        const uint32_t pi = base + (uint32_t)j;
        uint32_t color = rgb[j];
        rgb[j] = 0u;
        if (!f.ok[j]) continue;
        if (p.colorizeChunks != 0) color = bc.batchIndex * 1234567u;
        else if (p.colorizeOverdraw != 0) color = 0x00010101u;
        uint32_t r8 = color & 0xffu, g8 = (color >> 8) & 0xffu, b8 = (color >> 16) & 0xffu;
        if (p.applyTint != 0) {
            const float4 t = c.tints ? gload_rec(c.tints + pi) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (t.w < 0.0f) {  // discard sentinel, ColorPass/Shaders.metal:119
                f.ok[j] = false;
                continue;
            }
            const float w = saturate(t.w);
            const float ox = (float)r8 * kInv255, oy = (float)g8 * kInv255, oz = (float)b8 * kInv255;
            const float mx = ox + (saturate(t.x) - ox) * w;
            const float my = oy + (saturate(t.y) - oy) * w;
            const float mz = oz + (saturate(t.z) - oz) * w;
            r8 = f2u(saturate(mx) * 255.0f + 0.5f);
            g8 = f2u(saturate(my) * 255.0f + 0.5f);
            b8 = f2u(saturate(mz) * 255.0f + 0.5f);
        }
        rgb[j] = r8 | (g8 << 8) | (b8 << 16);
    }
}

// A chunk's colour sums (same routes as depth_frags).  Single-pixel points:
// runs of a thread's consecutive fragments on one tile pixel (Morton order)
// are summed in registers and added once.
template <bool kR0, int PPT>
__device__ __forceinline__ void color_frags(const ChunkFrags<PPT>& f, const uint32_t* rgb, const ChunkBox& box,
                                   const PassParams& p, const uint32_t* __restrict__ depth, uint32_t* red,
                                   uint32_t* green, uint32_t* blue, uint32_t* count, ColorLds L) {
    int runK = -1;
    unsigned long long runRGN = 0ull;
    uint32_t runB = 0u;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        if (!f.ok[j]) continue;
        const unsigned long long rgn = color_rgn(rgb[j]);
        const uint32_t b8 = (rgb[j] >> 16) & 0xffu;
        if (kR0) {
            const int k = box.lds ? box.local(f.px[j], f.py[j]) : f.py[j] * p.width + f.px[j];
            const float thr = box.lds ? __uint_as_float(lds_load(L.depth + k)) : color_threshold(gload(depth + k), p);
            if (!(f.z[j] >= thr)) continue;
            if (k == runK) {
                runRGN += rgn;
                runB += b8;
                continue;
            }
            if (runK >= 0) color_flush(box.lds, runK, runRGN, runB, L, red, green, blue, count);
            runK = k;
            runRGN = rgn;
            runB = b8;
            continue;
        }
        const int rad = f.r[j];
        for (int oy = -rad; oy <= rad; oy++)
            for (int ox = -rad; ox <= rad; ox++) {
                if (!inside_footprint(ox, oy, rad)) continue;
                const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                if (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height) continue;
                const size_t pix = (size_t)ty * p.width + tx;
                const bool inTile = box.lds;
                const int k = inTile ? box.local(tx, ty) : 0;
                const float thr =
                    inTile ? __uint_as_float(lds_load(L.depth + k)) : color_threshold(gload(depth + pix), p);
                if (!(f.z[j] >= thr)) continue;
                if (inTile) {
                    lds_add(L.rgn + k, rgn);
                    lds_add(L.b + k, b8);
                } else {
                    gadd(red + pix, rgb[j] & 0xffu);
                    gadd(green + pix, (rgb[j] >> 8) & 0xffu);
                    gadd(blue + pix, b8);
                    gadd(count + pix, 1u);
                }
            }
    }
    if (kR0 && runK >= 0) color_flush(box.lds, runK, runRGN, runB, L, red, green, blue, count);
}

// color_frags for single-pixel points meeting in a batch tile (the proven
// rectangle holds every fragment): one branch per fragment (accepted or not)
// instead of one per test.
template <int PPT>
__device__ __forceinline__ void color_frags_tile(const ChunkFrags<PPT>& f, const uint32_t* rgb, const ChunkBox& box,
                                                 const PassParams& p, const uint32_t* __restrict__ depth,
                                                 uint32_t* red, uint32_t* green, uint32_t* blue, uint32_t* count,
                                                 ColorLds L) {
    // every fragment's tile slot (slot 0 for points without one) and its
    // threshold, read before the first sum is added
    int ks[PPT];
    float thr[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        const int k = box.local(f.px[j], f.py[j]);
        ks[j] = f.ok[j] ? k : 0;
    }
#pragma unroll
    for (int j = 0; j < PPT; j++) thr[j] = __uint_as_float(lds_load(L.depth + ks[j]));
    int runK = -1;
    unsigned long long runRGN = 0ull;
    uint32_t runB = 0u;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        const int k = ks[j];
        const bool pass = f.ok[j] & (f.z[j] >= thr[j]);
        if (!pass) continue;
        const unsigned long long rgn = color_rgn(rgb[j]);
        const uint32_t b8 = (rgb[j] >> 16) & 0xffu;
        if (k == runK) {
            runRGN += rgn;
            runB += b8;
            continue;
        }
        if (runK >= 0) {
            lds_add(L.rgn + runK, runRGN);
            lds_add(L.b + runK, runB);
        }
        runK = k;
        runRGN = rgn;
        runB = b8;
    }
    if (runK >= 0) {
        lds_add(L.rgn + runK, runRGN);
        lds_add(L.b + runK, runB);
    }
}

// Stages the tile's acceptance thresholds and clears its sums.
template <typename G>
__device__ __forceinline__ float color_stage(const ChunkBox& box, const PassParams& p, const uint32_t* __restrict__ depth,
                                   ColorLds L) {
    // stages each pixel's color_threshold; also returns the smallest of the
    // thread's pixels (+inf where no pixel has a depth)
    const int area = box.w * box.h;
    float tmin = __builtin_inff();
    for (int i = G::rank(); i < area; i += G::kSize) {
        const float thr = color_threshold(gload(depth + box.pixel(i, p.width)), p);
        lds_store(L.depth + i, __float_as_uint(thr));
        lds_store(L.rgn + i, 0ull);
        lds_store(L.b + i, 0u);
        tmin = fminf(tmin, thr);
    }
    return tmin;
}

// Adds a tile's sums (one batch's, packed as color_rgn describes) to the planes.
template <typename G>
__device__ __forceinline__ void color_tile_flush(const ChunkBox& box, const PassParams& p, ColorLds L, uint32_t* red,
                                        uint32_t* green, uint32_t* blue, uint32_t* count) {
    const int area = box.w * box.h;
    for (int i = G::rank(); i < area; i += G::kSize) {
        const unsigned long long rgn = lds_load(L.rgn + i);
        const uint32_t n = (uint32_t)(rgn >> 48);
        if (n == 0u) continue;
        const size_t pix = box.pixel(i, p.width);
        gadd(red + pix, (uint32_t)rgn & 0xffffffu);
        gadd(green + pix, (uint32_t)(rgn >> 24) & 0xffffffu);
        gadd(blue + pix, lds_load(L.b + i));
        gadd(count + pix, n);
    }
}

template <bool kR0, bool kWave, int NT, int PPT, int kLevel, bool kDisp = false, int kTile = kColorTile>
__device__ __forceinline__ void color_chunks_level(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t lo,
                                    const uint32_t* __restrict__ depth, uint32_t* red, uint32_t* green,
                                    uint32_t* blue, uint32_t* count, ColorLds LAll,
        const ChunkBox& seg) {
    using G = Group<kWave, NT>;
    constexpr int kCap = kTile / G::kCount;
    const int slice = G::index() * kCap;
    const ColorLds L = color_lds(LAll.depth + slice, LAll.rgn + slice, LAll.b + slice, LAll.red);
    const bool vec = vec_ok<PPT>(bc);
    const uint32_t first = lo + (uint32_t)(G::index() * G::kSize * PPT);
    if (seg.lds) {
        // the tile's smallest threshold: one barrier makes both the staged
        // tile and the waves' minima visible (the reduction words are next
        // written by the following batch, after this one's flush barriers)
        float tmin = color_stage<G>(seg, p, depth, L);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) tmin = fminf(tmin, __shfl_xor(tmin, o));
        if (!kWave && __lane_id() == 0) lds_store(L.red + (threadIdx.x >> 6), __float_as_int(tmin));
        G::sync();
        if (!kWave) {
#pragma unroll
            for (int k = 0; k < NT / 64; k++) tmin = fminf(tmin, __int_as_float(lds_load(L.red + k)));
        }
        // every point of the batch lies behind every accepted depth of its
        // rectangle: no fragment can pass the test, nothing to add
        if (tmin > seg.zmax) return;
        // plain colours (no tint, no debug colours) are known for the whole
        // batch: the fragments' flags then cross no colour-stage branch
        const auto loop = [&](auto fix, auto plain) __attribute__((always_inline)) {
            for (uint32_t cb = first; cb < bc.active; cb += NT * PPT) {
                const uint32_t li0 = cb + G::rank() * PPT;
                ChunkRaw<PPT> raw;
                chunk_fetch<kLevel, PPT>(c, bc, li0, vec, raw);
                uint32_t rgb[PPT];
                color_fetch<PPT>(c, bc, li0, vec, rgb);
                ChunkFrags<PPT> f;
                chunk_decode<false, kR0, kLevel, PPT, kColorInside, kDisp, decltype(fix)::value>(c, p, bc, li0, raw, f);
                if constexpr (decltype(plain)::value) {
#pragma unroll
                    for (int j = 0; j < PPT; j++) rgb[j] &= 0x00ffffffu;
                } else {
                    chunk_colors<PPT>(c, p, bc, li0, f, rgb);
                }
                if (kR0) color_frags_tile<PPT>(f, rgb, seg, p, depth, red, green, blue, count, L);
                else color_frags<kR0, PPT>(f, rgb, seg, p, depth, red, green, blue, count, L);
            }
        };
        if (p.colorizeChunks == 0 && p.colorizeOverdraw == 0 && p.applyTint == 0) {
            with_fix<kColorInside && kR0 && !kDisp>(front_fix<kR0, kDisp, kColorInside>(bc),
                                                    [&](auto fix) __attribute__((always_inline)) { loop(fix, std::true_type{}); });
        } else {
            loop(std::integral_constant<int, kFixRuntime>{}, std::false_type{});
        }
        G::sync();
        color_tile_flush<G>(seg, p, L, red, green, blue, count);
        G::sync();
        return;
    }
    const auto loop = [&](auto fix, auto plain) __attribute__((always_inline)) {
        for (uint32_t cb = first; cb < bc.active; cb += NT * PPT) {
            const uint32_t li0 = cb + G::rank() * PPT;
            ChunkRaw<PPT> raw;
            chunk_fetch<kLevel, PPT>(c, bc, li0, vec, raw);
            uint32_t rgb[PPT];
            color_fetch<PPT>(c, bc, li0, vec, rgb);
            ChunkFrags<PPT> f;
            chunk_decode<false, kR0, kLevel, PPT, kColorInside, kDisp, decltype(fix)::value>(c, p, bc, li0, raw, f);
            if constexpr (decltype(plain)::value) {
#pragma unroll
                for (int j = 0; j < PPT; j++) rgb[j] &= 0x00ffffffu;
            } else {
                chunk_colors<PPT>(c, p, bc, li0, f, rgb);
            }
            const ChunkBox box = chunk_box<kWave, NT, PPT>(f, p, kCap, L.red);
            if (box.frags == 0) continue;
            if (box.lds) {
                color_stage<G>(box, p, depth, L);
                G::sync();
                // the chunk's rectangle holds every one of its fragments
                if (kR0) color_frags_tile<PPT>(f, rgb, box, p, depth, red, green, blue, count, L);
                else color_frags<kR0, PPT>(f, rgb, box, p, depth, red, green, blue, count, L);
                G::sync();
                color_tile_flush<G>(box, p, L, red, green, blue, count);
                G::sync();
            } else {
                color_frags<kR0, PPT>(f, rgb, box, p, depth, red, green, blue, count, L);
            }
        }
    };
    if (p.colorizeChunks == 0 && p.colorizeOverdraw == 0 && p.applyTint == 0) {
        with_fix<kColorInside && kR0 && !kDisp>(front_fix<kR0, kDisp, kColorInside>(bc),
                                                [&](auto fix) __attribute__((always_inline)) { loop(fix, std::true_type{}); });
    } else {
        loop(std::integral_constant<int, kFixRuntime>{}, std::false_type{});
    }
}

template <bool kR0, bool kWave, int NT, int PPT, bool kDisp = false, int kTile = kColorTile>
__device__ __forceinline__ void color_chunks(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t lo,
                                    const uint32_t* __restrict__ depth, uint32_t* red, uint32_t* green,
                                    uint32_t* blue, uint32_t* count, ColorLds LAll) {
    using G = Group<kWave, NT>;
    constexpr int kCap = kTile / G::kCount;
    const ChunkBox seg = segment_box(p, bc, bc.active > lo ? (bc.active - lo) / G::kCount : 0u, kCap);
    if (seg.none) return;  // no point of the batch reaches the screen
    // one straight-line copy of the body per position precision
    if (bc.level == 0) color_chunks_level<kR0, kWave, NT, PPT, 0, kDisp, kTile>(c, p, bc, lo, depth, red, green, blue, count, LAll, seg);
    else if (bc.level == 1) color_chunks_level<kR0, kWave, NT, PPT, 1, kDisp, kTile>(c, p, bc, lo, depth, red, green, blue, count, LAll, seg);
    else color_chunks_level<kR0, kWave, NT, PPT, 2, kDisp, kTile>(c, p, bc, lo, depth, red, green, blue, count, LAll, seg);
}

// ---------------------------------------------------------------- nearest pass body
// nearestDepthUpdate + nearestIndexUpdate (NearestDepth/Shaders.metal:14-91,
// NearestIndex/Shaders.metal:14-94) as one max of depth << 32 | ~index.
template <bool kR0, int PPT>
__device__ __forceinline__ void nearest_frags(const ChunkFrags<PPT>& f, uint32_t base, const ChunkBox& box,
                                     const PassParams& p, unsigned long long* lds,
                                     unsigned long long* __restrict__ packed) {
    int lastK = -1;
    unsigned long long lastKey = 0ull;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        if (!f.ok[j]) continue;
        const uint32_t pi = base + (uint32_t)j;
        // max of depth << 32 | ~index == the two-pass rule: deepest
        // reverse-Z depth, then the lowest point index among equal depths
        const unsigned long long key =
            ((unsigned long long)depth_to_uint_hw(f.z[j]) << 32) | (unsigned long long)(~pi);
        if (kR0) {
            // as in the depth pass: skip a max this thread already exceeded
            const int k = box.lds ? box.local(f.px[j], f.py[j]) : f.py[j] * p.width + f.px[j];
            if (k == lastK && key <= lastKey) continue;
            lastKey = (k == lastK) ? (key > lastKey ? key : lastKey) : key;
            lastK = k;
            if (box.lds) lds_max(lds + k, key);
            else gmax(packed + (size_t)k, key);
            continue;
        }
        const int r = f.r[j];
        for (int oy = -r; oy <= r; oy++)
            for (int ox = -r; ox <= r; ox++) {
                if (!inside_footprint(ox, oy, r)) continue;
                const int tx = f.px[j] + ox, ty = f.py[j] + oy;
                if (tx < 0 || tx >= p.width || ty < 0 || ty >= p.height) continue;
                if (box.lds) lds_max(lds + box.local(tx, ty), key);
                else gmax(packed + (size_t)ty * p.width + tx, key);
            }
    }
}

// nearest_frags for single-pixel points meeting in a batch tile (the proven
// rectangle holds every fragment), with no per-fragment branch (as
// depth_frags_tile): a fragment that is invalid or already exceeded by this
// thread aims its max at the lane's spare word, which nothing reads.
template <int PPT>
__device__ __forceinline__ void nearest_frags_tile(const ChunkFrags<PPT>& f, uint32_t base, const ChunkBox& box,
                                                   const PassParams& p, unsigned long long* lds,
                                                   unsigned long long* spare,
                                                   unsigned long long* __restrict__ packed) {
    int lastK = -1;
    unsigned long long lastKey = 0ull;
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        // max of depth << 32 | ~index == the two-pass rule: deepest
        // reverse-Z depth, then the lowest point index among equal depths
        const unsigned long long key =
            ((unsigned long long)depth_to_uint_hw(f.z[j]) << 32) | (unsigned long long)(~(base + (uint32_t)j));
        const bool in = f.ok[j];
        const int k = box.local(f.px[j], f.py[j]);
        const bool fresh = in & !((k == lastK) & (key <= lastKey));
        lastKey = fresh ? key : lastKey;
        lastK = fresh ? k : lastK;
        // the spare words are never read, so any value may land there
        lds_max(fresh ? lds + k : spare, key);
    }
}

template <typename G>
__device__ __forceinline__ void nearest_flush(const ChunkBox& box, const PassParams& p, const unsigned long long* lds,
                                     unsigned long long* __restrict__ packed) {
    const int area = box.w * box.h;
    for (int i = G::rank(); i < area; i += G::kSize) {
        const unsigned long long k = lds_load(lds + i);
        if (k == 0ull) continue;
        gmax(packed + box.pixel(i, p.width), k);
    }
}

template <bool kR0, bool kWave, int NT, int PPT, int kLevel, int kTile>
__device__ __forceinline__ void nearest_chunks_level(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t lo,
                                      unsigned long long* __restrict__ packed, unsigned long long* ldsAll, int* red,
        const ChunkBox& seg) {
    using G = Group<kWave, NT>;
    constexpr int kCap = kTile / G::kCount;
    unsigned long long* lds = ldsAll + G::index() * kCap;
    const bool vec = vec_ok<PPT>(bc);
    const uint32_t first = lo + (uint32_t)(G::index() * G::kSize * PPT);
    if (seg.lds) {
        const int area = seg.w * seg.h;
        for (int i = G::rank(); i < area; i += G::kSize) lds_store(lds + i, 0ull);
        G::sync();
        unsigned long long* spare = ldsAll + kTile + __lane_id();
        const auto loop = [&](auto fix) __attribute__((always_inline)) {
            for (uint32_t cb = first; cb < bc.active; cb += NT * PPT) {
                const uint32_t li0 = cb + G::rank() * PPT;
                ChunkFrags<PPT> f;
                chunk_front<true, kR0, kLevel, PPT, false, decltype(fix)::value>(c, p, bc, li0, vec, f);
                if (kR0) nearest_frags_tile<PPT>(f, bc.first + li0, seg, p, lds, spare, packed);
                else nearest_frags<kR0, PPT>(f, bc.first + li0, seg, p, lds, packed);
            }
        };
        with_fix<kR0>(front_fix<kR0, false>(bc), loop);
        G::sync();
        nearest_flush<G>(seg, p, lds, packed);
        G::sync();
        return;
    }
    const auto loop = [&](auto fix) __attribute__((always_inline)) {
        for (uint32_t cb = first; cb < bc.active; cb += NT * PPT) {
            const uint32_t li0 = cb + G::rank() * PPT;
            ChunkFrags<PPT> f;
            chunk_front<true, kR0, kLevel, PPT, false, decltype(fix)::value>(c, p, bc, li0, vec, f);
            const ChunkBox box = chunk_box<kWave, NT, PPT>(f, p, kCap, red);
            if (box.frags == 0) continue;
            if (box.lds) {
                const int area = box.w * box.h;
                for (int i = G::rank(); i < area; i += G::kSize) lds_store(lds + i, 0ull);
                G::sync();
                // the chunk's rectangle holds every one of its fragments
                if (kR0) nearest_frags_tile<PPT>(f, bc.first + li0, box, p, lds, ldsAll + kTile + __lane_id(), packed);
                else nearest_frags<kR0, PPT>(f, bc.first + li0, box, p, lds, packed);
                G::sync();
                nearest_flush<G>(box, p, lds, packed);
                G::sync();
            } else {
                nearest_frags<kR0, PPT>(f, bc.first + li0, box, p, lds, packed);
            }
        }
    };
    with_fix<kR0>(front_fix<kR0, false>(bc), loop);
}

template <bool kR0, bool kWave, int NT, int PPT, int kTile = kNearestTile>
__device__ __forceinline__ void nearest_chunks(const CloudDev& c, const PassParams& p, const BatchCtx& bc, uint32_t lo,
                                      unsigned long long* __restrict__ packed, unsigned long long* ldsAll, int* red) {
    using G = Group<kWave, NT>;
    constexpr int kCap = kTile / G::kCount;
    const ChunkBox seg = segment_box(p, bc, bc.active > lo ? (bc.active - lo) / G::kCount : 0u, kCap);
    if (seg.none) return;  // no point of the batch reaches the screen
    // one straight-line copy of the body per position precision
    if (bc.level == 0) nearest_chunks_level<kR0, kWave, NT, PPT, 0, kTile>(c, p, bc, lo, packed, ldsAll, red, seg);
    else if (bc.level == 1) nearest_chunks_level<kR0, kWave, NT, PPT, 1, kTile>(c, p, bc, lo, packed, ldsAll, red, seg);
    else nearest_chunks_level<kR0, kWave, NT, PPT, 2, kTile>(c, p, bc, lo, packed, ldsAll, red, seg);
}

// ---------------------------------------------------------------- per-batch entry points
// A batch whose points all have radius 0 runs the single-pixel code.

template <bool kR0, bool kWave, int NT, int PPT, int kTile = kDepthTile, bool kDisp = false>
__device__ __forceinline__ void depth_batch(const CloudDev& c, const PassParams& p, const scr_visible_batch& vb,
                                            const DrawRecord* bb, uint32_t lo, uint32_t hi, uint32_t* __restrict__ depth,
                                            uint32_t* lds, int* red) {
    const BatchCtx bc = batch_ctx(c, p, vb, bb, hi);
    if (!kR0 && bc.rconst == 0) depth_chunks<true, kWave, NT, PPT, kTile, kDisp>(c, p, bc, lo, depth, lds, red);
    else depth_chunks<kR0, kWave, NT, PPT, kTile, kDisp>(c, p, bc, lo, depth, lds, red);
}

template <bool kR0, bool kWave, int NT, int PPT, bool kDisp = false, int kTile = kColorTile>
__device__ __forceinline__ void color_batch(const CloudDev& c, const PassParams& p, const scr_visible_batch& vb,
                                            const DrawRecord* bb, uint32_t lo, uint32_t hi,
                                            const uint32_t* __restrict__ depth, uint32_t* red, uint32_t* green,
                                            uint32_t* blue, uint32_t* count, ColorLds L) {
    const BatchCtx bc = batch_ctx(c, p, vb, bb, hi);
    if (!kR0 && bc.rconst == 0) color_chunks<true, kWave, NT, PPT, kDisp, kTile>(c, p, bc, lo, depth, red, green, blue, count, L);
    else color_chunks<kR0, kWave, NT, PPT, kDisp, kTile>(c, p, bc, lo, depth, red, green, blue, count, L);
}

template <bool kR0, bool kWave, int NT, int PPT, int kTile = kNearestTile>
__device__ __forceinline__ void nearest_batch(const CloudDev& c, const PassParams& p, const scr_visible_batch& vb,
                                              const DrawRecord* bb, uint32_t lo, uint32_t hi,
                                              unsigned long long* __restrict__ packed, unsigned long long* lds,
                                              int* red) {
    const BatchCtx bc = batch_ctx(c, p, vb, bb, hi);
    if (!kR0 && bc.rconst == 0) nearest_chunks<true, kWave, NT, PPT, kTile>(c, p, bc, lo, packed, lds, red);
    else nearest_chunks<kR0, kWave, NT, PPT, kTile>(c, p, bc, lo, packed, lds, red);
}

// ---------------------------------------------------------------- resolves

__device__ __forceinline__ uint32_t unorm8(float x) { return f2u(rintf(saturate(x) * 255.0f)); }

__device__ __forceinline__ uint32_t pack_rgba(float r, float g, float b, float a) {
    return unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (unorm8(a) << 24);
}

// resolveUpdate (count-average path) — Resolve/Shaders.metal:13-57
__device__ __forceinline__ void resolve_pixel(const PassParams& p, const uint32_t* __restrict__ planes, size_t stride, size_t i,
                                     uint32_t* __restrict__ rgba, float* __restrict__ depthf) {
    const uint32_t cnt = planes[4 * stride + i];
    if (cnt == 0u) {
        rgba[i] = pack_rgba(p.bg[0], p.bg[1], p.bg[2], 0.0f);
        depthf[i] = 0.0f;
        return;
    }
    const float inv = rcp((float)cnt);
    const float r = ((float)planes[stride + i] * inv) * kInv255;
    const float g = ((float)planes[2 * stride + i] * inv) * kInv255;
    const float b = ((float)planes[3 * stride + i] * inv) * kInv255;
    rgba[i] = pack_rgba(r, g, b, 1.0f);
    depthf[i] = uint_to_depth(planes[i]);
}

// resolveUpdate, coverage path (Resolve/Shaders.metal:27-45): the planes hold
// sum(colour * a * 4096) and sum(a * 4096); colour = their ratio, alpha =
// saturate(1 - e^-sum(a)).
__device__ __forceinline__ void resolve_coverage_pixel(const PassParams& p, const uint32_t* __restrict__ planes,
                                                       size_t stride, size_t i, uint32_t* __restrict__ rgba,
                                                       float* __restrict__ depthf) {
    const uint32_t a = planes[4 * stride + i];
    if (a == 0u) {
        rgba[i] = pack_rgba(p.bg[0], p.bg[1], p.bg[2], 0.0f);
        depthf[i] = 0.0f;
        return;
    }
    const float inv = rcp((float)a);
    const float sumA = (float)a * (1.0f / 4096.0f);
    const float alpha = saturate(1.0f - scr_expf(-sumA));
    rgba[i] = pack_rgba((float)planes[stride + i] * inv, (float)planes[2 * stride + i] * inv,
                        (float)planes[3 * stride + i] * inv, alpha);
    depthf[i] = uint_to_depth(planes[i]);
}

// holeFillUpdate — HoleFill/Shaders.metal:10-48, one pixel of an RGBA8
// plane: a texel with alpha >= 0.5 (byte >= 128) passes through; an empty one
// takes the mean of its valid 8-neighbours' colours (each byte read as
// byte * (1/255), summed in the shader's row-major order, divided as
// sum * rcp(count)) with alpha 1, or stays as it is when none is valid.
// c points at the texel in a row-major block of texels `pitch` wide whose
// ring of eight neighbours is present (texels outside the image held as 0).
__device__ __forceinline__ uint32_t hole_fill_texels(const uint32_t* c0, int pitch) {
    const uint32_t c = c0[0];
    if ((c >> 24) >= 128u) return c;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f, n = 0.0f;
#pragma unroll
    for (int oy = -1; oy <= 1; oy++)
#pragma unroll
        for (int ox = -1; ox <= 1; ox++) {
            if (ox == 0 && oy == 0) continue;
            const uint32_t q = c0[oy * pitch + ox];
            if ((q >> 24) < 128u) continue;
            sr += (float)(q & 0xffu) * kInv255;
            sg += (float)((q >> 8) & 0xffu) * kInv255;
            sb += (float)((q >> 16) & 0xffu) * kInv255;
            n += 1.0f;
        }
    if (!(n > 0.0f)) return c;
    const float r = rcp(n);
    return pack_rgba(sr * r, sg * r, sb * r, 1.0f);
}

// The composite of ComputeRasteriser.draw (Post.metal:23-50) for one pixel:
// cloud texel c (RGBA8) with reverse-Z depth d onto target texel *t and depth
// *td.  The depth mode discards where d <= 0 (no cloud), keeps the fragment
// where the depth compare holds, optionally stores d; blending as the options
// say, with UNORM8 texels read as byte / 255 and written as rint(saturate(x) * 255).
__device__ __forceinline__ float unorm8_in(uint32_t texel, int ch) {
    return (float)((texel >> (8 * ch)) & 0xffu) / 255.0f;
}

__device__ __forceinline__ void composite_pixel(uint32_t c, float d, uint32_t* t, float* td,
                                                const scr_composite_options& o) {
    if (o.mode == SCR_COMPOSITE_DEPTH) {
        if (d <= 0.0f) return;  // discard_fragment (Post.metal:44-46)
        const float cur = *td;
        bool pass = true;
        switch (o.depthCompare) {
            case SCR_COMPARE_GREATER: pass = d > cur; break;
            case SCR_COMPARE_GREATER_EQUAL: pass = d >= cur; break;
            case SCR_COMPARE_LESS: pass = d < cur; break;
            case SCR_COMPARE_LESS_EQUAL: pass = d <= cur; break;
            default: break;
        }
        if (!pass) return;
        if (o.writeDepth) *td = d;
    }
    if (o.blend == SCR_BLEND_REPLACE) {
        *t = c;
        return;
    }
    const uint32_t dst = *t;
    const float sa = unorm8_in(c, 3), k = 1.0f - sa;
    uint32_t out = 0u;
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        const float v = unorm8_in(c, ch) * sa + unorm8_in(dst, ch) * k;
        out |= unorm8(v) << (8 * ch);
    }
    out |= unorm8(sa + unorm8_in(dst, 3) * k) << 24;
    *t = out;
}

// nearestResolveUpdate — NearestResolve/Shaders.metal:8-38
__device__ __forceinline__ void nearest_resolve_pixel(const PassParams& p, const unsigned long long* __restrict__ packed,
                                             const uint32_t* __restrict__ colors, size_t i, uint32_t* __restrict__ rgba,
                                             float* __restrict__ depthf) {
    const unsigned long long k = packed[i];
    const uint32_t d = (uint32_t)(k >> 32);
    const uint32_t idx = ~(uint32_t)k;
    if (d == 0u || idx == 0xffffffffu) {
        rgba[i] = pack_rgba(p.bg[0], p.bg[1], p.bg[2], 0.0f);
        depthf[i] = 0.0f;
        return;
    }
    const uint32_t col = colors[idx];
    rgba[i] = pack_rgba((float)(col & 0xffu) * kInv255, (float)((col >> 8) & 0xffu) * kInv255,
                        (float)((col >> 16) & 0xffu) * kInv255, 1.0f);
    depthf[i] = uint_to_depth(d);
}

}  // namespace draw
}  // namespace scr
