// scr_pack.hip — GPU packer for gfx950: the GPU counterpart of
// PackedPointCloudFixtures.pack (PackedPointCloudFixtures.swift:30-167) and
// GPUPacker.encodePack (GPUPacker.swift:204-392).
//
// Stages (all on the context stream):
//   1 bounds       block min/max → order-preserving u32 atomics (a few per block)  GPUPacker/Shaders.metal:65-105
//   2 morton       30-bit (10 b/axis, the reference) or 63-bit (21 b/axis) keys, identity payload :109-126
//   3 radix sort   stable LSD, 8-bit digits: one all-digit histogram pass decides which
//                  passes are trivial; per pass tile histograms → device-wide exclusive
//                  scan (digit-major) → stable scatter: ranked in LDS per tile, staged in
//                  tile-sorted order, each digit's run written as consecutive addresses;
//                  large inputs in one kernel per pass with decoupled look-back
//                  instead of the histogram and scan kernels                     :134-215
//   4 gather       the one gather from the caller's order: 16-byte point records
//                  (position + RGBA8 colour, written with the keys) into Morton
//                  order                                                           :219-235
//   5 LOD levels   dense voxel grids, atomicMin "first claimer", wave run de-dup; one
//                  kernel per level takes that level's winners and claims the next :317-371
//   6-9 finalize   one workgroup per batch: last level's winners, cum counts, AABB,
//                  the RasterBatch record, stable 8-way counting scatter by level,
//                  quantize (p - min) / size (IEEE division, as the CPU pack) into three
//                  10-bit words, colours and source indices in final order  :243-309, 383-414
#include <algorithm>
#include <cstring>
#include <vector>

#include "scr_internal.hpp"

namespace scr {
namespace {

constexpr int kThreads = 256;
constexpr int kChunks = 16;                  // keys per thread per tile
constexpr int kTile = kThreads * kChunks;    // 4096 keys per tile
constexpr int kScanSeg = kThreads * 16;      // elements per scan block

__device__ inline uint32_t float_flip(float f) {
    const uint32_t x = __float_as_uint(f);
    const uint32_t mask = (uint32_t)(-(int32_t)(x >> 31)) | 0x80000000u;
    return x ^ mask;
}

__host__ __device__ inline float float_unflip(uint32_t u) {
    const uint32_t mask = ((u >> 31) - 1u) | 0x80000000u;
    return __builtin_bit_cast(float, u ^ mask);
}

__device__ inline uint32_t spread10(uint32_t v) {
    uint32_t x = v & 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ inline uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ inline uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// ---------------------------------------------------------------- 1 bounds
__global__ __launch_bounds__(kThreads) void k_bounds(const float4* __restrict__ pos, size_t n,
                                                     uint32_t* __restrict__ bounds) {
    float mn[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
    float mx[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const float4 p = pos[i];
        mn[0] = fminf(mn[0], p.x); mn[1] = fminf(mn[1], p.y); mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x); mx[1] = fmaxf(mx[1], p.y); mx[2] = fmaxf(mx[2], p.z);
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off));
        }
    // the block's waves combine in LDS; one wave lane per word updates the
    // six device words (a few atomics per block, not per wave)
    __shared__ float red[kThreads / 64][6];
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0)
        for (int a = 0; a < 3; a++) {
            red[w][a] = mn[a];
            red[w][3 + a] = mx[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float v = red[0][a];
        for (int k = 1; k < kThreads / 64; k++) v = a < 3 ? fminf(v, red[k][a]) : fmaxf(v, red[k][a]);
        if (a < 3) atomicMin(bounds + a, float_flip(v));
        else atomicMax(bounds + a, float_flip(v));
    }
}

__global__ void k_bounds_init(uint32_t* bounds) {
    if (threadIdx.x < 3) bounds[threadIdx.x] = 0xffffffffu;
    else if (threadIdx.x < 6) bounds[threadIdx.x] = 0u;
}

// ---------------------------------------------------------------- 2 morton
// packGather colour conversion: uint(clamp(c, 0, 1) * 255) per channel (truncating)
__device__ inline uint32_t pack_color(float4 c) {
    const uint32_t r = f2u(fminf(fmaxf(c.x, 0.0f), 1.0f) * 255.0f);
    const uint32_t g = f2u(fminf(fmaxf(c.y, 0.0f), 1.0f) * 255.0f);
    const uint32_t b = f2u(fminf(fmaxf(c.z, 0.0f), 1.0f) * 255.0f);
    const uint32_t a = f2u(fminf(fmaxf(c.w, 0.0f), 1.0f) * 255.0f);
    return r | (g << 8) | (b << 16) | (a << 24);
}

// Keys, the identity payload, and each point's 16-byte record (x, y, z and
// its RGBA8 colour in the w bits) in the caller's order: the gather after the
// sort then fetches one record per point instead of a position and a colour.
template <typename K>
__global__ __launch_bounds__(kThreads) void k_morton(const float4* __restrict__ pos, const float4* __restrict__ col,
                                                     size_t n, const uint32_t* __restrict__ bounds,
                                                     K* __restrict__ keys, uint32_t* __restrict__ idx,
                                                     float4* __restrict__ rec) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    constexpr bool wide = sizeof(K) == 8;
    const float qmax = wide ? 2097151.0f : 1023.0f;
    float mn[3], sc[3];
    for (int a = 0; a < 3; a++) {
        mn[a] = float_unflip(bounds[a]);
        const float mx = float_unflip(bounds[3 + a]);
        sc[a] = qmax / fmaxf(mx - mn[a], 0.000001f);
    }
    const float4 p = pos[i];
    const float v[3] = {p.x, p.y, p.z};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) q[a] = f2u(fminf(fmaxf((v[a] - mn[a]) * sc[a], 0.0f), qmax));
    if constexpr (wide)
        keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    else
        keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    idx[i] = (uint32_t)i;
    rec[i] = make_float4(p.x, p.y, p.z, __uint_as_float(pack_color(col[i])));
}

// ---------------------------------------------------------------- 3 radix sort
template <typename K>
__global__ __launch_bounds__(kThreads) void k_all_digit_hist(const K* __restrict__ keys, size_t n,
                                                             uint32_t* __restrict__ ghist) {
    constexpr int P = sizeof(K);
    __shared__ uint32_t h[P * 256];
    for (int i = threadIdx.x; i < P * 256; i += kThreads) h[i] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const K k = keys[i];
#pragma unroll
        for (int p = 0; p < P; p++) atomicAdd(&h[p * 256 + (uint32_t)((k >> (8 * p)) & 0xff)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P * 256; i += kThreads)
        if (h[i]) atomicAdd(ghist + i, h[i]);
}

template <typename K>
__global__ __launch_bounds__(kThreads) void k_tile_hist(const K* __restrict__ keys, size_t n, int shift,
                                                        uint32_t numTiles, uint32_t* __restrict__ thist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll
    for (int c = 0; c < kChunks; c++) {
        const size_t i = base + (size_t)c * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)((keys[i] >> shift) & 0xff)], 1u);
    }
    __syncthreads();
    thist[(size_t)threadIdx.x * numTiles + blockIdx.x] = h[threadIdx.x];
}

// Block-wide exclusive scan of one value per thread over NT threads; returns
// the block total.
template <int NT = kThreads>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t& excl, uint32_t* lds /*>= NT / 64*/) {
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if ((int)__lane_id() >= off) incl += t;
    }
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 63) lds[w] = incl;
    __syncthreads();
    uint32_t waveBase = 0, total = 0;
    for (int k = 0; k < NT / 64; k++) {
        const uint32_t s = lds[k];
        if (k < w) waveBase += s;
        total += s;
    }
    __syncthreads();
    excl = waveBase + incl - v;
    return total;
}

__global__ __launch_bounds__(kThreads) void k_scan_reduce(const uint32_t* __restrict__ in, size_t n,
                                                          uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[4];
    const size_t base = (size_t)blockIdx.x * kScanSeg + (size_t)threadIdx.x * 16;
    uint32_t s = 0;
    for (int j = 0; j < 16; j++)
        if (base + j < n) s += in[base + j];
    uint32_t ex;
    const uint32_t tot = block_exclusive_scan(s, ex, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kThreads) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t nseg) {
    __shared__ uint32_t lds[4];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t c = 0; c < nseg; c += kThreads) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t v = i < nseg ? sums[i] : 0u;
        uint32_t ex;
        const uint32_t tot = block_exclusive_scan(v, ex, lds);
        const uint32_t base = carry;
        if (i < nseg) sums[i] = base + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry = base + tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kThreads) void k_scan_apply(const uint32_t* __restrict__ in, size_t n,
                                                         const uint32_t* __restrict__ sums,
                                                         uint32_t* __restrict__ out) {
    __shared__ uint32_t lds[4];
    const size_t base = (size_t)blockIdx.x * kScanSeg + (size_t)threadIdx.x * 16;
    uint32_t v[16];
    uint32_t s = 0;
    for (int j = 0; j < 16; j++) {
        v[j] = base + j < n ? in[base + j] : 0u;
        s += v[j];
    }
    uint32_t ex;
    (void)block_exclusive_scan(s, ex, lds);
    uint32_t run = sums[blockIdx.x] + ex;
    for (int j = 0; j < 16; j++)
        if (base + j < n) {
            out[base + j] = run;
            run += v[j];
        }
}

// Stable scatter of one tile of kTile keys in two steps.  Rank: wave w owns
// keys [w * kTile/4, (w + 1) * kTile/4) of the tile and walks them 64 at a
// time in index order; lanes with equal digits find each other with 8
// ballots, and a per-wave running count per digit in LDS gives each key its
// rank among the wave's earlier keys of that digit.  Stage: per-digit wave
// prefixes and the tile's digit starts (a block scan) place every key at its
// tile-sorted position in LDS.  Write: thread t writes the staged keys t,
// t + 256, ... to the digit's global offset plus their distance from the
// digit's tile start, so each digit's run leaves as consecutive addresses.
template <typename K>
__global__ __launch_bounds__(kThreads) void k_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      K* __restrict__ kout, uint32_t* __restrict__ vout, size_t n,
                                                      int shift, uint32_t numTiles,
                                                      const uint32_t* __restrict__ offsets) {
    constexpr int kWaves = kThreads / 64;
    constexpr int kPerWave = kTile / kWaves;
    __shared__ K sk[kTile];
    __shared__ uint32_t sv[kTile];
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t lstart[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t scanTmp[kWaves];
    const int tid = threadIdx.x;
    const int w = tid >> 6;
    const int lane = __lane_id();
    const size_t base = (size_t)blockIdx.x * kTile;
    const uint32_t cnt = (uint32_t)min((size_t)kTile, n - base);
    for (int j = lane; j < 256; j += 64) wcnt[w][j] = 0u;
    K kr[kChunks];
    uint32_t vr[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; c++) {
        const uint32_t li = (uint32_t)(w * kPerWave + c * 64 + lane);
        const bool valid = li < cnt;
        kr[c] = valid ? kin[base + li] : (K)0;
        vr[c] = valid ? vin[base + li] : 0u;
    }
    const uint64_t lt = lanemask_lt();
    uint32_t rk[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; c++) {
        const bool valid = (uint32_t)(w * kPerWave + c * 64 + lane) < cnt;
        const uint32_t d = (uint32_t)((kr[c] >> shift) & 0xff);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        // every lane of the digit reads the wave's count before its first
        // lane stores the new one: LDS serves one wave's accesses in order
        const uint32_t before = wcnt[w][d];
        rk[c] = before + rank;
        if (valid && rank == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < kWaves; k++) {
            const uint32_t v = wcnt[k][tid];
            wcnt[k][tid] = acc;
            acc += v;
        }
        uint32_t excl;
        (void)block_exclusive_scan(acc, excl, scanTmp);
        lstart[tid] = excl;
        gbase[tid] = offsets[(size_t)tid * numTiles + blockIdx.x] - excl;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kChunks; c++) {
        const bool valid = (uint32_t)(w * kPerWave + c * 64 + lane) < cnt;
        if (valid) {
            const uint32_t d = (uint32_t)((kr[c] >> shift) & 0xff);
            const uint32_t p = lstart[d] + wcnt[w][d] + rk[c];
            sk[p] = kr[c];
            sv[p] = vr[c];
        }
    }
    __syncthreads();
    for (uint32_t j = tid; j < cnt; j += kThreads) {
        const K k = sk[j];
        const uint32_t o = gbase[(uint32_t)((k >> shift) & 0xff)] + j;
        kout[o] = k;
        vout[o] = sv[j];
    }
}

// One-pass scatter with decoupled look-back.  Each workgroup takes the next
// tile number from a counter (so every lower tile's workgroup has started),
// ranks its tile as k_scatter does, publishes its per-digit counts
// (kAggregate), then one thread per digit walks back over the lower tiles'
// words, adding their counts until it meets a tile that has published its
// inclusive prefix (kPrefix), and publishes its own.  A tile's publication
// never waits on anything, so every walk ends; a walk that polls more than
// kLookbackPolls times sets *err and stops (the pack then fails).
constexpr uint64_t kAggregate = 1ull << 62;
constexpr uint64_t kPrefix = 2ull << 62;
constexpr uint64_t kCountMask = (1ull << 62) - 1;
constexpr int kLookbackPolls = 1 << 20;

__global__ __launch_bounds__(256) void k_digit_starts(const uint32_t* __restrict__ ghist, int P,
                                                      uint32_t* __restrict__ dstart) {
    __shared__ uint32_t lds[4];
    for (int p = 0; p < P; p++) {
        uint32_t excl;
        (void)block_exclusive_scan(ghist[p * 256 + threadIdx.x], excl, lds);
        dstart[p * 256 + threadIdx.x] = excl;
    }
}

// Shape of the look-back pass: NT threads of KPT keys each, NT * KPT keys per
// tile, the number of lower tiles whose words one look-back step reads at
// once, and the waves per SIMD the kernel is compiled for.  The tile is
// staged in tile-sorted order in LDS; 64-bit keys and their payload are
// staged one after the other in the same words (two rounds).  Measured on
// 100M-point packs: 8,192-key tiles (each tile walks back over about half as
// many lower tiles as with 4,096 keys); 30-bit keys with 512 threads (72 KB of
// LDS, two workgroups per CU), 63-bit keys with 1,024 (82 KB, one per CU).
template <typename K>
struct SweepShape;
template <>
struct SweepShape<uint32_t> {
    static constexpr int kNT = 512, kKPT = 16, kWindow = 4, kWaves = 4;
};
template <>
struct SweepShape<uint64_t> {
    static constexpr int kNT = 1024, kKPT = 8, kWindow = 4, kWaves = 4;
};
template <typename K>
constexpr int sweep_tile() { return SweepShape<K>::kNT * SweepShape<K>::kKPT; }

template <typename K, int NT, int KPT, int kWindow, int kWavesPerSimd>
__global__ __launch_bounds__(NT, kWavesPerSimd) void k_onesweep(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                 K* __restrict__ kout, uint32_t* __restrict__ vout, size_t n,
                                                 int shift, const uint32_t* __restrict__ dstart,
                                                 uint64_t* __restrict__ status, uint32_t* __restrict__ tileCounter,
                                                 uint32_t* __restrict__ err) {
    constexpr int kWaves = NT / 64;
    constexpr int kTileKeys = NT * KPT;
    constexpr int kPerWave = kTileKeys / kWaves;
    constexpr bool kTwoRounds = sizeof(K) > 4;
    static_assert(NT >= 256, "one thread per digit");
    __shared__ K sk[kTileKeys];
    __shared__ uint32_t svOwn[kTwoRounds ? 1 : kTileKeys];
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t lstart[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t scanTmp[kWaves];
    __shared__ uint32_t tileShared;
    uint32_t* sv = kTwoRounds ? reinterpret_cast<uint32_t*>(sk) : svOwn;
    const int tid = threadIdx.x;
    const int w = tid >> 6;
    const int lane = __lane_id();
    if (tid == 0) tileShared = atomicAdd(tileCounter, 1u);
    for (int j = lane; j < 256; j += 64) wcnt[w][j] = 0u;
    __syncthreads();
    const uint32_t tile = tileShared;
    const size_t base = (size_t)tile * kTileKeys;
    const uint32_t cnt = (uint32_t)min((size_t)kTileKeys, n - base);
    K kr[KPT];
    uint32_t vr[KPT];
#pragma unroll
    for (int c = 0; c < KPT; c++) {
        const uint32_t li = (uint32_t)(w * kPerWave + c * 64 + lane);
        const bool valid = li < cnt;
        kr[c] = valid ? kin[base + li] : (K)0;
        vr[c] = valid ? vin[base + li] : 0u;
    }
    // rank among the wave's earlier keys of the same digit
    const uint64_t lt = lanemask_lt();
    uint32_t rk[KPT];
#pragma unroll
    for (int c = 0; c < KPT; c++) {
        const bool valid = (uint32_t)(w * kPerWave + c * 64 + lane) < cnt;
        const uint32_t d = (uint32_t)((kr[c] >> shift) & 0xff);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        const uint32_t before = wcnt[w][d];
        rk[c] = before + rank;
        if (valid && rank == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    uint32_t acc = 0;
    uint64_t excl = 0;
    if (tid < 256) {
#pragma unroll
        for (int k = 0; k < kWaves; k++) {
            const uint32_t v = wcnt[k][tid];
            wcnt[k][tid] = acc;
            acc += v;
        }
        // publish this tile's count of digit tid, then look back for the
        // count of the digit in all lower tiles
        uint64_t* mine = status + (size_t)tile * 256 + tid;
        if (tile == 0) {
            __hip_atomic_store(mine, kPrefix | acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(mine, kAggregate | acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // the words of kWindow lower tiles are read at once (one memory
            // round trip per window), then consumed in order: counts are added
            // up to the first inclusive prefix; a word not yet published stops
            // the window and is polled again
            int64_t j = (int64_t)tile - 1;
            int polls = 0;
            bool done = false;
            while (!done) {
                uint64_t v[kWindow];
#pragma unroll
                for (int k = 0; k < kWindow; k++)
                    v[k] = j - k >= 0 ? __hip_atomic_load(status + (size_t)(j - k) * 256 + tid, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                                      : kPrefix;   // below tile 0: nothing to add
                int used = kWindow;
#pragma unroll
                for (int k = 0; k < kWindow; k++) {
                    if (done || used < kWindow) continue;
                    if ((v[k] & ~kCountMask) == 0) {
                        used = k;
                        continue;
                    }
                    excl += v[k] & kCountMask;
                    if ((v[k] & ~kCountMask) == kPrefix) done = true;
                }
                if (done) break;
                j -= used;
                if (used < kWindow) {
                    if (++polls > kLookbackPolls) {
                        atomicOr(err, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __hip_atomic_store(mine, kPrefix | (excl + acc), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    uint32_t lex;
    (void)block_exclusive_scan<NT>(acc, lex, scanTmp);
    if (tid < 256) {
        lstart[tid] = lex;
        gbase[tid] = dstart[tid] + (uint32_t)excl - lex;
    }
    __syncthreads();
    // stage the keys (and 32-bit payloads) in tile-sorted order
#pragma unroll
    for (int c = 0; c < KPT; c++) {
        const bool valid = (uint32_t)(w * kPerWave + c * 64 + lane) < cnt;
        const uint32_t d = (uint32_t)((kr[c] >> shift) & 0xff);
        rk[c] += lstart[d] + wcnt[w][d];   // the key's place in the tile
        if (valid) {
            sk[rk[c]] = kr[c];
            if (!kTwoRounds) sv[rk[c]] = vr[c];
        }
    }
    __syncthreads();
    // thread t writes staged keys t, t + NT, ...: each digit's run leaves as
    // consecutive addresses.  o < n holds whenever every look-back completed;
    // after a failed one (err set, the pack fails) no write leaves the arrays.
    uint32_t o[KPT];
#pragma unroll
    for (int c = 0; c < KPT; c++) {
        const uint32_t j = (uint32_t)(c * NT + tid);
        o[c] = 0xffffffffu;
        if (j < cnt) {
            const K k = sk[j];
            o[c] = gbase[(uint32_t)((k >> shift) & 0xff)] + j;
            if (o[c] < n) {
                kout[o[c]] = k;
                if (!kTwoRounds) vout[o[c]] = sv[j];
            }
        }
    }
    if (kTwoRounds) {
        __syncthreads();
#pragma unroll
        for (int c = 0; c < KPT; c++)
            if ((uint32_t)(w * kPerWave + c * 64 + lane) < cnt) sv[rk[c]] = vr[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < KPT; c++) {
            const uint32_t j = (uint32_t)(c * NT + tid);
            if (j < cnt && o[c] < n) vout[o[c]] = sv[j];
        }
    }
}

// ---------------------------------------------------------------- 4 gather
// The one gather from the caller's order: the points' records (position and
// RGBA8 colour) into Morton order (every later stage reads inside a batch),
// levels reset to the finest.
__global__ __launch_bounds__(kThreads) void k_gather_morton(const float4* __restrict__ rec,
                                                            const uint32_t* __restrict__ idx, size_t n,
                                                            uint8_t maxLevel, float4* __restrict__ sp,
                                                            uint8_t* __restrict__ lv) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    sp[i] = rec[idx[i]];
    lv[i] = maxLevel;
}

// ---------------------------------------------------------------- 5 LOD
struct LodGrid {
    float voxelScale;   // 0.5^level
    float coarse;
    uint32_t dx, dy, dz;
};

// A level's cell transform: the cloud's minimum and the reciprocal voxel
// size, computed once per thread from the bounds words (the same arithmetic
// for every point).
struct LodXform {
    float mn[3];
    float inv;
};
__device__ inline LodXform lod_xform(const uint32_t* bounds, const LodGrid& g) {
    LodXform t;
    float ext[3];
    for (int a = 0; a < 3; a++) {
        t.mn[a] = float_unflip(bounds[a]);
        ext[a] = fmaxf(float_unflip(bounds[3 + a]) - t.mn[a], 0.000001f);
    }
    const float longest = fmaxf(ext[0], fmaxf(ext[1], ext[2]));
    const float voxel = (longest / g.coarse) * g.voxelScale;
    t.inv = 1.0f / fmaxf(voxel, 0.000001f);
    return t;
}

__device__ inline void lod_cell(const LodXform& t, const LodGrid& g, float4 p, uint32_t& cx, uint32_t& cy,
                                uint32_t& cz) {
    const float* mn = t.mn;
    const float inv = t.inv;
    int32_t x = f2i(floorf((p.x - mn[0]) * inv));
    int32_t y = f2i(floorf((p.y - mn[1]) * inv));
    int32_t z = f2i(floorf((p.z - mn[2]) * inv));
    x = x < 0 ? 0 : x;
    y = y < 0 ? 0 : y;
    z = z < 0 ? 0 : z;
    // (p - min) * inv <= extent * inv, so the clamps never bind; they keep
    // every cell index inside the grid for any input
    cx = min((uint32_t)x, g.dx - 1u);
    cy = min((uint32_t)y, g.dy - 1u);
    cz = min((uint32_t)z, g.dz - 1u);
}

__device__ inline uint64_t lod_lin(const LodXform& t, const LodGrid& g, float4 p) {
    uint32_t cx, cy, cz;
    lod_cell(t, g, p, cx, cy, cz);
    return (uint64_t)cx + (uint64_t)g.dx * ((uint64_t)cy + (uint64_t)g.dy * cz);
}

constexpr int kLodPts = 4;   // points per thread: index blockIdx * 1024 + k * 256 + tid

// Level 0 claims: every point offers its index to its cell's atomicMin; the
// previous point of the wave (one index lower) in the same cell holds a
// smaller index, so only the first of a run of equal cells offers.
__global__ __launch_bounds__(kThreads) void k_lod_claim(const float4* __restrict__ sp, const uint8_t* __restrict__ lv,
                                                        size_t n, uint8_t maxLevel, const uint32_t* __restrict__ bounds,
                                                        LodGrid g, uint32_t* __restrict__ grid) {
    const LodXform t = lod_xform(bounds, g);
    const size_t b0 = (size_t)blockIdx.x * (kThreads * kLodPts) + threadIdx.x;
    float4 p[kLodPts];
    uint8_t l[kLodPts];
#pragma unroll
    for (int k = 0; k < kLodPts; k++) {
        const size_t i = b0 + (size_t)k * kThreads;
        l[k] = i < n ? lv[i] : (uint8_t)0xff;
        p[k] = i < n ? sp[i] : float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
#pragma unroll
    for (int k = 0; k < kLodPts; k++) {
        const size_t i = b0 + (size_t)k * kThreads;
        const bool part = i < n && l[k] == maxLevel;
        const uint64_t lin = part ? lod_lin(t, g, p[k]) : ~0ull;
        const uint64_t prev = __shfl_up(lin, 1);
        const bool dup = __lane_id() > 0 && prev == lin;
        if (part && !dup) atomicMin(grid + lin, (uint32_t)i);
    }
}

// Level `level` takes the points that won their cell in gridL (the lowest
// Morton index of the cell's unclaimed points); the rest offer their index
// to their cell of the next level in gridN.
__global__ __launch_bounds__(kThreads) void k_lod_step(const float4* __restrict__ sp, uint8_t* __restrict__ lv,
                                                       size_t n, uint8_t maxLevel, uint8_t level,
                                                       const uint32_t* __restrict__ bounds, LodGrid gl,
                                                       const uint32_t* __restrict__ gridL, LodGrid gn,
                                                       uint32_t* __restrict__ gridN) {
    const LodXform tl = lod_xform(bounds, gl), tn = lod_xform(bounds, gn);
    const size_t b0 = (size_t)blockIdx.x * (kThreads * kLodPts) + threadIdx.x;
    float4 p[kLodPts];
    uint8_t l[kLodPts];
    uint32_t won[kLodPts];
#pragma unroll
    for (int k = 0; k < kLodPts; k++) {
        const size_t i = b0 + (size_t)k * kThreads;
        l[k] = i < n ? lv[i] : (uint8_t)0xff;
        p[k] = i < n ? sp[i] : float4{0.0f, 0.0f, 0.0f, 0.0f};
    }
#pragma unroll
    for (int k = 0; k < kLodPts; k++) {
        const size_t i = b0 + (size_t)k * kThreads;
        won[k] = (i < n && l[k] == maxLevel) ? gridL[lod_lin(tl, gl, p[k])] : 0xffffffffu;
    }
#pragma unroll
    for (int k = 0; k < kLodPts; k++) {
        const size_t i = b0 + (size_t)k * kThreads;
        bool part = i < n && l[k] == maxLevel;
        if (part && won[k] == (uint32_t)i) {
            lv[i] = level;
            part = false;
        }
        const uint64_t lin = part ? lod_lin(tn, gn, p[k]) : ~0ull;
        const uint64_t prev = __shfl_up(lin, 1);
        const bool dup = __lane_id() > 0 && prev == lin;
        if (part && !dup) atomicMin(gridN + lin, (uint32_t)i);
    }
}

// ---------------------------------------------------------------- 6-9 per batch
// One workgroup per batch (Morton positions [b * ppb, b * ppb + ppb)):
//   - the last LOD level takes its cell winners (when LOD runs);
//   - level histogram → cum counts, and the batch AABB (the batch holds the
//     same points in Morton and in final order, so min/max are the same);
//   - the RasterBatch record;
//   - a stable counting scatter by level in Morton order (3-bit digit, wave
//     ballots, per-wave counts in LDS), writing each point's final slot: level,
//     three 10-bit position words quantised against the batch AABB, RGBA8
//     colour, and (when asked) its source index.
__global__ __launch_bounds__(kThreads) void k_finalize(const float4* __restrict__ sp, uint8_t* __restrict__ lv, const uint32_t* __restrict__ srcIdx,
                                                       size_t n, uint32_t ppb, uint32_t slotBase, uint8_t maxLevel,
                                                       int lastLevel, const uint32_t* __restrict__ bounds, LodGrid g,
                                                       const uint32_t* __restrict__ grid,
                                                       scr_raster_batch* __restrict__ batches,
                                                       uint8_t* __restrict__ levelsOut, uint32_t* __restrict__ colorsOut,
                                                       uint32_t* __restrict__ low, uint32_t* __restrict__ med,
                                                       uint32_t* __restrict__ high, uint32_t* __restrict__ permOut) {
    constexpr int kWaves = kThreads / 64;
    __shared__ uint32_t cnt[8];
    __shared__ uint32_t running[8];
    __shared__ uint32_t wcount[kWaves][8];
    __shared__ float smn[kWaves][3], smx[kWaves][3];
    __shared__ float bmn[3], bsz[3];
    const int tid = threadIdx.x;
    const int w = tid >> 6;
    const size_t start = (size_t)blockIdx.x * ppb;
    const size_t end = min(start + (size_t)ppb, n);
    const LodXform tf = lastLevel >= 0 ? lod_xform(bounds, g) : LodXform{};
    if (tid < 8) cnt[tid] = 0;
    __syncthreads();
    float mn[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
    float mx[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    for (size_t i = start + tid; i < end; i += kThreads) {
        const float4 p = sp[i];
        uint8_t l = lv[i];
        if (lastLevel >= 0 && l == maxLevel && grid[lod_lin(tf, g, p)] == (uint32_t)i) {
            l = (uint8_t)lastLevel;
            lv[i] = l;   // re-read below by this same thread
        }
        atomicAdd(&cnt[l & 7u], 1u);
        mn[0] = fminf(mn[0], p.x); mn[1] = fminf(mn[1], p.y); mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x); mx[1] = fmaxf(mx[1], p.y); mx[2] = fmaxf(mx[2], p.z);
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off));
        }
    if (__lane_id() == 0)
        for (int a = 0; a < 3; a++) {
            smn[w][a] = mn[a];
            smx[w][a] = mx[a];
        }
    __syncthreads();
    if (tid == 0) {
        for (int a = 0; a < 3; a++) {
            for (int k = 1; k < kWaves; k++) {
                mn[a] = fminf(mn[a], smn[k][a]);
                mx[a] = fmaxf(mx[a], smx[k][a]);
            }
            bmn[a] = mn[a];
            bsz[a] = fmaxf(mx[a] - mn[a], 0.000001f);
        }
        uint32_t acc = 0, c[8];
        for (int l = 0; l < 8; l++) {
            running[l] = acc;
            acc += cnt[l];
            c[l] = acc;
        }
        scr_raster_batch b;
        b.state = 1;
        b.minX = mn[0]; b.minY = mn[1]; b.minZ = mn[2];
        b.maxX = mx[0]; b.maxY = mx[1]; b.maxZ = mx[2];
        b.numPoints = (uint32_t)(end - start);
        b.firstPoint = (slotBase + blockIdx.x) * ppb;
        b.fileIndex = 0;
        b.lodCum01 = c[0] | (c[1] << 16);
        b.lodCum23 = c[2] | (c[3] << 16);
        b.lodCum45 = c[4] | (c[5] << 16);
        b.lodCum67 = c[6] | (c[7] << 16);
        b.padding7 = 0;
        b.padding8 = 0;
        batches[blockIdx.x] = b;
    }
    __syncthreads();
    const uint64_t lt = lanemask_lt();
    for (size_t c = start; c < end; c += kThreads) {
        const size_t i = c + tid;
        const bool valid = i < end;
        const uint8_t lvl = valid ? lv[i] : 0;
        const uint32_t d = lvl & 7u;
        if (__lane_id() < 8) wcount[w][__lane_id()] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcount[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[d] + rank;
            for (int k = 0; k < w; k++) pos += wcount[k][d];
            const size_t o = start + pos;
            // the record: position, RGBA8 colour in the w bits
            const float4 p = sp[i];
            levelsOut[o] = lvl;
            colorsOut[o] = __float_as_uint(p.w);
            if (permOut) permOut[o] = srcIdx[i];
            // quantize (p - min) / size, IEEE division, as the CPU pack
            const float v[3] = {p.x, p.y, p.z};
            uint32_t q[3];
            for (int a = 0; a < 3; a++) {
                const float nrm = fminf(fmaxf((v[a] - bmn[a]) / bsz[a], 0.0f), 0.99999994f);
                q[a] = f2u(nrm * 1073741824.0f);  // Float(2^30 - 1) rounds to 2^30
            }
            low[o] = ((q[0] >> 20) & 1023u) | (((q[1] >> 20) & 1023u) << 10) | (((q[2] >> 20) & 1023u) << 20);
            med[o] = ((q[0] >> 10) & 1023u) | (((q[1] >> 10) & 1023u) << 10) | (((q[2] >> 10) & 1023u) << 20);
            high[o] = (q[0] & 1023u) | ((q[1] & 1023u) << 10) | ((q[2] & 1023u) << 20);
        }
        __syncthreads();
        if (tid < 8) {
            uint32_t add = 0;
            for (int k = 0; k < kWaves; k++) add += wcount[k][tid];
            running[tid] += add;
        }
        __syncthreads();
    }
}

unsigned blocks(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

// Bump allocator over the context scratch block.
struct Arena {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = (T*)(base + off);
        off += count * sizeof(T);
        return p;
    }
};

}  // namespace


// Scratch bytes of a pack of n points with LOD grids of grid0 and grid1
// cells (the six bounds words are not included).
static size_t pack_scratch_bytes(size_t n, bool wide, size_t grid0, size_t grid1) {
    const size_t numTiles = (n + kTile - 1) / kTile;
    const size_t scanN = 256 * numTiles;
    const size_t k = wide ? 8 : 4;
    size_t b = 0;
    b += 2 * n * k + 2 * n * 4;                 // keys, payload ping-pong
    b += 2 * n * 16;                            // point records in caller and in Morton order
    b += n;                                     // Morton-order levels
    b += 2 * scanN * 4 + (scanN / kScanSeg + 2) * 4;  // tile hist, offsets, scan sums
    const size_t sweepTile = (size_t)std::min(sweep_tile<uint32_t>(), sweep_tile<uint64_t>());
    b += std::max(scanN, 256 * ((n + sweepTile - 1) / sweepTile)) * 8;   // look-back words
    b += 8 * 256 * 4 + 16 * 4;                  // digit starts, counters
    b += 8 * 256 * 4 + 64;                      // digit histograms
    b += (grid0 + grid1) * 4;
    b += 32 * 256;                              // alignment slack
    return b;
}

// The dense grid of one LOD level: the device cell index's arithmetic on the
// bounds words hb (float_flip'ed min xyz, max xyz).
static LodGrid lod_grid(const uint32_t hb[6], int coarse, int level) {
    LodGrid g;
    g.voxelScale = std::pow(0.5f, (float)level);
    g.coarse = (float)coarse;
    float mn[3], ext[3];
    for (int a = 0; a < 3; a++) {
        mn[a] = float_unflip(hb[a]);
        ext[a] = std::fmax(float_unflip(hb[3 + a]) - mn[a], 0.000001f);
    }
    const float longest = std::fmax(ext[0], std::fmax(ext[1], ext[2]));
    const float voxel = (longest / g.coarse) * g.voxelScale;
    const float inv = 1.0f / std::fmax(voxel, 0.000001f);
    g.dx = (uint32_t)f2u(std::floor(ext[0] * inv)) + 1u;
    g.dy = (uint32_t)f2u(std::floor(ext[1] * inv)) + 1u;
    g.dz = (uint32_t)f2u(std::floor(ext[2] * inv)) + 1u;
    return g;
}

static size_t lod_cells(const LodGrid& g) { return (size_t)g.dx * g.dy * g.dz; }

// Working memory of one sort.
struct SortScratch {
    uint32_t* ghist;    // all-digit histogram, 8 x 256
    uint32_t* thist;    // tile histograms, digit-major (split passes)
    uint32_t* offs;     // their exclusive scan
    uint32_t* sums;     // scan block sums
    uint64_t* status;   // look-back words, one per (tile, digit)
    uint32_t* dstart;   // per pass exclusive digit starts
    uint32_t* counters; // per pass tile counters [0, 8), look-back failure flag [8]
};

// Inputs of at least this many points sort with the one-kernel look-back
// pass; smaller ones with the split histogram / scan / scatter pass
// (measured faster there: 30-bit keys below 2^24 points, 63-bit keys below
// 2^20).
template <typename K>
constexpr size_t lookback_min_points() { return sizeof(K) == 8 ? size_t(1) << 20 : size_t(1) << 24; }

template <typename K>
static int radix_sort(scr_ctx* ctx, K*& kin, uint32_t*& vin, K*& kout, uint32_t*& vout, size_t n,
                      const SortScratch& w, int* passesRun) {
    hipStream_t s = ctx->stream;
    constexpr int P = sizeof(K);
    const bool lookback = n >= lookback_min_points<K>();
    const uint32_t numTiles = (uint32_t)((n + kTile - 1) / kTile);
    const size_t scanN = (size_t)256 * numTiles;
    const uint32_t nseg = (uint32_t)((scanN + kScanSeg - 1) / kScanSeg);
    constexpr int kSweepNT = SweepShape<K>::kNT, kSweepKPT = SweepShape<K>::kKPT, kSweepWin = SweepShape<K>::kWindow,
                  kSweepWaves = SweepShape<K>::kWaves;
    const uint32_t sweepTiles = (uint32_t)((n + sweep_tile<K>() - 1) / sweep_tile<K>());
    hipError_t e = launch_fill_u32(s, w.ghist, P * 256, 0u);
    if (e == hipSuccess && lookback) e = launch_fill_u32(s, w.counters, 16, 0u);
    if (e != hipSuccess) return hip_error(ctx, e, "radix hist clear");
    hipLaunchKernelGGL(k_all_digit_hist<K>, dim3(std::min<unsigned>(blocks(n, kThreads), (unsigned)ctx->numCUs * 4)),
                       dim3(kThreads), 0, s, kin, n, w.ghist);
    if (lookback) hipLaunchKernelGGL(k_digit_starts, dim3(1), dim3(256), 0, s, w.ghist, P, w.dstart);
    uint32_t h[8 * 256];
    e = hipMemcpyAsync(h, w.ghist, (size_t)P * 256 * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, "radix hist");
    int run = 0;
    for (int p = 0; p < P; p++) {
        bool trivial = false;
        for (int d = 0; d < 256; d++)
            if (h[(size_t)p * 256 + d] == n) trivial = true;
        if (trivial) continue;  // every key has the same digit: the pass is the identity
        const int shift = 8 * p;
        if (lookback) {
            e = launch_fill_u32(s, (uint32_t*)w.status, (size_t)sweepTiles * 256 * 2, 0u);
            if (e != hipSuccess) return hip_error(ctx, e, "radix look-back clear");
            hipLaunchKernelGGL((k_onesweep<K, kSweepNT, kSweepKPT, kSweepWin, kSweepWaves>), dim3(sweepTiles), dim3(kSweepNT), 0, s, kin, vin,
                               kout, vout, n, shift, w.dstart + p * 256, w.status, w.counters + p, w.counters + 8);
        } else {
            hipLaunchKernelGGL(k_tile_hist<K>, dim3(numTiles), dim3(kThreads), 0, s, kin, n, shift, numTiles, w.thist);
            hipLaunchKernelGGL(k_scan_reduce, dim3(nseg), dim3(kThreads), 0, s, w.thist, scanN, w.sums);
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, s, w.sums, nseg);
            hipLaunchKernelGGL(k_scan_apply, dim3(nseg), dim3(kThreads), 0, s, w.thist, scanN, w.sums, w.offs);
            hipLaunchKernelGGL(k_scatter<K>, dim3(numTiles), dim3(kThreads), 0, s, kin, vin, kout, vout, n, shift,
                               numTiles, w.offs);
        }
        std::swap(kin, kout);
        std::swap(vin, vout);
        run++;
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_error(ctx, e, "radix sort");
    if (lookback) {
        uint32_t flag = 0;
        e = hipMemcpyAsync(&flag, w.counters + 8, 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_error(ctx, e, "radix sort");
        if (flag) return set_error(ctx, SCR_ERR_HIP, "radix sort: a tile's look-back did not complete");
    }
    if (passesRun) *passesRun = run;
    return SCR_OK;
}

// GPUPacker.init's LOD gate (GPUPacker.swift:35, 86-99): the dense grid has at
// most (coarse * 2^(L - 2) + 2)^3 cells (cells per axis floor(extent /
// voxel) + 1 at the finest claimed level); above 268,435,456 the packer
// skips LOD and leaves every level at the finest.
bool gpu_lod_enabled(int lodLevels, int coarse) {
    if (lodLevels <= 1) return false;
    const long long maxDim = (long long)coarse * (1ll << (lodLevels - 2)) + 2;
    return maxDim <= 1000000 && maxDim * maxDim * maxDim <= 268435456ll;
}

// Run the nine stages for n points into `dst`.  pos/col are device float4
// arrays; the stages work in the context's scratch, grown on demand.
int pack_points(scr_ctx* ctx, const float4* pos, const float4* col, size_t n, const scr_pack_options& opt,
                const PackDest& dst, uint32_t* hostSourceIndices) {
    hipStream_t s = ctx->stream;
    const bool wide = opt.keyBits == 63;
    const size_t ppb = dst.ppb;
    const size_t nb = (n + ppb - 1) / ppb;
    const int lodLevels = std::min(std::max(opt.lodLevels, 1), 8);
    const int coarse = std::max(opt.coarseVoxelDivisions, 1);
    const uint8_t maxLevel = (uint8_t)(lodLevels - 1);
    const bool lodOn = gpu_lod_enabled(lodLevels, coarse);
    hipError_t e = hipSuccess;

    // bounds first (the LOD grid size depends on them), in the context's
    // bounds words (allocated with the context)
    uint32_t* dBounds = ctx->packBounds;
    {
        PassScope ps(ctx, "pack_bounds");
        hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, dBounds);
        hipLaunchKernelGGL(k_bounds, dim3(std::min<unsigned>(blocks(n, kThreads), (unsigned)ctx->numCUs * 8)),
                           dim3(kThreads), 0, s, pos, n, dBounds);
    }
    uint32_t hb[6];
    e = hipMemcpyAsync(hb, dBounds, sizeof hb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, "pack bounds");
    // dense LOD grids: level l claims in grid l % 2 (the finest claimed level
    // has the most cells; the gate bounds it)
    size_t gridCells[2] = {0, 0};
    if (lodOn)
        for (int level = 0; level < lodLevels - 1; level++)
            gridCells[level & 1] = std::max(gridCells[level & 1], lod_cells(lod_grid(hb, coarse, level)));
    const size_t need = pack_scratch_bytes(n, wide, gridCells[0], gridCells[1]);
    if (ctx->scratchBytes < need) {
        if (ctx->scratch) (void)hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratchBytes = 0;
        e = hipMalloc(&ctx->scratch, need);
        if (e != hipSuccess) return hip_error(ctx, e, "pack scratch alloc");
        ctx->scratchBytes = need;
    }
    Arena ar{(char*)ctx->scratch};
    const size_t numTiles = (n + kTile - 1) / kTile;
    const size_t scanN = 256 * numTiles;
    void* keysA = ar.take<char>(n * (wide ? 8 : 4));
    void* keysB = ar.take<char>(n * (wide ? 8 : 4));
    uint32_t* idxA = ar.take<uint32_t>(n);
    uint32_t* idxB = ar.take<uint32_t>(n);
    float4* rec = ar.take<float4>(n);
    float4* sp = ar.take<float4>(n);
    uint8_t* lvM = ar.take<uint8_t>(n);
    SortScratch sw;
    sw.thist = ar.take<uint32_t>(scanN);
    sw.offs = ar.take<uint32_t>(scanN);
    sw.sums = ar.take<uint32_t>(scanN / kScanSeg + 2);
    sw.ghist = ar.take<uint32_t>(8 * 256);
    {
        const size_t sweepTile = (size_t)std::min(sweep_tile<uint32_t>(), sweep_tile<uint64_t>());
        sw.status = ar.take<uint64_t>(std::max(scanN, 256 * ((n + sweepTile - 1) / sweepTile)));
    }
    sw.dstart = ar.take<uint32_t>(8 * 256);
    sw.counters = ar.take<uint32_t>(16);
    uint32_t* grid[2] = {ar.take<uint32_t>(gridCells[0]), ar.take<uint32_t>(gridCells[1])};

    // 2 Morton keys, 3 sort
    uint32_t* sortedIdx = nullptr;
    int rc = SCR_OK;
    if (wide) {
        {
            PassScope ps(ctx, "pack_morton");
            hipLaunchKernelGGL(k_morton<uint64_t>, dim3(blocks(n, kThreads)), dim3(kThreads), 0, s, pos, col, n,
                               dBounds, (uint64_t*)keysA, idxA, rec);
        }
        uint64_t *kin = (uint64_t*)keysA, *kout = (uint64_t*)keysB;
        uint32_t *vin = idxA, *vout = idxB;
        PassScope ps(ctx, "pack_sort");
        rc = radix_sort<uint64_t>(ctx, kin, vin, kout, vout, n, sw, nullptr);
        sortedIdx = vin;
    } else {
        {
            PassScope ps(ctx, "pack_morton");
            hipLaunchKernelGGL(k_morton<uint32_t>, dim3(blocks(n, kThreads)), dim3(kThreads), 0, s, pos, col, n,
                               dBounds, (uint32_t*)keysA, idxA, rec);
        }
        uint32_t *kin = (uint32_t*)keysA, *kout = (uint32_t*)keysB;
        uint32_t *vin = idxA, *vout = idxB;
        PassScope ps(ctx, "pack_sort");
        rc = radix_sort<uint32_t>(ctx, kin, vin, kout, vout, n, sw, nullptr);
        sortedIdx = vin;
    }
    if (rc != SCR_OK) return rc;
    uint32_t* permOut = sortedIdx == idxA ? idxB : idxA;

    // 4 positions and colours into Morton order
    {
        PassScope ps(ctx, "pack_gather");
        hipLaunchKernelGGL(k_gather_morton, dim3(blocks(n, kThreads)), dim3(kThreads), 0, s, rec, sortedIdx, n,
                           maxLevel, sp, lvM);
    }
    // 5 LOD levels 0 .. L-3 (the last claimed level, L-2, is taken inside the finalize)
    int lastLevel = -1;
    LodGrid lastGrid{};
    if (lodOn) {
        PassScope ps(ctx, "pack_lod");
        LodGrid g = lod_grid(hb, coarse, 0);
        e = launch_fill_u32(s, grid[0], lod_cells(g), 0xffffffffu);
        if (e != hipSuccess) return hip_error(ctx, e, "pack lod clear");
        hipLaunchKernelGGL(k_lod_claim, dim3(blocks(n, kThreads * kLodPts)), dim3(kThreads), 0, s, sp, lvM, n, maxLevel, dBounds,
                           g, grid[0]);
        for (int level = 0; level + 1 < lodLevels - 1; level++) {
            const LodGrid gn = lod_grid(hb, coarse, level + 1);
            e = launch_fill_u32(s, grid[(level + 1) & 1], lod_cells(gn), 0xffffffffu);
            if (e != hipSuccess) return hip_error(ctx, e, "pack lod clear");
            hipLaunchKernelGGL(k_lod_step, dim3(blocks(n, kThreads * kLodPts)), dim3(kThreads), 0, s, sp, lvM, n, maxLevel,
                               (uint8_t)level, dBounds, g, grid[level & 1], gn, grid[(level + 1) & 1]);
            g = gn;
        }
        lastLevel = lodLevels - 2;
        lastGrid = g;
    }
    // 6-9 per batch: last level, cum counts, AABB, record, bucketed final writes
    {
        PassScope ps(ctx, "pack_finalize");
        hipLaunchKernelGGL(k_finalize, dim3((unsigned)nb), dim3(kThreads), 0, s, sp, lvM, sortedIdx, n,
                           (uint32_t)ppb, dst.slotBase, maxLevel, lastLevel, dBounds, lastGrid,
                           lastLevel >= 0 ? grid[lastLevel & 1] : nullptr, dst.batches, dst.levels, dst.colors,
                           dst.low, dst.med, dst.high, hostSourceIndices ? permOut : nullptr);
    }
    e = hipGetLastError();
    if (e == hipSuccess && hostSourceIndices)
        e = hipMemcpyAsync(hostSourceIndices, permOut, n * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_error(ctx, e, "pack");
    return SCR_OK;
}

}  // namespace scr
