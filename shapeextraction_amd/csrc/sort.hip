// Stable LSD radix sort of (uint64 key, uint32 value) pairs, 8-bit digits.
//
// Replaces cub::DeviceRadixSort::SortKeys (src/RadixTree.cu:316-335).  The
// reference sorts keys only; this sort also carries the input index so that
// the sort permutation (and with it the duplicate representatives and the
// input-order scatter of results) is available.
//
// Three schedules share the tile ranking code:
//
// Hybrid (default for keys of seven or eight digits, e.g. 63-bit Morton
// codes): the four most significant digits by onesweep passes, then every
// run of equal top bits (a bucket: a Morton cell of 1/1024 of the cube's
// edge for 63-bit codes) sorted on its low bits locally - in LDS by the
// workgroup that owns its start when it is small, by one wave or one
// workgroup when larger (see the hybrid section below).
//
// Onesweep: one histogram kernel counts every digit of every pass
// in a single read of the keys; each pass is then ONE kernel in which tiles
// take tickets in launch order, publish their digit counts, and look back
// over the preceding tiles' published counts (decoupled look-back: one
// 32-bit word per (pass, tile, digit) holding a 2-bit state and a 30-bit
// count) to find their global output offsets.  The keys are read twice and
// written once per pass; the scatter is staged through LDS so that each
// digit's run leaves the block as contiguous stores.
//
// Three-kernel fallback (SX_SORT_THREE_KERNEL), per digit pass:
//   1. radix_hist_kernel     per-tile 256-bin digit histogram (LDS atomics),
//                            stored digit-major: hist[d * tiles + t]
//   2. device_scan           exclusive scan of that array = global output
//                            offset of (digit d, tile t) for a stable sort
//   3. radix_scatter_kernel  stable rank of every key inside its tile (wave
//                            ballot match on the digit + per-wave running
//                            digit counters in LDS), then scatter.
// A tile is 4096 keys: 4 waves x 16 rounds x 64 lanes, each wave owning a
// contiguous 1024-key span so that ranks follow input order.
#include <hip/hip_runtime.h>

#include <atomic>

#include "sx_internal.hpp"

namespace sx {


constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / kWave;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kWaveSpan = kSortItems * kWave;
constexpr int kRadix = 256;

__global__ void __launch_bounds__(kSortBlock) radix_hist_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                                int shift, uint32_t* __restrict__ hist, int n_tiles) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const int64_t i = base + j * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[static_cast<uint32_t>(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[static_cast<int64_t>(threadIdx.x) * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// Rank of this lane's digit among the wave's: *below = lanes under this one
// with this lane's digit, *total = lanes with it.  Built from the complement
// of the match set (one ballot per digit bit): with
// s = -1 where bit b of this lane's digit is set (0 otherwise), s ^ ballot(bit b)
// marks the lanes whose bit b differs, and OR over the bits marks every lane
// whose digit differs - two 32-bit XORs and ORs per bit instead of a 64-bit
// select and AND.
__device__ __forceinline__ void rank_digit(uint32_t d, uint32_t& below, uint32_t& total) {
    uint32_t nlo = 0u, nhi = 0u;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const int32_t sb = static_cast<int32_t>(d << (31 - b)) >> 31;
        const uint64_t bal = __ballot(sb != 0);
        nlo |= static_cast<uint32_t>(sb) ^ static_cast<uint32_t>(bal);
        nhi |= static_cast<uint32_t>(sb) ^ static_cast<uint32_t>(bal >> 32);
    }
    below = __builtin_amdgcn_mbcnt_hi(~nhi, __builtin_amdgcn_mbcnt_lo(~nlo, 0u));
    total = static_cast<uint32_t>(__popc(~nlo) + __popc(~nhi));
}

__global__ void __launch_bounds__(kSortBlock) radix_scatter_kernel(const uint64_t* __restrict__ keys_in,
                                                                   const uint32_t* __restrict__ vals_in, int64_t n,
                                                                   int shift, const uint32_t* __restrict__ offsets,
                                                                   int n_tiles, uint64_t* __restrict__ keys_out,
                                                                   uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t cnt[kSortWaves][kRadix];
    __shared__ uint32_t tile_base[kRadix];
    const int tid = static_cast<int>(threadIdx.x);
    const int w = tid / kWave;
    const unsigned lane = lane_id();
    tile_base[tid] = offsets[static_cast<int64_t>(tid) * n_tiles + blockIdx.x];
    for (int d = static_cast<int>(lane); d < kRadix; d += kWave) cnt[w][d] = 0;
    __syncthreads();

    const int64_t span = static_cast<int64_t>(blockIdx.x) * kSortTile + static_cast<int64_t>(w) * kWaveSpan;
    uint64_t k[kSortItems];
    uint32_t v[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = span + r * kWave + lane;
        const bool valid = i < n;
        k[r] = valid ? keys_in[i] : ~0ull;
        v[r] = valid ? (vals_in ? vals_in[i] : static_cast<uint32_t>(i)) : 0u;
        const uint32_t d = static_cast<uint32_t>(k[r] >> shift) & (kRadix - 1);
        uint32_t below, total;
        rank_digit(d, below, total);
        const uint32_t c = cnt[w][d];
        rank[r] = c + below;
        if (below + 1u == total) cnt[w][d] = c + total;
    }
    __syncthreads();
    {
        uint32_t run = 0;
#pragma unroll
        for (int wv = 0; wv < kSortWaves; ++wv) {
            const uint32_t c = cnt[wv][tid];
            cnt[wv][tid] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t i = span + r * kWave + lane;
        if (i < n) {
            const uint32_t d = static_cast<uint32_t>(k[r] >> shift) & (kRadix - 1);
            const uint32_t pos = tile_base[d] + cnt[w][d] + rank[r];
            keys_out[pos] = k[r];
            vals_out[pos] = v[r];
        }
    }
}

struct HistIn {
    const uint32_t* h;
    __device__ int operator()(int64_t i) const { return static_cast<int>(h[i]); }
};
struct HistOut {
    uint32_t* h;
    __device__ void operator()(int64_t i, int ex, int) const { h[i] = static_cast<uint32_t>(ex); }
};

inline int64_t sort_tiles(int64_t n) { return (n + kSortTile - 1) / kSortTile; }


// ---------------------------------------------------------------------------
// Onesweep (Adinets & Merrill, "Onesweep: A Faster Least Significant Digit
// Radix Sort for GPUs", 2022), written for wave64:
//   onesweep_hist_kernel   one pass over the keys: histograms of every digit
//                          position at once; onesweep_base_kernel scans them
//   onesweep_pass_kernel   per digit position, one kernel: tiles take a
//                          ticket (so every tile a tile waits on is already
//                          running), rank their keys (wave ballots), publish
//                          per-digit counts, look back over earlier tiles for
//                          the exclusive prefix, stage the tile in LDS in digit
//                          order and write it out in coalesced runs.
// Look-back words pack a 2-bit flag (1 = tile count, 2 = inclusive prefix)
// with a 30-bit count, so one agent-scope atomic load/store carries both and
// no fence is needed.  Requires n < 2^30.
// ---------------------------------------------------------------------------
constexpr uint32_t kFlagAgg = 1u << 30, kFlagPrefix = 2u << 30, kCountMask = (1u << 30) - 1;
constexpr int kMaxPasses = 8;
constexpr uint32_t kDefaultSpinLimit = 1u << 22;

// A block takes chunks of kHistItems keys per thread: the chunk's loads are
// issued together, then counted, so a wave waits for one memory round trip
// per chunk rather than one per key.
constexpr int kHistItems = 16;

// Histograms of digits [first, first + passes), into ghist[p * kRadix + d]
// for p = 0 .. passes-1.
__global__ void __launch_bounds__(kSortBlock) onesweep_hist_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                                   int first, int passes,
                                                                   uint32_t* __restrict__ ghist) {
    __shared__ uint32_t h[kMaxPasses][kRadix];
    for (int i = threadIdx.x; i < kMaxPasses * kRadix; i += kSortBlock) (&h[0][0])[i] = 0;
    __syncthreads();
    constexpr int64_t kChunk = static_cast<int64_t>(kSortBlock) * kHistItems;
    for (int64_t c0 = static_cast<int64_t>(blockIdx.x) * kChunk; c0 < n; c0 += static_cast<int64_t>(gridDim.x) * kChunk) {
        uint64_t k[kHistItems];
#pragma unroll
        for (int j = 0; j < kHistItems; ++j) {
            const int64_t i = c0 + j * kSortBlock + threadIdx.x;
            k[j] = i < n ? keys[i] : 0ull;
        }
#pragma unroll
        for (int j = 0; j < kHistItems; ++j) {
            if (c0 + j * kSortBlock + threadIdx.x >= n) continue;
            for (int p = 0; p < passes; ++p)
                atomicAdd(&h[p][static_cast<uint32_t>(k[j] >> (8 * (first + p))) & (kRadix - 1)], 1u);
        }
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p) {
        const uint32_t c = h[p][threadIdx.x];
        if (c) atomicAdd(&ghist[p * kRadix + threadIdx.x], c);
    }
}

// One block per pass: exclusive scan of the 256 digit counts in place.
__global__ void __launch_bounds__(kRadix) onesweep_base_kernel(uint32_t* ghist) {
    uint32_t* h = ghist + blockIdx.x * kRadix;
    const int v = static_cast<int>(h[threadIdx.x]);
    int total;
    const int ex = block_excl_sum(v, &total);
    h[threadIdx.x] = static_cast<uint32_t>(ex);
}

// Pass tiles: 4096 keys over 512 threads (8 items each) - the same tile and
// LDS as 256 x 16, with twice the waves per CU to hide the ranking's LDS
// round trips and the look-back.
constexpr int kPassBlock = 512;
constexpr int kPassWaves = kPassBlock / kWave;
constexpr int kPassItems = kSortTile / kPassBlock;
constexpr int kPassSpan = kPassItems * kWave;  // contiguous keys per wave
constexpr int kLookWin = 4;                    // predecessor tiles read per look-back round trip
static_assert(kPassBlock >= kRadix, "one thread per digit");

__global__ void __launch_bounds__(kPassBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) onesweep_pass_kernel(const uint64_t* __restrict__ keys_in,
                                                                   const uint32_t* __restrict__ vals_in, int64_t n,
                                                                   int shift, const uint32_t* __restrict__ gbase,
                                                                   uint32_t* status, uint32_t* tile_counter,
                                                                   uint32_t spin_limit, int32_t* err,
                                                                   uint64_t* __restrict__ keys_out,
                                                                   uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t cnt[kPassWaves][kRadix];
    __shared__ uint32_t tile_start[kRadix];
    __shared__ uint32_t dest_base[kRadix];
    __shared__ int digit_wsum[kRadix / kWave];
    // the tile in digit order: keys first, then (same buffer) values
    __shared__ uint64_t s_buf[kSortTile];
    __shared__ uint8_t s_digit[kSortTile];
    __shared__ uint32_t s_tile;
    const int tid = static_cast<int>(threadIdx.x);
    const int w = tid / kWave;
    const unsigned lane = lane_id();
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int d = static_cast<int>(lane); d < kRadix; d += kWave) cnt[w][d] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const int64_t tile_base = static_cast<int64_t>(tile) * kSortTile;
    const int64_t span = tile_base + static_cast<int64_t>(w) * kPassSpan;
    uint64_t k[kPassItems];
    uint32_t v[kPassItems];
    uint32_t rank[kPassItems];
#pragma unroll
    for (int r = 0; r < kPassItems; ++r) {
        const int64_t i = span + r * kWave + lane;
        const bool valid = i < n;
        k[r] = valid ? keys_in[i] : ~0ull;
        v[r] = valid ? (vals_in ? vals_in[i] : static_cast<uint32_t>(i)) : 0u;
    }
#pragma unroll
    for (int r = 0; r < kPassItems; ++r) {
        const uint32_t d = static_cast<uint32_t>(k[r] >> shift) & (kRadix - 1);
        uint32_t below, total;
        rank_digit(d, below, total);
        const uint32_t c = cnt[w][d];
        rank[r] = c + below;
        if (below + 1u == total) cnt[w][d] = c + total;
    }
    __syncthreads();
    // per digit (thread tid < 256 = digit): wave offsets and the tile's count,
    // published at once so that later tiles can look back past this one
    const bool digit_thread = tid < kRadix;
    uint32_t run = 0, excl = 0;
    uint32_t* my = status + static_cast<int64_t>(tile) * kRadix + tid;
    if (digit_thread) {
#pragma unroll
        for (int wv = 0; wv < kPassWaves; ++wv) {
            const uint32_t c = cnt[wv][tid];
            cnt[wv][tid] = run;
            run += c;
        }
        __hip_atomic_store(my, (tile == 0 ? kFlagPrefix : kFlagAgg) | run, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    // exclusive scan of the tile's digit counts (waves 0-3 hold the digits)
    const int incl = digit_thread ? wave_incl_sum(static_cast<int>(run)) : 0;
    if (digit_thread && lane == kWave - 1) digit_wsum[w] = incl;
    __syncthreads();
    uint32_t start = 0;
    if (digit_thread) {
        int before = 0;
#pragma unroll
        for (int i = 0; i < kRadix / kWave; ++i) before += (i < w) ? digit_wsum[i] : 0;
        start = static_cast<uint32_t>(before + incl) - run;
        tile_start[tid] = start;
    }
    __syncthreads();
    // the tile in digit order in LDS (needs only tile-local offsets)
#pragma unroll
    for (int r = 0; r < kPassItems; ++r) {
        const uint32_t d = static_cast<uint32_t>(k[r] >> shift) & (kRadix - 1);
        const uint32_t pos = tile_start[d] + cnt[w][d] + rank[r];
        s_buf[pos] = k[r];
        rank[r] = pos;
    }
    // look-back: kLookWin earlier tiles per round trip (independent loads),
    // nearest first; t is the nearest tile not yet added
    if (digit_thread && tile != 0) {
        int64_t t = static_cast<int64_t>(tile) - 1;
        uint32_t spins = 0;
        bool found = false;
        while (!found) {
            uint32_t st[kLookWin];
#pragma unroll
            for (int lw = 0; lw < kLookWin; ++lw)
                st[lw] = t - lw >= 0 ? __hip_atomic_load(status + (t - lw) * kRadix + tid, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)
                                     : kFlagPrefix;  // before tile 0: an exclusive prefix of 0
            int used = 0;
            bool waiting = false;
#pragma unroll
            for (int lw = 0; lw < kLookWin; ++lw) {
                if (found || waiting) continue;
                if ((st[lw] & ~kCountMask) == 0) {
                    waiting = true;  // not published yet: reload from here
                    continue;
                }
                excl += st[lw] & kCountMask;
                ++used;
                if ((st[lw] & ~kCountMask) == kFlagPrefix) found = true;
            }
            t -= used;
            if (waiting) {
                // Tile t holds an earlier ticket, so it is running and will
                // publish; the bound only guards against a broken invariant.
                // Giving up leaves this tile's prefix short (its keys land
                // inside [0, n) but collide with others'): *err is raised so
                // that the caller's status reports the failed sort.
                if (++spins > spin_limit) {
                    atomicOr(err, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __hip_atomic_store(my, kFlagPrefix | (excl + run), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (digit_thread) dest_base[tid] = gbase[tid] + excl - start;  // output index = dest_base[d] + LDS slot
    __syncthreads();
    const int64_t valid = n - tile_base < kSortTile ? n - tile_base : kSortTile;
#pragma unroll
    for (int r = 0; r < kPassItems; ++r) {
        const int j = r * kPassBlock + tid;
        if (j < valid) {
            const uint64_t key = s_buf[j];
            const uint32_t d = static_cast<uint32_t>(key >> shift) & (kRadix - 1);
            keys_out[dest_base[d] + static_cast<uint32_t>(j)] = key;
            s_digit[j] = static_cast<uint8_t>(d);
        }
    }
    __syncthreads();
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_buf);
#pragma unroll
    for (int r = 0; r < kPassItems; ++r) s_vals[rank[r]] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPassItems; ++r) {
        const int j = r * kPassBlock + tid;
        if (j < valid) vals_out[dest_base[s_digit[j]] + static_cast<uint32_t>(j)] = s_vals[j];
    }
}

// ---------------------------------------------------------------------------
// Hybrid schedule for wide keys (seven or eight digit passes, e.g. 63-bit
// Morton codes):
//   1. the kTopPasses most significant digits by onesweep passes: the pairs
//      are then ordered by their top bits, ties in input order.  A run of
//      equal top bits is a bucket; what is left is to sort every bucket on
//      its low bits, stably;
//   2. bucket_local_kernel: each block owns the buckets that start in its
//      kLocTile positions and stages the tile with kLocSmall positions on
//      either side in LDS.  Two block scans give every position its bucket's
//      bounds; an element of an owned bucket of at most kLocSmall keys moves
//      to the bucket's start plus its rank (keys below it, or equal and
//      earlier).  The head of a larger owned bucket finds the bucket's end by
//      a search of the keys in HBM and lists it with its tile;
//   3. the larger buckets' elements are flagged, gathered in order, sorted by
//      onesweep over every digit with a device-side count (persistent blocks;
//      an empty set costs empty launches) and scattered back to the flagged
//      positions, which hold exactly those buckets in the same order.
// The result equals the LSD sort's, stable by key.  On the 8M-point scene
// every bucket of equal top 31 bits holds at most 10 codes, so the four
// low-digit passes become one local pass.
// ---------------------------------------------------------------------------
constexpr int kTopPasses = 4;
constexpr int kLocTile = 2048;
constexpr int kLocSmall = 64;  // largest bucket ranked in the local pass (also the window's halo)
constexpr int kLocBlock = 256;
constexpr int kLocWin = kLocTile + 2 * kLocSmall + 1;
constexpr int kLocPer = (kLocWin + kLocBlock - 1) / kLocBlock;  // window positions per thread in the scans
constexpr int kLocWaves = kLocBlock / kWave;
constexpr int16_t kNoPos = 0x7FFF;  // bucket start or end outside the window
static_assert(kLocWin < kNoPos, "window indices fit int16");
constexpr int kMidMax = 256;  // largest bucket sorted by one wave (rank by counting)
constexpr int kASlots = kLocTile / (kLocSmall + 1) + 1;  // such buckets that can start in one tile
constexpr int kBSlots = kLocTile / (kMidMax + 1) + 1;    // larger ones that can start in one tile

// Block-wide exclusive max-scan (identity -1) and exclusive suffix min-scan
// (identity kNoPos) of one int per thread, threads in order.
__device__ __forceinline__ int block_excl_max(int v, int* s_w) {
    const int lane = static_cast<int>(lane_id()), w = static_cast<int>(threadIdx.x) / kWave;
    int incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(incl, o, kWave);
        if (lane >= o) incl = max(incl, t);
    }
    if (lane == kWave - 1) s_w[w] = incl;
    __syncthreads();
    int ex = __shfl_up(incl, 1, kWave);
    if (lane == 0) ex = -1;
    for (int i = 0; i < w; ++i) ex = max(ex, s_w[i]);
    __syncthreads();
    return ex;
}
__device__ __forceinline__ int block_excl_suffix_min(int v, int* s_w) {
    const int lane = static_cast<int>(lane_id()), w = static_cast<int>(threadIdx.x) / kWave;
    int incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_down(incl, o, kWave);
        if (lane + o < kWave) incl = min(incl, t);
    }
    if (lane == 0) s_w[w] = incl;
    __syncthreads();
    int ex = __shfl_down(incl, 1, kWave);
    if (lane == kWave - 1) ex = kNoPos;
    for (int i = w + 1; i < kLocWaves; ++i) ex = min(ex, s_w[i]);
    __syncthreads();
    return ex;
}

// First position in (last, n) whose key's top bits differ from top, given
// that keys[last] has them (keys ordered by top bits): exponential, then
// binary search.
__device__ __forceinline__ int64_t bucket_end(const uint64_t* __restrict__ keys, int64_t n, int64_t last,
                                              uint64_t top, int lo_bits) {
    int64_t step = kLocSmall, hi = n;
    while (true) {
        const int64_t q = last + step;
        if (q >= n) break;
        if ((keys[q] >> lo_bits) != top) {
            hi = q;
            break;
        }
        last = q;
        step *= 2;
    }
    while (hi - last > 1) {  // keys[last] in the bucket, keys[hi] past it (or hi == n)
        const int64_t mid = last + (hi - last) / 2;
        if ((keys[mid] >> lo_bits) == top) last = mid; else hi = mid;
    }
    return hi;
}

// Lists of the buckets larger than kLocSmall, per local tile: class A (at
// most kMidMax keys) and class B (larger), each bucket as (start, end).
struct BucketLists {
    int2* a;           // [tile * kASlots + j]
    int2* b;           // [tile * kBSlots + j]
    int32_t* a_count;  // per tile (n_tiles + 1 entries: room for the scan's total)
    int32_t* b_count;  // per tile (n_tiles + 1)
    int2* a_flat;      // every A bucket, tile by tile
    int2* b_flat;      // every B bucket, tile by tile
    int32_t* totals;   // [0] A buckets, [1] B buckets, [2] B buckets taken (work counter)
};

// keys/vals: the pairs ordered by key >> lo_bits (ties in input order).
// Writes the owned small buckets' elements in final order to keys_out /
// vals_out and lists the owned larger buckets.  When a top pass gave up
// (*err set) the pairs are not ordered by their top bits: nothing is done
// and no bucket is listed (the sort is reported failed).
__global__ void __launch_bounds__(kLocBlock) bucket_local_kernel(const uint64_t* __restrict__ keys,
                                                                 const uint32_t* __restrict__ vals, int64_t n,
                                                                 int lo_bits, uint64_t* __restrict__ keys_out,
                                                                 uint32_t* __restrict__ vals_out, BucketLists bl,
                                                                 const int32_t* __restrict__ err) {
    __shared__ uint64_t s_key[kLocWin];
    __shared__ int16_t s_h[kLocWin];  // window index of the position's bucket start (kNoPos: not visible)
    __shared__ int16_t s_e[kLocWin];  // window index past the bucket's end (kNoPos: not visible)
    __shared__ int s_w[kLocWaves];
    __shared__ int s_nbig, s_na, s_nb;
    __shared__ int16_t s_heads[kASlots];
    const int tid = static_cast<int>(threadIdx.x);
    if (*err) {
        if (tid == 0) bl.a_count[blockIdx.x] = bl.b_count[blockIdx.x] = 0;
        return;
    }
    if (tid == 0) s_nbig = s_na = s_nb = 0;
    const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kLocTile;
    const int64_t lo = t0 > kLocSmall ? t0 - kLocSmall : 0;
    const int64_t hi = t0 + kLocTile + kLocSmall < n - 1 ? t0 + kLocTile + kLocSmall : n - 1;  // inclusive
    const int m = static_cast<int>(hi - lo + 1);
    for (int i = tid; i < m; i += kLocBlock) s_key[i] = keys[lo + i];
    __syncthreads();
    // a position is a bucket head when its top bits differ from its
    // predecessor's; the window's first position is one only at 0
    const int c0 = tid * kLocPer;
    bool hd[kLocPer];
    int last = -1, first = kNoPos;
#pragma unroll
    for (int r = 0; r < kLocPer; ++r) {
        const int i = c0 + r;
        hd[r] = i < m && (i == 0 ? lo == 0 : (s_key[i] >> lo_bits) != (s_key[i - 1] >> lo_bits));
        if (hd[r]) {
            last = i;
            if (first == kNoPos) first = i;
        }
    }
    int run = block_excl_max(last, s_w);
#pragma unroll
    for (int r = 0; r < kLocPer; ++r) {
        if (hd[r]) run = c0 + r;
        if (c0 + r < m) s_h[c0 + r] = static_cast<int16_t>(run < 0 ? kNoPos : run);
    }
    // past the window: the array's end is a bucket end, anything else unknown
    int next = block_excl_suffix_min(first, s_w);
    if (next == kNoPos && hi == n - 1) next = m;
#pragma unroll
    for (int r = kLocPer - 1; r >= 0; --r) {
        if (c0 + r < m) s_e[c0 + r] = static_cast<int16_t>(next);
        if (hd[r]) next = c0 + r;
    }
    __syncthreads();
    // owned small buckets: final position = start + rank.  Such a bucket lies
    // whole inside the window (its start is owned, its end at most kLocSmall
    // further).
    const int own0 = static_cast<int>(t0 - lo), own1 = own0 + kLocTile;  // owned starts [own0, own1)
    for (int i = own0 + tid; i < m && i < own1 + kLocSmall; i += kLocBlock) {
        const int h = s_h[i], e = s_e[i];
        if (h == kNoPos || e == kNoPos || h < own0 || h >= own1 || e - h > kLocSmall) continue;
        const uint64_t k = s_key[i];
        int rank = 0;
        for (int j = h; j < e; ++j) {
            const uint64_t o = s_key[j];
            rank += (o < k || (o == k && j < i)) ? 1 : 0;
        }
        const int64_t dst = lo + h + rank;
        keys_out[dst] = k;
        vals_out[dst] = vals[lo + i];
    }
    // owned larger buckets: their heads are collected, then each thread
    // finds one bucket's end (a search in HBM when it lies past the window)
    for (int i = own0 + tid; i < m && i < own1; i += kLocBlock) {
        const int e = s_e[i];
        if (s_h[i] != i || (e != kNoPos && e - i <= kLocSmall)) continue;
        s_heads[atomicAdd(&s_nbig, 1)] = static_cast<int16_t>(i);
    }
    __syncthreads();
    if (tid < s_nbig) {
        const int i = s_heads[tid], e = s_e[i];
        const int64_t end = e != kNoPos ? lo + e : bucket_end(keys, n, hi, s_key[i] >> lo_bits, lo_bits);
        const int2 r = make_int2(static_cast<int>(lo + i), static_cast<int>(end));
        if (r.y - r.x <= kMidMax) {
            bl.a[static_cast<int64_t>(blockIdx.x) * kASlots + atomicAdd(&s_na, 1)] = r;
        } else {
            bl.b[static_cast<int64_t>(blockIdx.x) * kBSlots + atomicAdd(&s_nb, 1)] = r;
        }
    }
    __syncthreads();
    if (tid == 0) {
        bl.a_count[blockIdx.x] = s_na;
        bl.b_count[blockIdx.x] = s_nb;
    }
}

// Larger buckets (more than kLocSmall keys), one workgroup each: a stable
// LSD radix sort of the low digits.  A bucket of at most kSegTile keys is
// loaded once and sorted in LDS; a larger one is sorted through HBM, each
// pass a digit histogram over the bucket, then its tiles ranked in order and
// scattered at the digit's running offset.
constexpr int kSegBlock = 256;
constexpr int kSegItems = 16;
constexpr int kSegTile = kSegBlock * kSegItems;
constexpr int kSegWaves = kSegBlock / kWave;
static_assert(kSegBlock == kRadix, "one thread per digit");

struct SegLds {
    uint32_t cnt[kSegWaves][kRadix];  // per-wave digit counts, then the waves' offsets
    uint32_t start[kRadix];           // tile-local first slot of each digit
    uint32_t count[kRadix];           // the tile's keys per digit
    uint32_t run[kRadix];             // HBM path: the digit's next output slot
    uint64_t keys[kSegTile];          // LDS path: the bucket in the current order
    uint16_t idx[kSegTile];           // LDS path: each key's index in the bucket
};

// Stable ranks of a tile: element r * 64 + lane of wave w's span (rows
// elements per wave, wave-contiguous spans) goes to tile slot pos[r].
__device__ __forceinline__ void seg_rank(const uint64_t (&k)[kSegItems], int rows, int shift,
                                         uint32_t (&pos)[kSegItems], SegLds& L) {
    const int tid = static_cast<int>(threadIdx.x), w = tid / kWave;
    const int lane = static_cast<int>(lane_id());
    for (int d = lane; d < kRadix; d += kWave) L.cnt[w][d] = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
    for (int r = 0; r < kSegItems; ++r) {
        if (r < rows) {
            const uint32_t d = static_cast<uint32_t>(k[r] >> shift) & (kRadix - 1);
            uint32_t below, total;
            rank_digit(d, below, total);
            const uint32_t c = L.cnt[w][d];
            pos[r] = c + below;
            if (below + 1u == total) L.cnt[w][d] = c + total;
        }
    }
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (int wv = 0; wv < kSegWaves; ++wv) {
        const uint32_t c = L.cnt[wv][tid];
        L.cnt[wv][tid] = run;
        run += c;
    }
    L.count[tid] = run;
    int total;
    L.start[tid] = static_cast<uint32_t>(block_excl_sum(static_cast<int>(run), &total));
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSegItems; ++r) {
        if (r < rows) {
            const uint32_t d = static_cast<uint32_t>(k[r] >> shift) & (kRadix - 1);
            pos[r] += L.start[d] + L.cnt[w][d];
        }
    }
}

// Exclusive scan of v[0 .. n) in place by one workgroup; returns the total
// (uniform).
constexpr int kSegBlocks = 768;  // three per CU
__device__ int block_scan_inplace(int32_t* v, int n) {
    int carry = 0;
    for (int base = 0; base < n; base += kScanBlock * kTilesPer) {
        const int i0 = base + static_cast<int>(threadIdx.x) * kTilesPer;
        int x[kTilesPer];
        int sum = 0;
#pragma unroll
        for (int j = 0; j < kTilesPer; ++j) {
            x[j] = i0 + j < n ? v[i0 + j] : 0;
            sum += x[j];
        }
        int total;
        int run = carry + block_excl_sum(sum, &total);
#pragma unroll
        for (int j = 0; j < kTilesPer; ++j) {
            if (i0 + j < n) v[i0 + j] = run;
            run += x[j];
        }
        carry += total;
    }
    return carry;
}

// One workgroup: the tiles' lists copied into flat arrays (tile order), the
// list lengths into totals[0..1] and the B work counter cleared.
__global__ void __launch_bounds__(kScanBlock) bucket_scan_kernel(BucketLists bl, int n_tiles) {
    const int na = block_scan_inplace(bl.a_count, n_tiles);
    const int nb = block_scan_inplace(bl.b_count, n_tiles);
    if (threadIdx.x == 0) {
        bl.a_count[n_tiles] = na;
        bl.b_count[n_tiles] = nb;
        bl.totals[0] = na;
        bl.totals[1] = nb;
        bl.totals[2] = 0;
    }
    __syncthreads();
    if (na == 0 && nb == 0) return;
    constexpr int kCopy = 8;  // entries whose loads are issued together
    for (int t = static_cast<int>(threadIdx.x); t < n_tiles; t += kScanBlock) {
        const int a0 = bl.a_count[t], a1 = bl.a_count[t + 1];
        for (int j0 = a0; j0 < a1; j0 += kCopy) {
            int2 v[kCopy];
#pragma unroll
            for (int c = 0; c < kCopy; ++c)
                if (j0 + c < a1) v[c] = bl.a[static_cast<int64_t>(t) * kASlots + (j0 + c - a0)];
#pragma unroll
            for (int c = 0; c < kCopy; ++c)
                if (j0 + c < a1) bl.a_flat[j0 + c] = v[c];
        }
        const int b0 = bl.b_count[t], b1 = bl.b_count[t + 1];
        for (int j = b0; j < b1; ++j) bl.b_flat[j] = bl.b[static_cast<int64_t>(t) * kBSlots + (j - b0)];
    }
}

// Wave-level sort of a bucket of at most kMidMax keys: each element's rank
// is the number of keys below it, or equal and earlier, counted against the
// bucket's low words staged in the wave's LDS (the keys of a bucket differ
// only below bit lo_bits <= 32).
__device__ __forceinline__ void sort_bucket_wave(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                 int2 r, uint32_t* s, uint64_t* __restrict__ keys_out,
                                                 uint32_t* __restrict__ vals_out) {
    constexpr int kRows = kMidMax / kWave;
    const int lane = static_cast<int>(lane_id());
    const int m = r.y - r.x;
    const int rows = (m + kWave - 1) / kWave;
    uint64_t k[kRows];
    uint32_t lo[kRows];
    int rank[kRows];
#pragma unroll
    for (int q = 0; q < kRows; ++q) {
        const int e = q * kWave + lane;
        k[q] = e < m ? keys[r.x + e] : ~0ull;
        lo[q] = static_cast<uint32_t>(k[q]);
        if (e < m) s[e] = lo[q];
        rank[q] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int j = 0; j < m; ++j) {
        const uint32_t o = s[j];
#pragma unroll
        for (int q = 0; q < kRows; ++q)
            if (q < rows) rank[q] += (o < lo[q] || (o == lo[q] && j < q * kWave + lane)) ? 1 : 0;
    }
#pragma unroll
    for (int q = 0; q < kRows; ++q) {
        const int e = q * kWave + lane;
        if (q < rows && e < m) {
            keys_out[r.x + rank[q]] = k[q];
            vals_out[r.x + rank[q]] = vals[r.x + e];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // s is refilled by the next bucket
}

// Class A: wave g of W sorts the A buckets [g * NA / W, (g + 1) * NA / W)
// of the flat list, one bucket at a time.
constexpr int kABlock = 256;
constexpr int kAWaves = kABlock / kWave;
constexpr int kABlocks = 2048;  // eight waves per SIMD
__global__ void __launch_bounds__(kABlock) bucket_wave_kernel(const uint64_t* __restrict__ keys,
                                                              const uint32_t* __restrict__ vals, BucketLists bl,
                                                              uint64_t* __restrict__ keys_out,
                                                              uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t s_lo[kAWaves][kMidMax];
    const int w = static_cast<int>(threadIdx.x) / kWave;
    const int64_t W = static_cast<int64_t>(gridDim.x) * kAWaves;
    const int64_t g = static_cast<int64_t>(blockIdx.x) * kAWaves + w;
    const int64_t NA = bl.totals[0];
    const int f0 = static_cast<int>(g * NA / W), f1 = static_cast<int>((g + 1) * NA / W);
    for (int f = f0; f < f1; ++f) sort_bucket_wave(keys, vals, bl.a_flat[f], s_lo[w], keys_out, vals_out);
}

// Class B: the workgroups take the B buckets one at a time from a counter.
// keys/vals: the pairs ordered by top bits; tmp*: two buffers of n pairs for
// the buckets sorted through HBM.
__global__ void __launch_bounds__(kSegBlock) bucket_segment_kernel(const uint64_t* __restrict__ keys,
                                                                   const uint32_t* __restrict__ vals, BucketLists bl,
                                                                   int lo_bits, uint64_t* __restrict__ keys_out,
                                                                   uint32_t* __restrict__ vals_out, uint64_t* tmp_k0,
                                                                   uint32_t* tmp_v0, uint64_t* tmp_k1,
                                                                   uint32_t* tmp_v1) {
    __shared__ SegLds L;
    __shared__ int s_next;
    const int tid = static_cast<int>(threadIdx.x), w = tid / kWave;
    const int lane = static_cast<int>(lane_id());
    const int passes = lo_bits / 8;
    const int NB = bl.totals[1];
    if (NB == 0) return;  // uniform
    for (;;) {
        if (tid == 0) s_next = atomicAdd(&bl.totals[2], 1);
        __syncthreads();
        const int j = s_next;
        __syncthreads();
        if (j >= NB) return;
        {
            const int2 r = bl.b_flat[j];
            const int m = r.y - r.x;
            uint64_t k[kSegItems];
            uint32_t pos[kSegItems];
            if (m <= kSegTile) {
                // in LDS: element r * 64 + lane of wave w is bucket element
                // w * span + r * 64 + lane; padding (past m) sorts last
                const int rows = (m + kSegBlock - 1) / kSegBlock;
                const int span = rows * kWave;
                uint32_t ix[kSegItems];
#pragma unroll
                for (int q = 0; q < kSegItems; ++q) {
                    const int e = w * span + q * kWave + lane;
                    k[q] = (q < rows && e < m) ? keys[r.x + e] : ~0ull;
                    ix[q] = static_cast<uint32_t>(e);
                }
                for (int p = 0; p < passes; ++p) {
                    seg_rank(k, rows, 8 * p, pos, L);
#pragma unroll
                    for (int q = 0; q < kSegItems; ++q) {
                        if (q < rows) {
                            L.keys[pos[q]] = k[q];
                            L.idx[pos[q]] = static_cast<uint16_t>(ix[q]);
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int q = 0; q < kSegItems; ++q) {
                        if (q < rows) {
                            const int e = w * span + q * kWave + lane;
                            k[q] = L.keys[e];
                            ix[q] = L.idx[e];
                        }
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int q = 0; q < kSegItems; ++q) {
                    const int e = w * span + q * kWave + lane;
                    if (q < rows && e < m) {
                        keys_out[r.x + e] = k[q];
                        vals_out[r.x + e] = vals[r.x + ix[q]];
                    }
                }
                continue;
            }
            // through HBM: keys -> tmp0 -> tmp1 -> tmp0 ... -> keys_out
            for (int p = 0; p < passes; ++p) {
                const uint64_t* sk = p == 0 ? keys : (p % 2 ? tmp_k0 : tmp_k1);
                const uint32_t* sv = p == 0 ? vals : (p % 2 ? tmp_v0 : tmp_v1);
                uint64_t* dk = p == passes - 1 ? keys_out : (p % 2 ? tmp_k1 : tmp_k0);
                uint32_t* dv = p == passes - 1 ? vals_out : (p % 2 ? tmp_v1 : tmp_v0);
                const int shift = 8 * p;
                L.run[tid] = 0;
                __syncthreads();
                for (int c0 = r.x; c0 < r.y; c0 += kSegTile) {  // loads issued a chunk at a time
                    uint64_t hk[kSegItems];
#pragma unroll
                    for (int q = 0; q < kSegItems; ++q) {
                        const int i = c0 + q * kSegBlock + tid;
                        hk[q] = i < r.y ? sk[i] : 0ull;
                    }
#pragma unroll
                    for (int q = 0; q < kSegItems; ++q)
                        if (c0 + q * kSegBlock + tid < r.y)
                            atomicAdd(&L.run[static_cast<uint32_t>(hk[q] >> shift) & (kRadix - 1)], 1u);
                }
                __syncthreads();
                {
                    int total;
                    const int ex = block_excl_sum(static_cast<int>(L.run[tid]), &total);
                    L.run[tid] = static_cast<uint32_t>(ex);
                }
                __syncthreads();
                for (int t0 = r.x; t0 < r.y; t0 += kSegTile) {
                    uint32_t v[kSegItems];
#pragma unroll
                    for (int q = 0; q < kSegItems; ++q) {
                        const int e = t0 + w * (kSegItems * kWave) + q * kWave + lane;
                        k[q] = e < r.y ? sk[e] : ~0ull;
                        v[q] = e < r.y ? sv[e] : 0u;
                    }
                    seg_rank(k, kSegItems, shift, pos, L);
#pragma unroll
                    for (int q = 0; q < kSegItems; ++q) {
                        const int e = t0 + w * (kSegItems * kWave) + q * kWave + lane;
                        if (e < r.y) {
                            const uint32_t d = static_cast<uint32_t>(k[q] >> shift) & (kRadix - 1);
                            const int64_t dst = r.x + L.run[d] + (pos[q] - L.start[d]);
                            if (dst < r.y) {  // always, for buckets of a valid list
                                dk[dst] = k[q];
                                dv[dst] = v[q];
                            }
                        }
                    }
                    __syncthreads();
                    L.run[tid] += L.count[tid];  // past the bucket's end only padding is counted, unused
                    __syncthreads();
                }
            }
        }
    }
}

inline int64_t loc_tiles(int64_t n) { return (n + kLocTile - 1) / kLocTile; }

// Workspace of the onesweep and hybrid schedules.  The look-back error flag
// comes first, where sx_sort_check finds it for either schedule; everything
// up to the look-back words of the passes that run is cleared by one fill per
// sort.
struct SortWs {
    int32_t* err;
    uint32_t* ghist;     // digit histograms, then their exclusive scans
    uint32_t* counters;  // one tile-ticket counter per pass
    uint32_t* status;    // look-back words: kRadix per tile and pass
    uint64_t* keys_tmp;
    uint32_t* vals_tmp;
    // hybrid only: the larger buckets listed per local tile, two buffers of
    // pairs for the buckets sorted through HBM
    BucketLists lists;
    uint64_t* seg_k[2];
    uint32_t* seg_v[2];
    size_t bytes;
};

SortWs sort_layout(void* ws, int64_t n, bool hybrid) {
    const int64_t tiles = sort_tiles(n);
    Carver c(ws);
    SortWs w{};
    w.err = c.take<int32_t>(1);
    w.ghist = c.take<uint32_t>(kMaxPasses * kRadix);
    w.counters = c.take<uint32_t>(kMaxPasses);
    w.status = c.take<uint32_t>(static_cast<size_t>(kMaxPasses) * kRadix * tiles);
    w.keys_tmp = c.take<uint64_t>(static_cast<size_t>(n));
    w.vals_tmp = c.take<uint32_t>(static_cast<size_t>(n));
    if (hybrid) {
        const size_t lt = static_cast<size_t>(loc_tiles(n));
        w.lists.a = c.take<int2>(lt * kASlots);
        w.lists.b = c.take<int2>(lt * kBSlots);
        w.lists.a_count = c.take<int32_t>(lt + 1);
        w.lists.b_count = c.take<int32_t>(lt + 1);
        w.lists.a_flat = c.take<int2>(lt * kASlots);
        w.lists.b_flat = c.take<int2>(lt * kBSlots);
        w.lists.totals = c.take<int32_t>(3);
        for (int i = 0; i < 2; ++i) {
            w.seg_k[i] = c.take<uint64_t>(static_cast<size_t>(n));
            w.seg_v[i] = c.take<uint32_t>(static_cast<size_t>(n));
        }
    }
    w.bytes = c.off;
    return w;
}

// Clears everything before the look-back words and those of the first
// `passes` passes.
int clear_sort_state(const SortWs& w, void* ws, int64_t n, int passes, hipStream_t st) {
    const size_t bytes = static_cast<size_t>(reinterpret_cast<char*>(w.status) - static_cast<char*>(ws)) +
                         static_cast<size_t>(passes) * kRadix * sort_tiles(n) * sizeof(uint32_t);
    return fill_u8(ws, 0, bytes, st);
}

// Spins a look-back waits for an earlier tile before it gives up (a host
// setting, read at launch).
std::atomic<uint32_t> g_sort_spin_limit{kDefaultSpinLimit};

// 512 histogram blocks at most: each block's flush atomics land on the same
// device-wide counters, so fewer, longer blocks contend less (8M keys: 51 us
// at 2048 blocks, 29 us at 512, 33 us at 256).
inline unsigned hist_blocks(int64_t n) {
    const unsigned hb = grid_for(n, kSortBlock * kHistItems);
    return hb > 512 ? 512 : hb;
}

// The passes over digits [first, first + passes), each pass's output the
// next one's input, starting from (keys_in, vals_in) and ending in
// (keys_out, vals_out); tmp buffers take the other half of the ping-pong.
int onesweep_passes(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int first, int passes,
                    uint32_t* ghist, uint32_t* counters, uint32_t* status, int32_t* err, uint64_t* keys_out,
                    uint32_t* vals_out, uint64_t* keys_tmp, uint32_t* vals_tmp, hipStream_t st) {
    const int64_t tiles = sort_tiles(n);
    onesweep_hist_kernel<<<hist_blocks(n), kSortBlock, 0, st>>>(keys_in, n, first, passes, ghist);
    SX_LAUNCH_CHECK("onesweep_hist_kernel");
    onesweep_base_kernel<<<passes, kRadix, 0, st>>>(ghist);
    SX_LAUNCH_CHECK("onesweep_base_kernel");
    const uint32_t spin_limit = g_sort_spin_limit.load();
    const uint64_t* src_k = keys_in;
    const uint32_t* src_v = vals_in;
    for (int p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint64_t* dst_k = to_out ? keys_out : keys_tmp;
        uint32_t* dst_v = to_out ? vals_out : vals_tmp;
        onesweep_pass_kernel<<<static_cast<unsigned>(tiles), kPassBlock, 0, st>>>(
            src_k, src_v, n, 8 * (first + p), ghist + p * kRadix, status + static_cast<size_t>(p) * kRadix * tiles,
            counters + p, spin_limit, err, dst_k, dst_v);
        SX_LAUNCH_CHECK("onesweep_pass_kernel");
        src_k = dst_k;
        src_v = dst_v;
    }
    return SX_OK;
}

int onesweep_launch(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int key_bits, uint64_t* keys_out,
                    uint32_t* vals_out, void* ws, int32_t* err, hipStream_t st) {
    const int passes = (key_bits + 7) / 8;
    const SortWs w = sort_layout(ws, n, false);
    SX_TRY(clear_sort_state(w, ws, n, passes, st));
    return onesweep_passes(keys_in, vals_in, n, 0, passes, w.ghist, w.counters, w.status, err ? err : w.err,
                           keys_out, vals_out, w.keys_tmp, w.vals_tmp, st);
}

int hybrid_launch(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int key_bits, uint64_t* keys_out,
                  uint32_t* vals_out, void* ws, int32_t* err, hipStream_t st) {
    const int passes = (key_bits + 7) / 8;
    const int top0 = passes - kTopPasses;  // first top digit
    const int lo_bits = 8 * top0;
    const int64_t lt = loc_tiles(n);
    const SortWs w = sort_layout(ws, n, true);
    SX_TRY(clear_sort_state(w, ws, n, kTopPasses, st));
    if (!err) err = w.err;
    // 1. the top digits, ending in keys_tmp (the bucket steps write keys_out)
    SX_TRY(onesweep_passes(keys_in, vals_in, n, top0, kTopPasses, w.ghist, w.counters, w.status, err,
                           w.keys_tmp, w.vals_tmp, keys_out, vals_out, st));
    // 2. small buckets in place, larger ones listed
    bucket_local_kernel<<<static_cast<unsigned>(lt), kLocBlock, 0, st>>>(w.keys_tmp, w.vals_tmp, n, lo_bits, keys_out,
                                                                         vals_out, w.lists, err);
    SX_LAUNCH_CHECK("bucket_local_kernel");
    // 3. larger buckets: equal shares of the lists, a wave per bucket of at
    // most kMidMax keys, a workgroup per larger one
    bucket_scan_kernel<<<1, kScanBlock, 0, st>>>(w.lists, static_cast<int>(lt));
    SX_LAUNCH_CHECK("bucket_scan_kernel");
    bucket_wave_kernel<<<kABlocks, kABlock, 0, st>>>(w.keys_tmp, w.vals_tmp, w.lists, keys_out, vals_out);
    SX_LAUNCH_CHECK("bucket_wave_kernel");
    bucket_segment_kernel<<<kSegBlocks, kSegBlock, 0, st>>>(w.keys_tmp, w.vals_tmp, w.lists, lo_bits, keys_out,
                                                            vals_out, w.seg_k[0], w.seg_v[0], w.seg_k[1], w.seg_v[1]);
    SX_LAUNCH_CHECK("bucket_segment_kernel");
    return SX_OK;
}

size_t sort_ws_bytes_3k(int64_t n) {
    const int64_t tiles = sort_tiles(n);
    Carver c(nullptr);
    c.take<uint64_t>(static_cast<size_t>(n));
    c.take<uint32_t>(static_cast<size_t>(n));
    c.take<uint32_t>(static_cast<size_t>(kRadix * tiles));
    c.take<char>(scan_ws_bytes(kRadix * tiles));
    return c.off;
}

int sort_3k_launch(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int key_bits, uint64_t* keys_out,
                   uint32_t* vals_out, void* ws, hipStream_t st) {
    if (n <= 0) return SX_OK;
    if (n >= (1ll << 31)) return fail(SX_EINVAL, "sort: n must be below 2^31");
    const int64_t tiles = sort_tiles(n);
    Carver c(ws);
    uint64_t* keys_tmp = c.take<uint64_t>(static_cast<size_t>(n));
    uint32_t* vals_tmp = c.take<uint32_t>(static_cast<size_t>(n));
    uint32_t* hist = c.take<uint32_t>(static_cast<size_t>(kRadix * tiles));
    void* scan_ws = c.take<char>(scan_ws_bytes(kRadix * tiles));
    const int passes = (key_bits + 7) / 8;
    const uint64_t* src_k = keys_in;
    const uint32_t* src_v = vals_in;
    for (int p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint64_t* dst_k = to_out ? keys_out : keys_tmp;
        uint32_t* dst_v = to_out ? vals_out : vals_tmp;
        const int shift = 8 * p;
        radix_hist_kernel<<<static_cast<unsigned>(tiles), kSortBlock, 0, st>>>(src_k, n, shift, hist,
                                                                               static_cast<int>(tiles));
        SX_LAUNCH_CHECK("radix_hist_kernel");
        SX_TRY(device_scan(HistIn{hist}, HistOut{hist}, nullptr, kRadix * tiles, nullptr, scan_ws, st));
        radix_scatter_kernel<<<static_cast<unsigned>(tiles), kSortBlock, 0, st>>>(
            src_k, src_v, n, shift, hist, static_cast<int>(tiles), dst_k, dst_v);
        SX_LAUNCH_CHECK("radix_scatter_kernel");
        src_k = dst_k;
        src_v = dst_v;
    }
    return SX_OK;
}


// Below 2^30 keys: the hybrid schedule for keys of seven or eight digits
// (at least one local tile of them), else onesweep, unless onesweep alone or
// the three-kernel passes (histogram, scan, scatter) are asked for; those
// also serve larger inputs.
inline bool hybrid_applies(int64_t n, int key_bits) { return key_bits > 48 && n >= kLocTile && loc_tiles(n) < (1ll << 30); }

size_t sort_ws_bytes(int64_t n) {
    return n < (1ll << 30) ? sort_layout(nullptr, n, true).bytes : sort_ws_bytes_3k(n);
}

// err (nullable): device int32 that a failed onesweep look-back sets to
// nonzero (without it the flag stays in the workspace, SortWs::err).
int sort_pairs_launch(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int key_bits, uint64_t* keys_out,
                      uint32_t* vals_out, void* ws, int algo, hipStream_t st, int32_t* err) {
    if (n <= 0) return SX_OK;
    if (n < (1ll << 30) && algo == SX_SORT_HYBRID && hybrid_applies(n, key_bits))
        return hybrid_launch(keys_in, vals_in, n, key_bits, keys_out, vals_out, ws, err, st);
    if (n < (1ll << 30) && (algo == SX_SORT_ONESWEEP || algo == SX_SORT_HYBRID))
        return onesweep_launch(keys_in, vals_in, n, key_bits, keys_out, vals_out, ws, err, st);
    return sort_3k_launch(keys_in, vals_in, n, key_bits, keys_out, vals_out, ws, st);
}

}  // namespace sx

using sx::fail;

extern "C" int sx_sort_pairs(const uint64_t* d_keys_in, const uint32_t* d_vals_in, int64_t n, int key_bits,
                             uint64_t* d_keys_out, uint32_t* d_vals_out, void* ws, size_t* ws_bytes, void* stream) {
    if (!ws_bytes) return fail(SX_EINVAL, "sx_sort_pairs: ws_bytes is null");
    if (n < 0) return fail(SX_EINVAL, "sx_sort_pairs: negative n");
    const size_t need = sx::sort_ws_bytes(n);
    if (!ws) {
        *ws_bytes = need;
        return SX_OK;
    }
    if (*ws_bytes < need) return fail(SX_ENOMEM, "sx_sort_pairs: workspace too small");
    if (key_bits < 1 || key_bits > 64) return fail(SX_EINVAL, "sx_sort_pairs: key_bits must be in [1, 64]");
    if (n > 0 && (!d_keys_in || !d_keys_out || !d_vals_out)) return fail(SX_EINVAL, "sx_sort_pairs: null pointer");
    if (d_keys_in == d_keys_out || (d_vals_in && d_vals_in == d_vals_out))
        return fail(SX_EINVAL, "sx_sort_pairs: input and output buffers must differ");
    return sx::sort_pairs_launch(d_keys_in, d_vals_in, n, key_bits, d_keys_out, d_vals_out, ws, SX_SORT_HYBRID,
                                 sx::as_stream(stream), nullptr);
}

extern "C" int sx_sort_check(const void* ws, int64_t n, void* stream) {
    if (!ws || n < 0) return fail(SX_EINVAL, "sx_sort_check: null workspace or negative n");
    if (n == 0 || n >= (1ll << 30)) return SX_OK;  // no look-back ran
    int32_t h = 0;
    hipStream_t st = sx::as_stream(stream);
    SX_HIP(hipMemcpyAsync(&h, sx::sort_layout(const_cast<void*>(ws), n, false).err, sizeof(h), hipMemcpyDeviceToHost,
                          st));
    SX_HIP(hipStreamSynchronize(st));
    return h ? fail(SX_EINTERNAL, "sx_sort_pairs: a look-back gave up waiting for an earlier tile; the sort is invalid")
             : SX_OK;
}

extern "C" int sx_sort_set_spin_limit(uint32_t spins) {
    sx::g_sort_spin_limit.store(spins);
    return SX_OK;
}
