// Stable LSD radix sort of (uint64 key, uint32 value) pairs, 8-bit digits.
//
// Replaces cub::DeviceRadixSort::SortKeys (src/RadixTree.cu:316-335).  The
// reference sorts keys only; this sort also carries the input index so that
// the sort permutation (and with it the duplicate representatives and the
// input-order scatter of results) is available.
//
// Two schedules share the tile ranking code:
//
// Onesweep (default): one histogram kernel counts every digit of every pass
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

__global__ void __launch_bounds__(kSortBlock) onesweep_hist_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                                   int passes, uint32_t* __restrict__ ghist) {
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
                atomicAdd(&h[p][static_cast<uint32_t>(k[j] >> (8 * p)) & (kRadix - 1)], 1u);
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

size_t onesweep_ws_bytes(int64_t n, int passes) {
    const int64_t tiles = sort_tiles(n);
    Carver c(nullptr);
    c.take<uint64_t>(static_cast<size_t>(n));
    c.take<uint32_t>(static_cast<size_t>(n));
    c.take<uint32_t>(kMaxPasses * kRadix);
    c.take<uint32_t>(kMaxPasses + 1);  // per-pass tile tickets + a look-back error flag
    c.take<uint32_t>(static_cast<size_t>(passes) * kRadix * tiles);
    return c.off;
}

// The look-back error flag inside a workspace laid out for n keys.
int32_t* onesweep_ws_flag(void* ws, int64_t n) {
    Carver c(ws);
    c.take<uint64_t>(static_cast<size_t>(n));
    c.take<uint32_t>(static_cast<size_t>(n));
    c.take<uint32_t>(kMaxPasses * kRadix);
    return reinterpret_cast<int32_t*>(c.take<uint32_t>(kMaxPasses + 1) + kMaxPasses);
}

// Spins a look-back waits for an earlier tile before it gives up (a host
// setting, read at launch).
std::atomic<uint32_t> g_sort_spin_limit{kDefaultSpinLimit};

int onesweep_launch(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int key_bits, uint64_t* keys_out,
                    uint32_t* vals_out, void* ws, int32_t* err, hipStream_t st) {
    const int passes = (key_bits + 7) / 8;
    const int64_t tiles = sort_tiles(n);
    Carver c(ws);
    uint64_t* keys_tmp = c.take<uint64_t>(static_cast<size_t>(n));
    uint32_t* vals_tmp = c.take<uint32_t>(static_cast<size_t>(n));
    uint32_t* ghist = c.take<uint32_t>(kMaxPasses * kRadix);
    uint32_t* counters = c.take<uint32_t>(kMaxPasses + 1);
    uint32_t* status = c.take<uint32_t>(static_cast<size_t>(passes) * kRadix * tiles);
    // ghist, counters and status are contiguous: clear them in one fill
    const size_t clear = reinterpret_cast<char*>(status + static_cast<size_t>(passes) * kRadix * tiles) -
                         reinterpret_cast<char*>(ghist);
    SX_TRY(fill_u8(ghist, 0, clear, st));
    // 512 blocks at most: each block's 2048 flush atomics land on the same
    // 2048 device-wide counters, so fewer, longer blocks contend less
    // (8M keys: 51 us at 2048 blocks, 29 us at 512, 33 us at 256)
    unsigned hb = grid_for(n, kSortBlock * kHistItems);
    if (hb > 512) hb = 512;
    onesweep_hist_kernel<<<hb, kSortBlock, 0, st>>>(keys_in, n, passes, ghist);
    SX_LAUNCH_CHECK("onesweep_hist_kernel");
    onesweep_base_kernel<<<passes, kRadix, 0, st>>>(ghist);
    SX_LAUNCH_CHECK("onesweep_base_kernel");
    if (!err) err = reinterpret_cast<int32_t*>(counters + kMaxPasses);
    const uint32_t spin_limit = g_sort_spin_limit.load();
    const uint64_t* src_k = keys_in;
    const uint32_t* src_v = vals_in;
    for (int p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint64_t* dst_k = to_out ? keys_out : keys_tmp;
        uint32_t* dst_v = to_out ? vals_out : vals_tmp;
        onesweep_pass_kernel<<<static_cast<unsigned>(tiles), kPassBlock, 0, st>>>(
            src_k, src_v, n, 8 * p, ghist + p * kRadix, status + static_cast<size_t>(p) * kRadix * tiles,
            counters + p, spin_limit, err, dst_k, dst_v);
        SX_LAUNCH_CHECK("onesweep_pass_kernel");
        src_k = dst_k;
        src_v = dst_v;
    }
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


// Onesweep below 2^30 keys unless the three-kernel passes (histogram, scan,
// scatter) are asked for; those also serve larger inputs.
size_t sort_ws_bytes(int64_t n) {
    return n < (1ll << 30) ? onesweep_ws_bytes(n, kMaxPasses) : sort_ws_bytes_3k(n);
}

// err (nullable): device int32 that a failed onesweep look-back sets to
// nonzero (without it the flag stays in the workspace, onesweep_ws_flag).
int sort_pairs_launch(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int key_bits, uint64_t* keys_out,
                      uint32_t* vals_out, void* ws, int algo, hipStream_t st, int32_t* err) {
    if (n <= 0) return SX_OK;
    if (n < (1ll << 30) && algo == SX_SORT_ONESWEEP)
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
    return sx::sort_pairs_launch(d_keys_in, d_vals_in, n, key_bits, d_keys_out, d_vals_out, ws, SX_SORT_ONESWEEP,
                                 sx::as_stream(stream), nullptr);
}

extern "C" int sx_sort_check(const void* ws, int64_t n, void* stream) {
    if (!ws || n < 0) return fail(SX_EINVAL, "sx_sort_check: null workspace or negative n");
    if (n == 0 || n >= (1ll << 30)) return SX_OK;  // no look-back ran
    int32_t h = 0;
    hipStream_t st = sx::as_stream(stream);
    SX_HIP(hipMemcpyAsync(&h, sx::onesweep_ws_flag(const_cast<void*>(ws), n), sizeof(h), hipMemcpyDeviceToHost, st));
    SX_HIP(hipStreamSynchronize(st));
    return h ? fail(SX_EINTERNAL, "sx_sort_pairs: a look-back gave up waiting for an earlier tile; the sort is invalid")
             : SX_OK;
}

extern "C" int sx_sort_set_spin_limit(uint32_t spins) {
    sx::g_sort_spin_limit.store(spins);
    return SX_OK;
}
