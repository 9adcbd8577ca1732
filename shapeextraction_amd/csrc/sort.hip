// Stable LSD radix sort of (uint64 key, uint32 value) pairs, 8-bit digits.
//
// Replaces cub::DeviceRadixSort::SortKeys (src/RadixTree.cu:316-335).  The
// reference sorts keys only; this sort also carries the input index so that
// the sort permutation (and with it the duplicate representatives and the
// input-order scatter of results) is available.
//
// Per digit pass:
//   1. radix_hist_kernel     per-tile 256-bin digit histogram (LDS atomics),
//                            stored digit-major: hist[d * tiles + t]
//   2. device_scan           exclusive scan of that array = global output
//                            offset of (digit d, tile t) for a stable sort
//   3. radix_scatter_kernel  stable rank of every key inside its tile (wave
//                            ballot match on the digit + per-wave running
//                            digit counters in LDS), then scatter.
// A tile is 2048 keys: 4 waves x 8 rounds x 64 lanes, each wave owning a
// contiguous 512-key span so that ranks follow input order.
#include <hip/hip_runtime.h>

#include "sx_internal.hpp"

namespace sx {

constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / kWave;
constexpr int kSortItems = 8;
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

// Lanes of the wave whose digit equals this lane's digit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d) {
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const uint32_t bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
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

    const uint64_t lt = (1ull << lane) - 1ull;
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
        const uint64_t m = match_digit(d);
        const uint32_t c = cnt[w][d];
        rank[r] = c + static_cast<uint32_t>(__popcll(m & lt));
        if (((m >> lane) >> 1) == 0) cnt[w][d] = c + static_cast<uint32_t>(__popcll(m));
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

size_t sort_ws_bytes(int64_t n) {
    const int64_t tiles = sort_tiles(n);
    Carver c(nullptr);
    c.take<uint64_t>(static_cast<size_t>(n));
    c.take<uint32_t>(static_cast<size_t>(n));
    c.take<uint32_t>(static_cast<size_t>(kRadix * tiles));
    c.take<char>(scan_ws_bytes(kRadix * tiles));
    return c.off;
}

int sort_pairs_launch(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int key_bits, uint64_t* keys_out,
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
    return sx::sort_pairs_launch(d_keys_in, d_vals_in, n, key_bits, d_keys_out, d_vals_out, ws, sx::as_stream(stream));
}
