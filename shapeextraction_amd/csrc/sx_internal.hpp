// Internal definitions shared by the HIP translation units: error macro,
// launch helpers, the traversal node layout and the device-wide scan.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "sx_device.hpp"
#include "sx_status.hpp"

#define SX_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t sx_e_ = (call);                                                              \
        if (sx_e_ != hipSuccess)                                                                \
            return sx::fail(SX_EHIP, std::string(#call) + " failed: " + hipGetErrorString(sx_e_)); \
    } while (0)

#define SX_LAUNCH_CHECK(name)                                                                   \
    do {                                                                                        \
        hipError_t sx_e_ = hipGetLastError();                                                   \
        if (sx_e_ != hipSuccess)                                                                \
            return sx::fail(SX_EHIP, std::string(name) + " launch failed: " + hipGetErrorString(sx_e_)); \
    } while (0)

#define SX_TRY(expr)                      \
    do {                                  \
        int sx_rc_ = (expr);              \
        if (sx_rc_ != SX_OK) return sx_rc_; \
    } while (0)

namespace sx {

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g < 1 ? 1 : g);
}

inline size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

// ---------------------------------------------------------------------------
// Batched clouds.  Cloud c's unique points occupy slots [base[c], base[c] +
// size[c]) of the index arrays, base[c] a multiple of kCloudAlign, so every
// 64-slot tile of queries lies in one cloud and a cloud's internal radix
// nodes carry the same numbers as a standalone build shifted by base[c].  The
// default (all null) is one cloud of U points at slot 0.
// ---------------------------------------------------------------------------
constexpr int kCloudAlign = 64;
struct CloudMap {
    const int32_t* tile_cloud = nullptr;  // cloud of each kCloudAlign-slot tile
    const int32_t* base = nullptr;
    const int32_t* size = nullptr;
    // (first slot, unique points) of the cloud holding `slot`; U is the point
    // count of the single-cloud case
    __device__ __forceinline__ int2 at(int64_t slot, int U) const {
        if (!tile_cloud) return make_int2(0, U);
        const int c = tile_cloud[slot / kCloudAlign];
        return make_int2(base[c], size[c]);
    }
};

// Cloud of input point i when cloud c holds points [off[c], off[c+1]): the
// last cloud whose first point is <= i (empty clouds are skipped).
__device__ __forceinline__ int cloud_of(const int64_t* __restrict__ off, int n_clouds, int64_t i) {
    int lo = 0, hi = n_clouds - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Workgroups are dispatched round-robin over the XCDs (8 on MI355X), each
// with its own L2.  xcd_block() renumbers them so that every XCD receives a
// contiguous run of block indices: neighbouring tiles (Morton neighbours,
// sharing candidate points and tree nodes) then meet in the same L2.  A
// bijection on [0, gridDim.x) for any grid size.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

// Carves 256-byte aligned sub-buffers out of one workspace allocation.  With
// base == nullptr it only measures.
struct Carver {
    char* base;
    size_t off = 0;
    explicit Carver(void* b) : base(static_cast<char*>(b)) {}
    template <typename T>
    T* take(size_t count) {
        size_t o = off;
        off = align_up(off + count * sizeof(T));
        return base ? reinterpret_cast<T*>(base + o) : nullptr;
    }
};

// Traversal record of internal radix-tree node i (a BVH node over the sorted
// unique points): the boxes of its two children plus, in the .w lanes, the
// leaf range [a, b] it covers, its split gamma (left child covers [a, gamma],
// right child [gamma+1, b]).  64 bytes, written whole by the node's own
// thread; parents live in a separate array (ParentRef).
struct __align__(16) TravNode {
    float4 lo_l;  // w: a
    float4 hi_l;  // w: gamma
    float4 lo_r;  // w: b
    float4 hi_r;  // w: 0
};
static_assert(sizeof(TravNode) == 64, "TravNode must be 64 bytes");

__device__ __forceinline__ int w_int(float w) { return __float_as_int(w); }

// Parent of each internal node of the traversal tree (-1 at a cloud's root),
// read by the per-lane search when it climbs back up: int32 at
// p[node * stride] (the radix tree's parent array, or the .x of (parent,
// prefix) pairs).
struct ParentRef {
    const int32_t* p = nullptr;
    int stride = 1;
    __device__ __forceinline__ int at(int node) const { return p[static_cast<int64_t>(node) * stride]; }
};

// The arrays a batch's learned-radius features read (slot numbering; octree
// links in batch-octree numbering) and the bound on its rows.
struct BatchFeatures {
    const TravNode* trav = nullptr;
    const int2* aux = nullptr;        // (parent octree node, points in the cell)
    const uint8_t* pn = nullptr;      // prefix lengths
    const int32_t* leaf_parent = nullptr;
    const int32_t* leaf_oct = nullptr;
    const sx_otnode* oct = nullptr;
    int64_t oct_cap = 0;
    const float* bounds = nullptr;    // per cloud, stride 4
    const int32_t* slots = nullptr;   // device slot count
    CloudMap cm;
    int64_t limit = 0;                // slot capacity
};

// ---------------------------------------------------------------------------
// Device-wide scan of int32 values (reduce, then scan of tile sums, then scan
// within tiles).  In: int operator()(int64_t i) const for i < n.  Out:
// void operator()(int64_t i, int exclusive, int value) const.  The element
// count is n_max, or min(*d_n, n_max) when d_n is given (a device count
// produced by an earlier kernel).  The total is written to *d_total.
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int wave_incl_sum(int v) {
    const int lane = static_cast<int>(lane_id());
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive block scan for kScanBlock threads; returns the block total in *total.
__device__ __forceinline__ int block_excl_sum(int v, int* total) {
    __shared__ int wsum[kScanBlock / kWave];
    const int lane = static_cast<int>(lane_id());
    const int w = static_cast<int>(threadIdx.x) / kWave;
    int incl = wave_incl_sum(v);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < kScanBlock / kWave; ++i) {
        int s = wsum[i];
        before += (i < w) ? s : 0;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

__device__ __forceinline__ int64_t scan_count(const int32_t* d_n, int64_t n_max) {
    if (!d_n) return n_max;
    int64_t n = *d_n;
    return n < 0 ? 0 : (n < n_max ? n : n_max);
}

// Element i of a tile is visited by wave i / (kScanItems * kWave) in row
// (i / kWave) % kScanItems at lane i % kWave: every load and store of a row
// is one contiguous run (coalesced), and a tile is still kScanTile
// consecutive elements.
__device__ __forceinline__ int64_t scan_index(int64_t tile_begin, int row) {
    const int w = static_cast<int>(threadIdx.x) / kWave;
    return tile_begin + static_cast<int64_t>(w) * (kScanItems * kWave) + row * kWave + lane_id();
}

template <typename In>
__global__ void __launch_bounds__(kScanBlock) scan_reduce_kernel(In in, const int32_t* d_n, int64_t n_max,
                                                                 int* tile_sums) {
    const int64_t n = scan_count(d_n, n_max);
    const int64_t tile_begin = static_cast<int64_t>(blockIdx.x) * kScanTile;
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t i = scan_index(tile_begin, j);
        if (i < n) s += in(i);
    }
    int total;
    block_excl_sum(s, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One block: exclusive scan of the tile sums in place, total to *d_total
// (0 instead when *d_zero_if is nonzero: a failed earlier stage).  Each
// thread takes kTilesPer consecutive sums per round (their loads issued
// together), so 8M elements (4096 tiles) take one block scan, not sixteen.
constexpr int kTilesPer = 16;
template <int kBlock = kScanBlock>
__global__ void __launch_bounds__(kBlock) scan_tiles_kernel(int* tile_sums, int n_tiles, int32_t* d_total,
                                                            const int32_t* d_zero_if) {
    int carry = 0;
    for (int base = 0; base < n_tiles; base += kBlock * kTilesPer) {
        const int first = base + static_cast<int>(threadIdx.x) * kTilesPer;
        int v[kTilesPer];
        int s = 0;
#pragma unroll
        for (int j = 0; j < kTilesPer; ++j) {
            v[j] = (first + j < n_tiles) ? tile_sums[first + j] : 0;
            s += v[j];
        }
        int total;
        int run = carry + block_excl_sum(s, &total);
#pragma unroll
        for (int j = 0; j < kTilesPer; ++j) {
            if (first + j < n_tiles) tile_sums[first + j] = run;
            run += v[j];
        }
        carry += total;
    }
    if (threadIdx.x == 0 && d_total) *d_total = (d_zero_if && *d_zero_if) ? 0 : carry;
}

template <typename In, typename Out>
__global__ void __launch_bounds__(kScanBlock) scan_apply_kernel(In in, Out out, const int32_t* d_n, int64_t n_max,
                                                                const int* tile_prefix) {
    const int64_t n = scan_count(d_n, n_max);
    const int64_t tile_begin = static_cast<int64_t>(blockIdx.x) * kScanTile;
    if (tile_begin >= n) return;
    int vals[kScanItems], excl[kScanItems];
    int run = 0;  // this wave's running sum over its rows
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t i = scan_index(tile_begin, j);
        vals[j] = (i < n) ? in(i) : 0;
        const int incl = wave_incl_sum(vals[j]);
        excl[j] = run + incl - vals[j];
        run += __shfl(incl, kWave - 1, kWave);
    }
    // the waves' sums in order give each wave its offset inside the tile
    int total;
    const int wave_base = block_excl_sum(lane_id() == kWave - 1 ? run : 0, &total);
    const int base = __shfl(wave_base, kWave - 1, kWave) + tile_prefix[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const int64_t i = scan_index(tile_begin, j);
        if (i < n) out(i, base + excl[j], vals[j]);
    }
}

inline int64_t scan_tiles(int64_t n_max) { return (n_max + kScanTile - 1) / kScanTile; }
inline size_t scan_ws_bytes(int64_t n_max) { return align_up(static_cast<size_t>(scan_tiles(n_max) + 1) * sizeof(int)); }

template <typename In, typename Out>
int device_scan(In in, Out out, const int32_t* d_n, int64_t n_max, int32_t* d_total, void* ws, hipStream_t st,
                const int32_t* d_zero_if = nullptr) {
    if (n_max <= 0) return fail(SX_EINVAL, "device_scan: empty range");
    const int64_t tiles = scan_tiles(n_max);
    if (tiles > (1ll << 30)) return fail(SX_EINVAL, "device_scan: too many tiles");
    int* tile_sums = static_cast<int*>(ws);
    scan_reduce_kernel<In><<<static_cast<unsigned>(tiles), kScanBlock, 0, st>>>(in, d_n, n_max, tile_sums);
    SX_LAUNCH_CHECK("scan_reduce_kernel");
    scan_tiles_kernel<kScanBlock><<<1, kScanBlock, 0, st>>>(tile_sums, static_cast<int>(tiles), d_total, d_zero_if);
    SX_LAUNCH_CHECK("scan_tiles_kernel");
    scan_apply_kernel<In, Out><<<static_cast<unsigned>(tiles), kScanBlock, 0, st>>>(in, out, d_n, n_max, tile_sums);
    SX_LAUNCH_CHECK("scan_apply_kernel");
    return SX_OK;
}

// Host-side launchers implemented in the .hip files.
int fill_u8(void* dst, uint8_t value, size_t bytes, hipStream_t st);
int fill_i32(int32_t* dst, int32_t value, int64_t n, hipStream_t st);

}  // namespace sx
