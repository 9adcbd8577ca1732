// sx_batch: many independent clouds of uneven size through one set of
// launches.  Every stage runs once over all clouds (no per-cloud launches), so
// the whole batch is a short, fixed sequence of kernels that is captured into
// one hipGraph.  Each cloud's results equal those of a standalone build of
// that cloud (the reference builds clouds one by one: src/main.cu:80-107).
//
//   per-cloud cube       seg_bbox_kernel: one block per cloud
//   codes                seg_encode_kernel with the point's cloud's cube
//   segmented sort       stable radix sort by code, then by cloud id: the
//                        order is (cloud, code, input index), so every cloud's
//                        points stay in their input block [off[c], off[c+1])
//   duplicates           removed within a cloud only (scan over the pair)
//   cloud table          unique points per cloud; cloud c's unique points go
//                        to slots [base[c], base[c] + size[c]), base[c] a
//                        multiple of 64, so every 64-query tile is one cloud's
//   radix trees          karras_kernel per slot with the cloud map: cloud c's
//                        internal node i sits at slot base[c] + i, exactly a
//                        standalone tree shifted by base[c]
//   boxes, kNN, normals  the single-cloud kernels with the cloud map
//
//   octrees (optional)   one scan of the radix tree's edge counts over all
//                        slots, then oct_gather_kernel with the cloud map:
//                        cloud c's nodes are one run from oct_base[c], its
//                        links relative to the cloud, so the run is the
//                        standalone octree byte for byte; the feature links
//                        follow, so the learned radius runs on batches too
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "sx_internal.hpp"

namespace sx {
size_t seg_bbox_ws_bytes(int max_clouds);
int seg_bbox_launch(const float* x, const float* y, const float* z, const int64_t* off, int n_clouds, float* bounds,
                    int32_t* d_bad, float* partial, hipStream_t st);
int seg_encode_launch(const float* x, const float* y, const float* z, int64_t n, const int64_t* off, int n_clouds,
                      const float* bounds, int code_bits, uint64_t* codes, hipStream_t st);
int seg_decode_launch(const uint64_t* codes, const int32_t* d_slots, int64_t n_max, CloudMap cm, const float* bounds,
                      int code_bits, float4* pts16, hipStream_t st);
size_t sort_ws_bytes(int64_t n);
int sort_pairs_launch(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int key_bits, uint64_t* keys_out,
                      uint32_t* vals_out, void* ws, int algo, hipStream_t st, int32_t* err);
int karras_launch(const uint64_t* codes, const int32_t* d_count, int64_t n_max, int code_bits, int root_mode,
                  uint8_t* prefix_n, int32_t* left_child, uint8_t* hll, uint8_t* hlr, int32_t* parent,
                  int32_t* node_end, int32_t* leaf_parent, int2* up, TravNode* trav, const float4* pts,
                  const float* bounds, hipStream_t st, CloudMap cm, int32_t* oct_count = nullptr,
                  bool claimed_parents = false);
int knn_tiles_launch(const float4* pts, const TravNode* nodes, ParentRef par, const int32_t* d_counts, int64_t q_start,
                     int64_t n_q, int k, const float* radius2, int bucket, int32_t* out, int2* queue, int32_t* queue_count,
                     int32_t* deferred_total, hipStream_t st, CloudMap cm, int queue_blocks, int32_t* cand_total);
int normals_launch(const float4* pts, const int32_t* d_counts, const int32_t* nn, int64_t q_start, int64_t n_q, int k,
                   int precision, float vx, float vy, float vz, float* normals, uint8_t* flags, hipStream_t st,
                   CloudMap cm);
int batch_octree_launch(const uint64_t* codes, const int32_t* d_slots, int64_t slot_cap, int code_bits,
                        const float* bounds, const uint8_t* pn, const int32_t* left_child, const uint8_t* hll,
                        const uint8_t* hlr, const int32_t* counts, int32_t* offsets, CloudMap cm, int n_clouds,
                        int32_t* oct_base, int32_t* oct_size, sx_otnode* nodes, int64_t cap, int32_t* d_oct_count,
                        void* scan_ws, hipStream_t st);
int feature_links_launch(const int32_t* d_count, int64_t n_max, const int32_t* parent, const int32_t* leaf_parent,
                         const int32_t* node_end, const int32_t* offsets, const int32_t* counts, int64_t cap,
                         int32_t* leaf_oct, int2* aux, hipStream_t st, CloudMap cm);

// counts[] slots (the first four are cleared by every build; kBSortErr is
// the sorts' look-back error flag)
enum { kBSlots = 0, kBUnique = 1, kBBad = 2, kBSortErr = 3, kBDeferred = 4, kBQueue = 5, kBOct = 6 };

// Cloud of each element in code order: a binary search of the cloud offsets,
// staged in LDS when they fit (each block handles kCloudKeysItems elements
// per thread, so the staging is amortised); a gather of per-input cloud ids
// would touch one cache line per element instead.
constexpr int kCloudKeysLds = 2048;
constexpr int kCloudKeysItems = 8;
__global__ void __launch_bounds__(256) cloud_keys_kernel(const uint32_t* __restrict__ perm, int64_t n,
                                                         const int64_t* __restrict__ off, int n_clouds,
                                                         uint64_t* __restrict__ keys) {
    __shared__ int64_t s_off[kCloudKeysLds];
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x * kCloudKeysItems + threadIdx.x;
    if (n_clouds <= kCloudKeysLds) {
        for (int c = threadIdx.x; c < n_clouds; c += blockDim.x) s_off[c] = off[c];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kCloudKeysItems; ++r) {
            const int64_t i = i0 + static_cast<int64_t>(r) * blockDim.x;
            if (i < n) keys[i] = static_cast<uint64_t>(cloud_of(s_off, n_clouds, perm[i]));
        }
    } else {
        for (int r = 0; r < kCloudKeysItems; ++r) {
            const int64_t i = i0 + static_cast<int64_t>(r) * blockDim.x;
            if (i < n) keys[i] = static_cast<uint64_t>(cloud_of(off, n_clouds, perm[i]));
        }
    }
}

__global__ void __launch_bounds__(256) gather_codes_kernel(const uint64_t* __restrict__ codes,
                                                           const uint32_t* __restrict__ perm, int64_t n,
                                                           uint64_t* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = codes[perm[i]];
}

// First occurrence of a (cloud, code) pair in (cloud, code) order.
struct BatchUniqueIn {
    const uint64_t* cloud;
    const uint64_t* code;
    __device__ int operator()(int64_t i) const {
        return (i == 0 || cloud[i] != cloud[i - 1] || code[i] != code[i - 1]) ? 1 : 0;
    }
};
struct BatchUniqueOut {
    int32_t* uos;  // batch-wide unique id of each sorted position
    __device__ void operator()(int64_t i, int ex, int v) const { uos[i] = ex + v - 1; }
};

// One block: unique points per cloud, 64-aligned first slots and the slot
// total.  Cloud c's sorted positions are its input block [off[c], off[c+1]).
__global__ void __launch_bounds__(kScanBlock) cloud_table_kernel(const int64_t* __restrict__ off, int n_clouds,
                                                                 const int32_t* __restrict__ uos,
                                                                 int32_t* __restrict__ size, int32_t* __restrict__ base,
                                                                 int32_t* __restrict__ counts) {
    int carry = 0;
    for (int c0 = 0; c0 < n_clouds; c0 += kScanBlock) {
        const int c = c0 + static_cast<int>(threadIdx.x);
        int u = 0;
        if (c < n_clouds && off[c + 1] > off[c]) u = uos[off[c + 1] - 1] - uos[off[c]] + 1;
        const int slots = (u + kCloudAlign - 1) / kCloudAlign * kCloudAlign;
        int total;
        const int ex = block_excl_sum(slots, &total);
        if (c < n_clouds) {
            size[c] = u;
            base[c] = carry + ex;
        }
        carry += total;
    }
    if (threadIdx.x == 0) {
        // a failed sort leaves no slots: no later stage runs on unsorted keys
        const bool ok = counts[kBSortErr] == 0;
        counts[kBSlots] = ok ? carry : 0;
        counts[kBUnique] = ok && n_clouds > 0 ? uos[off[n_clouds] - 1] + 1 : 0;
    }
}

// Cloud of every 64-slot tile: the last cloud whose first slot is <= the tile's.
__global__ void __launch_bounds__(256) tile_cloud_kernel(const int32_t* __restrict__ base, int n_clouds,
                                                         const int32_t* __restrict__ counts, int64_t max_tiles,
                                                         int32_t* __restrict__ tile_cloud) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= max_tiles || t * kCloudAlign >= counts[kBSlots]) return;
    int lo = 0, hi = n_clouds - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (static_cast<int64_t>(base[mid]) <= t * kCloudAlign) lo = mid; else hi = mid - 1;
    }
    tile_cloud[t] = lo;
}

// Unique (cloud, code) pairs to their cloud's slots; rep = the first input
// index of the pair, relative to the cloud's first input point.
__global__ void __launch_bounds__(256) compact_kernel(int64_t n, const uint64_t* __restrict__ cloud,
                                                      const uint64_t* __restrict__ code,
                                                      const uint32_t* __restrict__ perm,
                                                      const int32_t* __restrict__ uos, const int64_t* __restrict__ off,
                                                      const int32_t* __restrict__ base, uint64_t* __restrict__ ucodes,
                                                      uint32_t* __restrict__ rep) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(i == 0 || cloud[i] != cloud[i - 1] || code[i] != code[i - 1])) return;
    const int c = static_cast<int>(cloud[i]);
    const int64_t slot = base[c] + (uos[i] - uos[off[c]]);
    ucodes[slot] = code[i];
    rep[slot] = static_cast<uint32_t>(perm[i] - off[c]);
}

// A captured batch pipeline.  Its kernels write into the owner's buffers, so
// the owner tracks its live graphs: replacing the octree storage waits for
// them to go, and a destroyed batch leaves its graphs unusable.
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    sx_batch* owner = nullptr;  // null once the batch is destroyed
};

}  // namespace sx

using sx::fail;

struct sx_batch {
    int64_t cap = 0;        // input points
    int max_clouds = 0;
    int64_t slot_cap = 0;   // cap + (kCloudAlign - 1) * max_clouds
    int code_bits = 63;
    int bucket = 16;
    int sort_algo = SX_SORT_HYBRID;
    int queue_blocks = 0;   // deferred-query workers; 0 = default
    void* mem = nullptr;
    hipEvent_t built_event = nullptr;  // recorded after every eager build
    // per input point / sorted position
    uint64_t *codes_in = nullptr, *codes_a = nullptr, *keys_c = nullptr, *keys_s = nullptr, *codes_s = nullptr;
    uint32_t *perm_a = nullptr, *perm = nullptr;
    int32_t* uos = nullptr;
    // per cloud
    int64_t* off = nullptr;
    float* bounds = nullptr;
    float* bbox_partial = nullptr;
    int32_t *size = nullptr, *base = nullptr;
    // per slot
    int32_t* tile_cloud = nullptr;
    uint64_t* ucodes = nullptr;
    uint32_t* rep = nullptr;
    float4* pts16 = nullptr;
    uint8_t* pn = nullptr;
    int32_t *lc = nullptr, *parent = nullptr;
    int2* queue = nullptr;  // deferred queries: (slot, k-th distance bound bits)
    sx::TravNode* trav = nullptr;
    int32_t* counts = nullptr;
    void* ws = nullptr;
    // per-cloud octrees (own allocation, present when enabled)
    void* oct_mem = nullptr;
    int64_t oct_cap = 0;
    uint8_t *hll = nullptr, *hlr = nullptr;                // per slot
    int32_t *oct_cnt = nullptr, *oct_off = nullptr, *leaf_parent = nullptr, *node_end = nullptr,
            *leaf_oct = nullptr;                           // per slot
    int32_t *oct_base = nullptr, *oct_size = nullptr;      // per cloud
    sx_otnode* oct = nullptr;                              // per octree node
    int2* oct_aux = nullptr;                               // per octree node: (parent, points)
    // host state
    int n_clouds = 0;
    int64_t n_input = 0;
    bool built = false;
    std::vector<sx::BatchGraph*> graphs;  // captured graphs not yet destroyed
    bool octree() const { return oct_mem != nullptr; }
    sx::CloudMap map() const { return sx::CloudMap{tile_cloud, base, size}; }
};

namespace {
size_t batch_ws_bytes(int64_t cap, int64_t slot_cap) {
    // the octree's count scan runs over every slot
    return std::max({sx::sort_ws_bytes(cap), sx::scan_ws_bytes(cap), sx::scan_ws_bytes(slot_cap)});
}

size_t octree_layout(sx_batch* b, void* base) {
    sx::Carver c(base);
    const int64_t s = b->slot_cap, mc = b->max_clouds, o = b->oct_cap;
    b->hll = c.take<uint8_t>(s);
    b->hlr = c.take<uint8_t>(s);
    b->oct_cnt = c.take<int32_t>(s);
    b->oct_off = c.take<int32_t>(s);
    b->leaf_parent = c.take<int32_t>(s);
    b->node_end = c.take<int32_t>(s);
    b->leaf_oct = c.take<int32_t>(s);
    b->oct_base = c.take<int32_t>(mc);
    b->oct_size = c.take<int32_t>(mc);
    b->oct = c.take<sx_otnode>(o);
    b->oct_aux = c.take<int2>(o);
    return c.off;
}

size_t batch_layout(sx_batch* b, void* base) {
    sx::Carver c(base);
    const int64_t n = b->cap, s = b->slot_cap, mc = b->max_clouds;
    b->codes_in = c.take<uint64_t>(n);
    b->codes_a = c.take<uint64_t>(n);
    b->keys_c = c.take<uint64_t>(n);
    b->keys_s = c.take<uint64_t>(n);
    b->codes_s = c.take<uint64_t>(n);
    b->perm_a = c.take<uint32_t>(n);
    b->perm = c.take<uint32_t>(n);
    b->uos = c.take<int32_t>(n);
    b->off = c.take<int64_t>(mc + 1);
    b->bounds = c.take<float>(4 * mc);
    b->bbox_partial = c.take<float>(sx::seg_bbox_ws_bytes(static_cast<int>(mc)) / sizeof(float));
    b->size = c.take<int32_t>(mc);
    b->base = c.take<int32_t>(mc);
    b->tile_cloud = c.take<int32_t>(s / sx::kCloudAlign + 1);
    b->ucodes = c.take<uint64_t>(s);
    b->rep = c.take<uint32_t>(s);
    b->pts16 = c.take<float4>(s);
    b->pn = c.take<uint8_t>(s);
    b->lc = c.take<int32_t>(s);
    b->parent = c.take<int32_t>(s);
    b->queue = c.take<int2>(s);
    b->trav = c.take<sx::TravNode>(s);
    b->counts = c.take<int32_t>(8);
    b->ws = c.take<char>(batch_ws_bytes(n, s));
    return c.off;
}

int check_batch(const sx_batch* b) {
    if (!b) return fail(SX_EINVAL, "sx_batch: null batch");
    if (!b->built) return fail(SX_ESTATE, "sx_batch: build the batch first");
    return SX_OK;
}

// Orders `stream` after the last eager build, whichever stream it ran on.
int after_build(const sx_batch* b, void* stream) {
    SX_HIP(hipStreamWaitEvent(sx::as_stream(stream), b->built_event, 0));
    return SX_OK;
}

int cloud_bits(int n_clouds) {
    int bits = 1;
    while ((1ll << bits) < n_clouds) ++bits;
    return bits;
}

int batch_build(sx_batch* b, const float* x, const float* y, const float* z, hipStream_t st) {
    const int64_t n = b->n_input;
    const int nc = b->n_clouds;
    const int cb = b->code_bits;
    // slot / unique / non-finite counts; the deferred-query total accumulates
    SX_TRY(sx::fill_i32(b->counts, 0, sx::kBDeferred, st));
    SX_TRY(sx::seg_bbox_launch(x, y, z, b->off, nc, b->bounds, b->counts + sx::kBBad, b->bbox_partial, st));
    SX_TRY(sx::seg_encode_launch(x, y, z, n, b->off, nc, b->bounds, cb, b->codes_in, st));
    // (cloud, code, input index) order: by code, then stably by cloud
    SX_TRY(sx::sort_pairs_launch(b->codes_in, nullptr, n, cb, b->codes_a, b->perm_a, b->ws, b->sort_algo, st,
                                 b->counts + sx::kBSortErr));
    sx::cloud_keys_kernel<<<sx::grid_for(n, 256 * sx::kCloudKeysItems), 256, 0, st>>>(b->perm_a, n, b->off, nc,
                                                                                      b->keys_c);
    SX_LAUNCH_CHECK("cloud_keys_kernel");
    SX_TRY(sx::sort_pairs_launch(b->keys_c, b->perm_a, n, cloud_bits(nc), b->keys_s, b->perm, b->ws, b->sort_algo,
                                 st, b->counts + sx::kBSortErr));
    sx::gather_codes_kernel<<<sx::grid_for(n, 256), 256, 0, st>>>(b->codes_in, b->perm, n, b->codes_s);
    SX_LAUNCH_CHECK("gather_codes_kernel");
    SX_TRY(sx::device_scan(sx::BatchUniqueIn{b->keys_s, b->codes_s}, sx::BatchUniqueOut{b->uos}, nullptr, n, nullptr,
                           b->ws, st));
    sx::cloud_table_kernel<<<1, sx::kScanBlock, 0, st>>>(b->off, nc, b->uos, b->size, b->base, b->counts);
    SX_LAUNCH_CHECK("cloud_table_kernel");
    const int64_t max_tiles = b->slot_cap / sx::kCloudAlign + 1;
    sx::tile_cloud_kernel<<<sx::grid_for(max_tiles, 256), 256, 0, st>>>(b->base, nc, b->counts, max_tiles,
                                                                        b->tile_cloud);
    SX_LAUNCH_CHECK("tile_cloud_kernel");
    sx::compact_kernel<<<sx::grid_for(n, 256), 256, 0, st>>>(n, b->keys_s, b->codes_s, b->perm, b->uos, b->off,
                                                              b->base, b->ucodes, b->rep);
    SX_LAUNCH_CHECK("compact_kernel");
    const int32_t* slots = b->counts + sx::kBSlots;
    SX_TRY(sx::seg_decode_launch(b->ucodes, slots, b->slot_cap, b->map(), b->bounds, cb, b->pts16, st));
    if (!b->octree()) {
        SX_TRY(sx::karras_launch(b->ucodes, slots, b->slot_cap, cb, SX_ROOT_TEXTBOOK, b->pn, b->lc, nullptr, nullptr,
                                 b->parent, nullptr, nullptr, nullptr, b->trav, b->pts16, b->bounds, st, b->map()));
        return SX_OK;
    }
    // with octrees: the radix-tree kernel also writes the leaf flags, range
    // ends, leaf parents and octree edge counts
    SX_TRY(sx::karras_launch(b->ucodes, slots, b->slot_cap, cb, SX_ROOT_TEXTBOOK, b->pn, b->lc, b->hll, b->hlr,
                             b->parent, b->node_end, b->leaf_parent, nullptr, b->trav, b->pts16, b->bounds, st,
                             b->map(), b->oct_cnt));
    SX_TRY(sx::batch_octree_launch(b->ucodes, slots, b->slot_cap, cb, b->bounds, b->pn, b->lc, b->hll, b->hlr,
                                   b->oct_cnt, b->oct_off, b->map(), nc, b->oct_base, b->oct_size, b->oct, b->oct_cap,
                                   b->counts + sx::kBOct, b->ws, st));
    SX_TRY(sx::feature_links_launch(slots, b->slot_cap, b->parent, b->leaf_parent, b->node_end, b->oct_off,
                                    b->oct_cnt, b->oct_cap, b->leaf_oct, b->oct_aux, st, b->map()));
    return SX_OK;
}

// Rows are written for slots < min(n_rows, slot count): n_rows is the row
// capacity of the caller's buffers.
// radius2: per-slot squared search radius (learned seeds), or null.
int batch_knn(sx_batch* b, int k, int32_t* out, int64_t n_rows, hipStream_t st, const float* radius2 = nullptr) {
    sx::ParentRef par;
    par.p = b->parent;
    return sx::knn_tiles_launch(b->pts16, b->trav, par, b->counts + sx::kBSlots, 0, n_rows, k, radius2, b->bucket,
                                out, b->queue, b->counts + sx::kBQueue, b->counts + sx::kBDeferred, st, b->map(),
                                b->queue_blocks, nullptr);
}

int batch_normals(sx_batch* b, const int32_t* nn, int k, int precision, float vx, float vy, float vz, float* normals,
                  uint8_t* flags, int64_t n_rows, hipStream_t st) {
    return sx::normals_launch(b->pts16, b->counts + sx::kBSlots, nn, 0, n_rows, k, precision, vx, vy, vz,
                              normals, flags, st, b->map());
}

bool valid_k(int k) { return k == 8 || k == 16 || k == 24 || k == 32; }
}  // namespace

namespace sx {
int batch_feature_arrays(sx_batch* b, BatchFeatures* out, void* stream) {
    SX_TRY(check_batch(b));
    if (!b->octree())
        return fail(SX_ESTATE, "batch features need the per-cloud octrees: call sx_batch_set_octree, then build");
    // ordered after the last build, whichever stream it ran on
    SX_TRY(after_build(b, stream));
    out->trav = b->trav;
    out->aux = b->oct_aux;
    out->pn = b->pn;
    out->leaf_parent = b->leaf_parent;
    out->leaf_oct = b->leaf_oct;
    out->oct = b->oct;
    out->oct_cap = b->oct_cap;
    out->bounds = b->bounds;
    out->slots = b->counts + kBSlots;
    out->cm = b->map();
    out->limit = b->slot_cap;
    return SX_OK;
}
}  // namespace sx

namespace {

}  // namespace

extern "C" {

int sx_batch_create(sx_batch** out, int64_t max_points, int max_clouds, int code_bits) {
    if (!out) return fail(SX_EINVAL, "sx_batch_create: null output");
    *out = nullptr;
    if (max_points < 1 || max_points >= (1ll << 30)) return fail(SX_EINVAL, "sx_batch_create: max_points must be in [1, 2^30)");
    if (max_clouds < 1 || max_clouds > (1 << 16)) return fail(SX_EINVAL, "sx_batch_create: max_clouds must be in [1, 65536]");
    if (code_bits != 63 && code_bits != 30) return fail(SX_EINVAL, "sx_batch_create: code_bits must be 30 or 63");
    auto* b = new sx_batch();
    b->cap = max_points;
    b->max_clouds = max_clouds;
    b->slot_cap = max_points + static_cast<int64_t>(sx::kCloudAlign - 1) * max_clouds;
    b->code_bits = code_bits;
    const size_t bytes = batch_layout(b, nullptr);
    if (hipMalloc(&b->mem, bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return fail(SX_ENOMEM, "sx_batch_create: device allocation failed");
    }
    batch_layout(b, b->mem);
    int rc = sx::fill_i32(b->counts, 0, 8, nullptr);
    if (rc == SX_OK && hipEventCreateWithFlags(&b->built_event, hipEventDisableTiming) != hipSuccess)
        rc = fail(SX_EHIP, "sx_batch_create: event creation failed");
    if (rc == SX_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(SX_EHIP, "sx_batch_create: sync failed");
    if (rc != SX_OK) {
        if (b->built_event) (void)hipEventDestroy(b->built_event);
        (void)hipFree(b->mem);
        delete b;
        return rc;
    }
    *out = b;
    return SX_OK;
}

int sx_batch_destroy(sx_batch* b) {
    if (!b) return SX_OK;
    for (sx::BatchGraph* g : b->graphs) g->owner = nullptr;  // left unusable (see sx_batch_graph_launch)
    if (b->built_event) SX_HIP(hipEventDestroy(b->built_event));
    if (b->oct_mem) SX_HIP(hipFree(b->oct_mem));
    if (b->mem) SX_HIP(hipFree(b->mem));
    delete b;
    return SX_OK;
}

int sx_batch_set_octree(sx_batch* b, float oct_factor) {
    if (!b) return fail(SX_EINVAL, "sx_batch_set_octree: null batch");
    if (!(oct_factor >= 0.0f) || oct_factor > 64.0f)
        return fail(SX_EINVAL, "sx_batch_set_octree: oct_factor must be in [0, 64]");
    if (!b->graphs.empty())
        return fail(SX_ESTATE, "sx_batch_set_octree: destroy the batch's graphs first (they write into the storage "
                               "this call replaces)");
    SX_HIP(hipDeviceSynchronize());  // no build or query may still use the old storage
    if (b->oct_mem) {
        SX_HIP(hipFree(b->oct_mem));
        b->oct_mem = nullptr;
        b->oct_cap = 0;
    }
    b->built = false;
    if (oct_factor == 0.0f) return SX_OK;
    b->oct_cap = std::max<int64_t>(16, static_cast<int64_t>(static_cast<double>(oct_factor) *
                                                            static_cast<double>(b->slot_cap)));
    const size_t bytes = octree_layout(b, nullptr);
    void* mem = nullptr;
    if (hipMalloc(&mem, bytes) != hipSuccess) {
        (void)hipGetLastError();
        b->oct_cap = 0;
        return fail(SX_ENOMEM, "sx_batch_set_octree: device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    b->oct_mem = mem;
    octree_layout(b, mem);
    return SX_OK;
}

int sx_batch_set_clouds(sx_batch* b, const int64_t* h_offsets, int n_clouds, void* stream) {
    if (!b || !h_offsets) return fail(SX_EINVAL, "sx_batch_set_clouds: null argument");
    if (n_clouds < 1 || n_clouds > b->max_clouds) return fail(SX_EINVAL, "sx_batch_set_clouds: bad cloud count");
    if (h_offsets[0] != 0) return fail(SX_EINVAL, "sx_batch_set_clouds: offsets must start at 0");
    for (int c = 0; c < n_clouds; ++c)
        if (h_offsets[c + 1] < h_offsets[c]) return fail(SX_EINVAL, "sx_batch_set_clouds: offsets must not decrease");
    const int64_t n = h_offsets[n_clouds];
    if (n < 1 || n > b->cap) return fail(SX_ERANGE, "sx_batch_set_clouds: point total outside [1, max_points]");
    hipStream_t st = sx::as_stream(stream);
    SX_HIP(hipMemcpyAsync(b->off, h_offsets, sizeof(int64_t) * (n_clouds + 1), hipMemcpyHostToDevice, st));
    SX_HIP(hipStreamSynchronize(st));
    b->n_clouds = n_clouds;
    b->n_input = n;
    b->built = false;
    return SX_OK;
}

int sx_batch_set_tuning(sx_batch* b, int bucket, int sort_algo, int queue_blocks) {
    if (!b) return fail(SX_EINVAL, "sx_batch_set_tuning: null batch");
    if (bucket < 1 || bucket > 256) return fail(SX_EINVAL, "sx_batch_set_tuning: bucket must be in [1, 256]");
    if (sort_algo != SX_SORT_ONESWEEP && sort_algo != SX_SORT_THREE_KERNEL && sort_algo != SX_SORT_HYBRID)
        return fail(SX_EINVAL, "sx_batch_set_tuning: bad sort algorithm");
    if (queue_blocks < 0 || queue_blocks > (1 << 16)) return fail(SX_EINVAL, "sx_batch_set_tuning: bad queue_blocks");
    b->bucket = bucket;
    b->sort_algo = sort_algo;
    b->queue_blocks = queue_blocks;
    return SX_OK;
}

int sx_batch_build(sx_batch* b, const float* d_x, const float* d_y, const float* d_z, void* stream) {
    if (!b) return fail(SX_EINVAL, "sx_batch_build: null batch");
    if (b->n_clouds < 1) return fail(SX_ESTATE, "sx_batch_build: set the clouds first");
    if (!d_x || !d_y || !d_z) return fail(SX_EINVAL, "sx_batch_build: null coordinates");
    SX_TRY(batch_build(b, d_x, d_y, d_z, sx::as_stream(stream)));
    SX_HIP(hipEventRecord(b->built_event, sx::as_stream(stream)));
    b->built = true;
    return SX_OK;
}

int sx_batch_get_view(sx_batch* b, sx_batch_view* view, void* stream) {
    SX_TRY(check_batch(b));
    if (!view) return fail(SX_EINVAL, "sx_batch_get_view: null view");
    hipStream_t st = sx::as_stream(stream);
    SX_HIP(hipEventSynchronize(b->built_event));  // whichever stream the build ran on
    int32_t h[8];
    SX_HIP(hipMemcpyAsync(h, b->counts, sizeof(h), hipMemcpyDeviceToHost, st));
    SX_HIP(hipStreamSynchronize(st));
    if (h[sx::kBSortErr] != 0)
        return fail(SX_EINTERNAL, "sx_batch_get_view: a sort of the build gave up waiting in a look-back; "
                                  "the batch is empty and must be rebuilt");
    if (h[sx::kBBad] != 0) return fail(SX_EDATA, "sx_batch_get_view: non-finite input coordinates");
    if (b->octree() && (h[sx::kBOct] < 0 || h[sx::kBOct] > b->oct_cap))
        return fail(SX_ERANGE, "sx_batch_get_view: the octrees need " + std::to_string(h[sx::kBOct]) +
                                   " nodes, capacity is " + std::to_string(b->oct_cap) + " (raise oct_factor)");
    view->n_clouds = b->n_clouds;
    view->n_input = b->n_input;
    view->n_slots = h[sx::kBSlots];
    view->n_unique = h[sx::kBUnique];
    view->cloud_offsets = b->off;
    view->cloud_base = b->base;
    view->cloud_size = b->size;
    view->cloud_bounds = b->bounds;
    view->codes = b->ucodes;
    view->rep = b->rep;
    view->points = reinterpret_cast<const float*>(b->pts16);
    view->prefix_n = b->pn;
    view->left_child = b->lc;
    view->parent = b->parent;
    const bool oct = b->octree();
    view->n_oct = oct ? h[sx::kBOct] : 0;
    view->oct_nodes = oct ? b->oct : nullptr;
    view->cloud_oct_base = oct ? b->oct_base : nullptr;
    view->cloud_oct_size = oct ? b->oct_size : nullptr;
    view->has_leaf_left = oct ? b->hll : nullptr;
    view->has_leaf_right = oct ? b->hlr : nullptr;
    return SX_OK;
}

int sx_batch_knn(sx_batch* b, int k, int32_t* d_out, int64_t n_rows, void* stream) {
    SX_TRY(check_batch(b));
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_batch_knn: k must be 8, 16, 24 or 32");
    if (!d_out || reinterpret_cast<uintptr_t>(d_out) % 16 != 0)
        return fail(SX_EINVAL, "sx_batch_knn: output must be a 16-byte aligned device buffer");
    if (n_rows < 0 || n_rows > b->slot_cap) return fail(SX_EINVAL, "sx_batch_knn: n_rows outside [0, slot capacity]");
    SX_TRY(after_build(b, stream));
    return batch_knn(b, k, d_out, n_rows, sx::as_stream(stream));
}

int sx_batch_knn_radius(sx_batch* b, int k, const float* d_radius2, int32_t* d_out, int64_t n_rows, void* stream) {
    SX_TRY(check_batch(b));
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_batch_knn_radius: k must be 8, 16, 24 or 32");
    if (!d_radius2) return fail(SX_EINVAL, "sx_batch_knn_radius: null radius2");
    if (!d_out || reinterpret_cast<uintptr_t>(d_out) % 16 != 0)
        return fail(SX_EINVAL, "sx_batch_knn_radius: output must be a 16-byte aligned device buffer");
    if (n_rows < 0 || n_rows > b->slot_cap)
        return fail(SX_EINVAL, "sx_batch_knn_radius: n_rows outside [0, slot capacity]");
    SX_TRY(after_build(b, stream));
    return batch_knn(b, k, d_out, n_rows, sx::as_stream(stream), d_radius2);
}

int sx_batch_normals(sx_batch* b, const int32_t* d_nn, int k, int precision, float vx, float vy, float vz,
                     float* d_normals, uint8_t* d_flags, int64_t n_rows, void* stream) {
    SX_TRY(check_batch(b));
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_batch_normals: k must be 8, 16, 24 or 32");
    if (!d_nn || !d_normals) return fail(SX_EINVAL, "sx_batch_normals: null buffer");
    if (n_rows < 0 || n_rows > b->slot_cap)
        return fail(SX_EINVAL, "sx_batch_normals: n_rows outside [0, slot capacity]");
    SX_TRY(after_build(b, stream));
    return batch_normals(b, d_nn, k, precision, vx, vy, vz, d_normals, d_flags, n_rows, sx::as_stream(stream));
}

int sx_batch_deferred(sx_batch* b, int64_t* deferred, int reset, void* stream) {
    if (!b || !deferred) return fail(SX_EINVAL, "sx_batch_deferred: null argument");
    hipStream_t st = sx::as_stream(stream);
    int32_t h = 0;
    SX_HIP(hipMemcpyAsync(&h, b->counts + sx::kBDeferred, 4, hipMemcpyDeviceToHost, st));
    SX_HIP(hipStreamSynchronize(st));
    *deferred = h;
    if (reset) SX_TRY(sx::fill_i32(b->counts + sx::kBDeferred, 0, 1, st));
    return SX_OK;
}

int sx_batch_graph_create(sx_batch* b, const float* d_x, const float* d_y, const float* d_z, int k, int32_t* d_nn,
                          float* d_normals, int64_t n_rows, void** graph, void* stream) {
    if (!b || !graph) return fail(SX_EINVAL, "sx_batch_graph_create: null argument");
    *graph = nullptr;
    if (b->n_clouds < 1) return fail(SX_ESTATE, "sx_batch_graph_create: set the clouds first");
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_batch_graph_create: k must be 8, 16, 24 or 32");
    if (!d_x || !d_y || !d_z || !d_nn || !d_normals || reinterpret_cast<uintptr_t>(d_nn) % 16 != 0)
        return fail(SX_EINVAL, "sx_batch_graph_create: null or misaligned buffer");
    if (n_rows < 0 || n_rows > b->slot_cap)
        return fail(SX_EINVAL, "sx_batch_graph_create: n_rows outside [0, slot capacity]");
    hipStream_t st = sx::as_stream(stream);
    if (!st) return fail(SX_EINVAL, "sx_batch_graph_create: capture needs a created stream, not the null stream");
    auto* g = new sx::BatchGraph();
    SX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = batch_build(b, d_x, d_y, d_z, st);
    if (rc == SX_OK) rc = batch_knn(b, k, d_nn, n_rows, st);
    if (rc == SX_OK) rc = batch_normals(b, d_nn, k, SX_PREC_FP32, 0.0f, 0.0f, 0.0f, d_normals, nullptr, n_rows, st);
    hipError_t e = hipStreamEndCapture(st, &g->graph);
    if (rc == SX_OK && e != hipSuccess) rc = fail(SX_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (rc == SX_OK) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fail(SX_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    if (rc != SX_OK) {
        (void)hipGetLastError();
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc;
    }
    b->built = true;
    g->owner = b;
    b->graphs.push_back(g);
    *graph = g;
    return SX_OK;
}

int sx_batch_graph_launch(void* graph, void* stream) {
    if (!graph) return fail(SX_EINVAL, "sx_batch_graph_launch: null graph");
    auto* g = static_cast<sx::BatchGraph*>(graph);
    if (!g->owner) return fail(SX_ESTATE, "sx_batch_graph_launch: the graph's batch was destroyed");
    SX_HIP(hipGraphLaunch(g->exec, sx::as_stream(stream)));
    return SX_OK;
}

int sx_batch_graph_destroy(void* graph) {
    if (!graph) return SX_OK;
    auto* g = static_cast<sx::BatchGraph*>(graph);
    if (g->owner) {
        auto& v = g->owner->graphs;
        v.erase(std::remove(v.begin(), v.end(), g), v.end());
    }
    if (g->exec) SX_HIP(hipGraphExecDestroy(g->exec));
    if (g->graph) SX_HIP(hipGraphDestroy(g->graph));
    delete g;
    return SX_OK;
}

// Times the kNN of the built batch for every (bucket, deferred-worker grid)
// pair in the candidate lists (reps runs each after one warm-up), keeps the
// fastest on the batch and reports it.  ms (nullable) receives n_buckets x
// n_queue times in milliseconds, row-major.
int sx_batch_autotune(sx_batch* b, int k, int32_t* d_scratch_nn, int64_t n_rows, const int* h_buckets, int n_buckets,
                      const int* h_queue_blocks, int n_queue, int reps, float* h_ms, int* best_bucket,
                      int* best_queue_blocks, void* stream) {
    SX_TRY(check_batch(b));
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_batch_autotune: k must be 8, 16, 24 or 32");
    if (!d_scratch_nn || !h_buckets || !h_queue_blocks || n_buckets < 1 || n_queue < 1 || reps < 1 || n_rows < 0 ||
        n_rows > b->slot_cap)
        return fail(SX_EINVAL, "sx_batch_autotune: bad argument");
    SX_TRY(after_build(b, stream));
    hipStream_t st = sx::as_stream(stream);
    hipEvent_t e0, e1;
    SX_HIP(hipEventCreate(&e0));
    SX_HIP(hipEventCreate(&e1));
    const int keep_bucket = b->bucket, keep_queue = b->queue_blocks;
    float best = 1e30f;
    int bb = keep_bucket, bq = keep_queue;
    int rc = SX_OK;
    for (int i = 0; i < n_buckets && rc == SX_OK; ++i) {
        for (int j = 0; j < n_queue && rc == SX_OK; ++j) {
            rc = sx_batch_set_tuning(b, h_buckets[i], b->sort_algo, h_queue_blocks[j]);
            if (rc == SX_OK) rc = batch_knn(b, k, d_scratch_nn, n_rows, st);  // warm-up
            float ms = 0.0f;
            if (rc == SX_OK && hipEventRecord(e0, st) != hipSuccess) rc = fail(SX_EHIP, "sx_batch_autotune: event");
            for (int r = 0; r < reps && rc == SX_OK; ++r) rc = batch_knn(b, k, d_scratch_nn, n_rows, st);
            if (rc == SX_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = fail(SX_EHIP, "sx_batch_autotune: timing failed");
            ms /= static_cast<float>(reps);
            if (h_ms) h_ms[i * n_queue + j] = ms;
            if (rc == SX_OK && ms < best) {
                best = ms;
                bb = h_buckets[i];
                bq = h_queue_blocks[j];
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != SX_OK) {
        b->bucket = keep_bucket;
        b->queue_blocks = keep_queue;
        return rc;
    }
    b->bucket = bb;
    b->queue_blocks = bq;
    if (best_bucket) *best_bucket = bb;
    if (best_queue_blocks) *best_queue_blocks = bq;
    return SX_OK;
}

}  // extern "C"
