// sx_index: one cloud's spatial index, built by chaining every stage on one
// stream with device-resident counts (no host synchronisation until a caller
// asks for the view).  Replaces the RT::RadixTree and OT::Octree constructors
// (src/RadixTree.cu:305-382, src/Octree.cu:213-319), which synchronise after
// every kernel and read counts through managed memory.
//
// HBM layout (U = unique points, O = octree nodes, capacity-sized):
//   codes_in[N] u64 -> (sort) codes_sorted[N] u64 + perm[N] u32
//   ucodes[U] u64, rep[U] u32, unique_of_input[N] i32
//   points12[3U] f32 (API, Octree::u_points), points16[U] float4 (kernels)
//   RT::Nodes arrays: prefix_n/has_leaf_left/has_leaf_right[U] u8,
//                     left_child/parent[U] i32
//   trav[U] TravNode (64 B: child boxes + range/split/parent) for kNN
//   oct[O] sx_otnode (56 B, Octree::u_nodes)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "sx_internal.hpp"

namespace sx {
// launchers from the other translation units
size_t bbox_ws_bytes();
int bbox_launch(const float* x, const float* y, const float* z, int64_t n, float* bounds, int32_t* d_bad, void* ws,
                hipStream_t st, int32_t* clear = nullptr, int n_clear = 0);
int pack_points12_launch(const float4* pts16, const int32_t* d_count, int64_t n_max, float* pts12, hipStream_t st);
int encode_launch(const float* x, const float* y, const float* z, int64_t n, const float* bounds, int code_bits,
                  uint64_t* codes, hipStream_t st);
int decode_launch(const uint64_t* codes, const int32_t* d_count, int64_t n_max, const float* bounds, int code_bits,
                  float* pts12, float4* pts16, hipStream_t st);
size_t sort_ws_bytes(int64_t n);
int sort_pairs_launch(const uint64_t* keys_in, const uint32_t* vals_in, int64_t n, int key_bits, uint64_t* keys_out,
                      uint32_t* vals_out, void* ws, int algo, hipStream_t st, int32_t* err);
int unique_launch(const uint64_t* sorted, const uint32_t* vals, int64_t n, uint64_t* uniq, uint32_t* rep,
                  int32_t* uos, int32_t* uoi, int32_t* d_count, void* ws, hipStream_t st, const int32_t* sort_err);
int karras_launch(const uint64_t* codes, const int32_t* d_count, int64_t n_max, int code_bits, int root_mode,
                  uint8_t* prefix_n, int32_t* left_child, uint8_t* hll, uint8_t* hlr, int32_t* parent,
                  int32_t* node_end, int32_t* leaf_parent, int2* up, TravNode* trav, const float4* pts,
                  const float* bounds, hipStream_t st, CloudMap cm = CloudMap(), int32_t* oct_count = nullptr,
                  bool claimed_parents = false);
size_t octree_ws_bytes(int64_t n_max);
int octree_launch(const uint64_t* codes, const int32_t* d_count, int64_t n_max, int code_bits, const float* bounds,
                  const uint8_t* pn, const int32_t* left_child, const uint8_t* hll, const uint8_t* hlr,
                  const int32_t* parent, sx_otnode* nodes, int64_t cap, int32_t* d_oct_count, int32_t* leaf_oct,
                  void* ws, hipStream_t st, int32_t* keep_offsets = nullptr, int32_t* keep_counts = nullptr,
                  bool counts_given = false);
int feature_links_launch(const int32_t* d_count, int64_t n_max, const int32_t* parent, const int32_t* leaf_parent,
                         const int32_t* node_end, const int32_t* offsets, const int32_t* counts, int64_t cap,
                         int32_t* leaf_oct, int2* aux, hipStream_t st, CloudMap cm = CloudMap());
int knn_launch(const float4* pts, const TravNode* nodes, ParentRef par, const int32_t* d_counts, int64_t q_start,
               int64_t n_q, const float* queries, int k, const float* radius2, int bucket, int32_t* out, int32_t* reruns,
               hipStream_t st, CloudMap cm = CloudMap());
int knn_tiles_launch(const float4* pts, const TravNode* nodes, ParentRef par, const int32_t* d_counts, int64_t q_start,
                     int64_t n_q, int k, const float* radius2, int bucket, int32_t* out, int2* queue, int32_t* queue_count,
                     int32_t* deferred_total, hipStream_t st, CloudMap cm = CloudMap(), int queue_blocks = 0,
                     int32_t* cand_total = nullptr);
int scatter_i32_launch(const int32_t* src, const uint32_t* idx, int64_t n, int32_t* dst, hipStream_t st);
int normals_launch(const float4* pts, const int32_t* d_counts, const int32_t* nn, int64_t q_start, int64_t n_q, int k,
                   int precision, float vx, float vy, float vz, float* normals, uint8_t* flags, hipStream_t st,
                   CloudMap cm = CloudMap());
}  // namespace sx

// counts[] slots: unique points, octree nodes, non-finite inputs, the sort's
// look-back error flag (the first four are cleared by every build), queries
// the kNN deferred to the per-lane search (cumulative), the last call's queue
// size, candidates the tile search broadcast (cumulative, when counting is on)
enum { kCntUnique = 0, kCntOct = 1, kCntBad = 2, kCntSortErr = 3, kCntReruns = 4, kCntQueue = 5, kCntCand = 6 };



struct sx_index {
    int64_t cap = 0;
    int64_t oct_cap = 0;
    int code_bits = 63;
    int root_mode = SX_ROOT_TEXTBOOK;
    int bucket = 32;  // kNN leaf-range size (8M scene: 2.78 ms at 16, 2.72 at 32; gmm16m equal)
    int knn_algo = SX_KNN_ALGO_TILES;
    int sort_algo = SX_SORT_HYBRID;
    bool knn_stats = false;  // count the candidates the kNN tiles broadcast
    void* mem = nullptr;
    float* bounds = nullptr;
    int32_t* counts = nullptr;
    uint64_t* codes_in = nullptr;
    uint64_t* codes_sorted = nullptr;
    uint64_t* ucodes = nullptr;
    uint32_t* perm = nullptr;
    uint32_t* rep = nullptr;
    int32_t* uos = nullptr;
    int32_t* uoi = nullptr;
    float* pts12 = nullptr;
    float4* pts16 = nullptr;
    uint8_t* pn = nullptr;
    uint8_t* hll = nullptr;
    uint8_t* hlr = nullptr;
    int32_t* lc = nullptr;
    int32_t* parent = nullptr;
    int32_t* leaf_parent = nullptr;
    int2* up = nullptr;  // (parent, prefix length) per traversal-tree node, reference root only
    int32_t* leaf_oct = nullptr;
    int2* oct_aux = nullptr;  // (parent octree node, points in the cell) per octree node
    int32_t* node_end = nullptr;  // the far end of each radix node's leaf range
    int32_t* oct_off = nullptr;   // first octree node of each radix node's chain
    int32_t* oct_cnt = nullptr;   // its chain length
    bool feat_valid = false;      // leaf_oct / oct_aux hold the last build's links
    int2* knn_queue = nullptr;  // deferred queries: (index, k-th distance bound bits)
    sx::TravNode* trav = nullptr;
    sx_otnode* oct = nullptr;
    void* ws_bbox = nullptr;
    void* ws = nullptr;
    int32_t* nn_buf = nullptr;
    int64_t nn_buf_len = 0;
    // host state
    int64_t n_input = 0;
    bool built = false;
    bool from_codes = false;  // the last build took sorted codes (sx_index_build_codes)
    bool uoi_valid = false;   // unique_of_input holds the last build's map
    bool pts12_valid = false; // the 12-byte point copy holds the last build's points
    bool counts_valid = false;
    hipEvent_t built_event = nullptr;  // recorded after every build
    int64_t n_unique = 0;
    int64_t n_oct = 0;
    float host_bounds[3] = {0, 0, 0};
};

namespace {

size_t index_layout(sx_index* ix, void* base) {
    sx::Carver c(base);
    const size_t n = static_cast<size_t>(ix->cap);
    ix->bounds = c.take<float>(4);
    ix->counts = c.take<int32_t>(8);
    ix->codes_in = c.take<uint64_t>(n);
    ix->codes_sorted = c.take<uint64_t>(n);
    ix->ucodes = c.take<uint64_t>(n);
    ix->perm = c.take<uint32_t>(n);
    ix->rep = c.take<uint32_t>(n);
    ix->uos = c.take<int32_t>(n);
    ix->uoi = c.take<int32_t>(n);
    ix->pts12 = c.take<float>(3 * n);
    ix->pts16 = c.take<float4>(n);
    ix->pn = c.take<uint8_t>(n);
    ix->hll = c.take<uint8_t>(n);
    ix->hlr = c.take<uint8_t>(n);
    ix->lc = c.take<int32_t>(n);
    ix->parent = c.take<int32_t>(n);
    ix->leaf_parent = c.take<int32_t>(n);
    ix->up = c.take<int2>(n);
    ix->leaf_oct = c.take<int32_t>(n);
    ix->knn_queue = c.take<int2>(n);
    ix->trav = c.take<sx::TravNode>(n);
    ix->oct = c.take<sx_otnode>(static_cast<size_t>(ix->oct_cap));
    ix->oct_aux = c.take<int2>(static_cast<size_t>(ix->oct_cap));
    ix->node_end = c.take<int32_t>(n);
    ix->oct_off = c.take<int32_t>(n);
    ix->oct_cnt = c.take<int32_t>(n);
    ix->ws_bbox = c.take<char>(sx::bbox_ws_bytes());
    const size_t ws = std::max({sx::sort_ws_bytes(ix->cap), sx::scan_ws_bytes(ix->cap), sx::octree_ws_bytes(ix->cap)});
    ix->ws = c.take<char>(ws);
    return c.off;
}

// Parents of the traversal tree: the radix tree's own array with the
// textbook root; with the reference root the traversal tree is the separate
// textbook tree, whose parents are the .x of its (parent, prefix) pairs.
sx::ParentRef trav_parents(const sx_index* ix) {
    sx::ParentRef r;
    if (ix->root_mode == SX_ROOT_TEXTBOOK) {
        r.p = ix->parent;
        r.stride = 1;
    } else {
        r.p = reinterpret_cast<const int32_t*>(ix->up);
        r.stride = 2;
    }
    return r;
}

int check_index(const sx_index* ix) {
    if (!ix) return sx::fail(SX_EINVAL, "null index");
    if (!ix->built) return sx::fail(SX_ESTATE, "index has not been built");
    return SX_OK;
}

// Query positions are checked against the unique count once a view has read
// it, else against the input count; the kernels give positions at or beyond
// the unique count rows of -1 (neighbours) or NaN (normals).
int check_queries(const sx_index* ix, int64_t q_start, int64_t n_q) {
    const int64_t limit = ix->counts_valid ? ix->n_unique : ix->n_input;
    if (q_start < 0 || n_q < 0 || q_start + n_q > limit)
        return sx::fail(SX_EINVAL, "query range [" + std::to_string(q_start) + ", " + std::to_string(q_start + n_q) +
                                       ") outside the index points (" + std::to_string(limit) + ")");
    return SX_OK;
}

bool valid_k(int k) { return k == 8 || k == 16 || k == 24 || k == 32; }

// unique_of_input of the last build (scattered from the sorted positions the
// first time a caller needs it).
int input_order_map(sx_index* ix, hipStream_t st) {
    if (ix->uoi_valid) return SX_OK;
    SX_TRY(sx::scatter_i32_launch(ix->uos, ix->perm, ix->n_input, ix->uoi, st));
    ix->uoi_valid = true;
    return SX_OK;
}

// The 12-byte point copy of the last build (written the first time a view
// needs it; the searches read the float4 copy).
int points12(sx_index* ix, hipStream_t st) {
    if (ix->pts12_valid) return SX_OK;
    SX_TRY(sx::pack_points12_launch(ix->pts16, ix->counts + kCntUnique, ix->n_input, ix->pts12, st));
    ix->pts12_valid = true;
    return SX_OK;
}

// Orders `stream` after the last build, whichever stream it ran on.
int after_build(const sx_index* ix, void* stream) {
    SX_HIP(hipStreamWaitEvent(sx::as_stream(stream), ix->built_event, 0));
    return SX_OK;
}

}  // namespace

namespace sx {
// Arrays the learned-radius feature kernel reads (mlp.hip).
int index_feature_arrays(sx_index* ix, const TravNode** trav, const int2** up, const int2** oct_aux,
                         const uint8_t** pn, const int32_t** leaf_parent, const int32_t** leaf_oct,
                         const sx_otnode** oct, int64_t* oct_cap, const float** bounds, const int32_t** counts,
                         int64_t* limit, void* stream) {
    SX_TRY(check_index(ix));
    SX_TRY(after_build(ix, stream));
    // textbook root: the octree links give the counts and the API's prefix
    // array the depth; otherwise the traversal tree's (parent, prefix) pairs
    const bool textbook = ix->root_mode == SX_ROOT_TEXTBOOK;
    if (!ix->feat_valid) {
        SX_TRY(feature_links_launch(ix->counts + kCntUnique, ix->n_input, ix->parent, ix->leaf_parent, ix->node_end,
                                    ix->oct_off, ix->oct_cnt, ix->oct_cap, ix->leaf_oct, ix->oct_aux,
                                    as_stream(stream)));
        ix->feat_valid = true;
    }
    *trav = ix->trav;
    *up = textbook ? nullptr : ix->up;
    *oct_aux = textbook ? ix->oct_aux : nullptr;
    *pn = textbook ? ix->pn : nullptr;
    *leaf_parent = ix->leaf_parent;
    *leaf_oct = ix->leaf_oct;
    *oct = ix->oct;
    *oct_cap = ix->oct_cap;
    *bounds = ix->bounds;
    *counts = ix->counts;
    *limit = ix->counts_valid ? ix->n_unique : ix->n_input;
    return SX_OK;
}
}  // namespace sx

// Neighbour rows of sx_estimate_normals_noalloc: a file-scope device array,
// one copy per GPU for the life of the process (64 MiB), shared by every
// call on that GPU.  Host code never passes the symbol itself to a kernel:
// it passes the device address hipGetSymbolAddress gives for the current
// device (noalloc_rows below).
constexpr int64_t kNoAllocEntries = int64_t(1) << 24;  // n_q * k int32 entries at most
static __device__ __align__(256) int32_t scratch[kNoAllocEntries];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (no allocation).
int noalloc_rows(int32_t** out) {
    int dev = 0;
    SX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return sx::fail(SX_EINVAL, "sx_estimate_normals_noalloc: device id too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return sx::fail(SX_EHIP, std::string("sx_estimate_normals_noalloc: hipGetSymbolAddress failed: ") +
                                         hipGetErrorString(e));
        g_scratch[dev] = static_cast<int32_t*>(ptr);
    }
    *out = g_scratch[dev];
    return SX_OK;
}
}  // namespace

namespace sx {

// Build from given codes: copy them, identity permutation and
// representatives, and count codes that are not strictly increasing or do
// not fit in code_bits into *bad.
__global__ void __launch_bounds__(256) codes_in_kernel(const uint64_t* __restrict__ in, int64_t n, int code_bits,
                                                       uint64_t* __restrict__ ucodes, uint32_t* __restrict__ perm,
                                                       uint32_t* __restrict__ rep, int32_t* __restrict__ uos,
                                                       int32_t* bad) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool wrong = false;
    if (i < n) {
        const uint64_t c = in[i];
        wrong = (c >> code_bits) != 0 || (i > 0 && in[i - 1] >= c);
        ucodes[i] = c;
        perm[i] = static_cast<uint32_t>(i);
        rep[i] = static_cast<uint32_t>(i);
        uos[i] = static_cast<int32_t>(i);
    }
    const uint64_t m = __ballot(wrong);
    if (m && lane_id() == static_cast<unsigned>(__ffsll(static_cast<unsigned long long>(m)) - 1))
        atomicAdd(bad, __popcll(m));
}

// The cube and the unique count of a build from codes; no points when a code
// was rejected (every later stage is sized by the count).
__global__ void codes_count_kernel(int64_t n, float mn, float range, float* bounds, int32_t* counts, int bad_slot) {
    bounds[0] = mn;
    bounds[1] = __fadd_rn(mn, range);
    bounds[2] = range;
    counts[0] = counts[bad_slot] ? 0 : static_cast<int32_t>(n);
}

// dst row r = src row map[r] (rows of `words` 32-bit words).
__global__ void __launch_bounds__(256) gather_rows_kernel(const uint32_t* __restrict__ src, int64_t words,
                                                          const int32_t* __restrict__ map, int64_t n_rows,
                                                          uint32_t* __restrict__ dst) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_rows * words) return;
    const int64_t r = i / words, w = i - r * words;
    dst[i] = src[static_cast<int64_t>(map[r]) * words + w];
}

}  // namespace sx

namespace {
// Decode, radix tree (+ traversal boxes) and octree of the unique codes,
// sized by counts[kCntUnique] (capacity n).
int build_trees(sx_index* ix, int64_t n, hipStream_t st) {
    const int cb = ix->code_bits;
    // the 12-byte point copy (API layout) is written when a view asks for it
    SX_TRY(sx::decode_launch(ix->ucodes, ix->counts + kCntUnique, n, ix->bounds, cb, nullptr, ix->pts16, st));
    ix->pts12_valid = false;
    if (ix->root_mode == SX_ROOT_TEXTBOOK) {
        SX_TRY(sx::karras_launch(ix->ucodes, ix->counts + kCntUnique, n, cb, SX_ROOT_TEXTBOOK, ix->pn, ix->lc, ix->hll,
                                 ix->hlr, ix->parent, ix->node_end, ix->leaf_parent, nullptr, ix->trav, ix->pts16,
                                 ix->bounds, st, sx::CloudMap(), ix->oct_cnt, true));
    } else {
        SX_TRY(sx::karras_launch(ix->ucodes, ix->counts + kCntUnique, n, cb, SX_ROOT_REFERENCE, ix->pn, ix->lc,
                                 ix->hll, ix->hlr, ix->parent, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                 st));
        SX_TRY(sx::karras_launch(ix->ucodes, ix->counts + kCntUnique, n, cb, SX_ROOT_TEXTBOOK, nullptr, nullptr,
                                 nullptr, nullptr, nullptr, nullptr, ix->leaf_parent, ix->up, ix->trav, ix->pts16,
                                 ix->bounds, st));
    }
    // With the textbook root the octree and the traversal tree are one tree:
    // the feature links (leaf -> octree node, each octree node's parent and
    // point count) are derived from the kept chain offsets when features are
    // first asked for.  With the reference root the gather places each leaf.
    const bool textbook = ix->root_mode == SX_ROOT_TEXTBOOK;
    SX_TRY(sx::octree_launch(ix->ucodes, ix->counts + kCntUnique, n, cb, ix->bounds, ix->pn, ix->lc, ix->hll, ix->hlr,
                             ix->parent, ix->oct, ix->oct_cap, ix->counts + kCntOct,
                             textbook ? nullptr : ix->leaf_oct, ix->ws, st, ix->oct_off, ix->oct_cnt,
                             textbook));
    ix->feat_valid = !textbook;
    return SX_OK;
}

}  // namespace

using sx::fail;

extern "C" {

int sx_index_create(sx_index** out, int64_t capacity, int code_bits, int root_mode, float oct_factor) {
    if (!out) return fail(SX_EINVAL, "sx_index_create: null out");
    *out = nullptr;
    if (capacity <= 0 || capacity >= (1ll << 31) - 1) return fail(SX_EINVAL, "sx_index_create: capacity must be in [1, 2^31-1)");
    if (code_bits != 30 && code_bits != 63) return fail(SX_EINVAL, "sx_index_create: code_bits must be 30 or 63");
    if (root_mode != SX_ROOT_TEXTBOOK && root_mode != SX_ROOT_REFERENCE)
        return fail(SX_EINVAL, "sx_index_create: bad root_mode");
    sx_index* ix = new sx_index();
    ix->cap = capacity;
    ix->code_bits = code_bits;
    ix->root_mode = root_mode;
    const double f = oct_factor > 0 ? oct_factor : 3.0;
    ix->oct_cap = std::max<int64_t>(16, static_cast<int64_t>(f * static_cast<double>(capacity)));
    const size_t bytes = index_layout(ix, nullptr);
    hipError_t e = hipMalloc(&ix->mem, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        delete ix;
        return fail(SX_ENOMEM, "sx_index_create: hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    index_layout(ix, ix->mem);
    int rc = sx::fill_u8(ix->bounds, 0, 256, nullptr);                    // bounds + counts
    if (rc == SX_OK && hipEventCreateWithFlags(&ix->built_event, hipEventDisableTiming) != hipSuccess)
        rc = fail(SX_EHIP, "sx_index_create: event creation failed");
    if (rc == SX_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(SX_EHIP, "sx_index_create: sync failed");
    if (rc != SX_OK) {
        if (ix->built_event) (void)hipEventDestroy(ix->built_event);
        (void)hipFree(ix->mem);
        delete ix;
        return rc;
    }
    *out = ix;
    return SX_OK;
}

int sx_index_destroy(sx_index* ix) {
    if (!ix) return SX_OK;
    int rc = SX_OK;
    if (ix->nn_buf && hipFree(ix->nn_buf) != hipSuccess) rc = fail(SX_EHIP, "sx_index_destroy: hipFree failed");
    if (ix->mem && hipFree(ix->mem) != hipSuccess) rc = fail(SX_EHIP, "sx_index_destroy: hipFree failed");
    if (ix->built_event && hipEventDestroy(ix->built_event) != hipSuccess)
        rc = fail(SX_EHIP, "sx_index_destroy: hipEventDestroy failed");
    delete ix;
    return rc;
}

int sx_index_set_bucket(sx_index* ix, int bucket) {
    if (!ix) return fail(SX_EINVAL, "sx_index_set_bucket: null index");
    if (bucket < 1 || bucket > 256) return fail(SX_EINVAL, "sx_index_set_bucket: bucket must be in [1, 256]");
    ix->bucket = bucket;
    return SX_OK;
}

int sx_index_set_algorithms(sx_index* ix, int knn_algo, int sort_algo) {
    if (!ix) return fail(SX_EINVAL, "sx_index_set_algorithms: null index");
    if (knn_algo != SX_KNN_ALGO_TILES && knn_algo != SX_KNN_ALGO_LANE)
        return fail(SX_EINVAL, "sx_index_set_algorithms: bad kNN algorithm");
    if (sort_algo != SX_SORT_ONESWEEP && sort_algo != SX_SORT_THREE_KERNEL && sort_algo != SX_SORT_HYBRID)
        return fail(SX_EINVAL, "sx_index_set_algorithms: bad sort algorithm");
    ix->knn_algo = knn_algo;
    ix->sort_algo = sort_algo;
    return SX_OK;
}

int sx_index_build(sx_index* ix, const float* d_x, const float* d_y, const float* d_z, int64_t n, void* stream) {
    if (!ix) return fail(SX_EINVAL, "sx_index_build: null index");
    if (n <= 0 || n > ix->cap) return fail(SX_EINVAL, "sx_index_build: n must be in [1, capacity]");
    if (!d_x || !d_y || !d_z) return fail(SX_EINVAL, "sx_index_build: null coordinates");
    hipStream_t st = sx::as_stream(stream);
    ix->built = false;
    ix->counts_valid = false;
    ix->from_codes = false;
    ix->uoi_valid = false;
    const int cb = ix->code_bits;
    // unique / octree / non-finite counts and the sort flag are cleared by the
    // bounds kernel; the deferred-query total accumulates across builds until
    // sx_index_knn_reruns resets it
    SX_TRY(sx::bbox_launch(d_x, d_y, d_z, n, ix->bounds, ix->counts + kCntBad, ix->ws_bbox, st, ix->counts,
                           kCntReruns));
    SX_TRY(sx::encode_launch(d_x, d_y, d_z, n, ix->bounds, cb, ix->codes_in, st));
    SX_TRY(sx::sort_pairs_launch(ix->codes_in, nullptr, n, cb, ix->codes_sorted, ix->perm, ix->ws, ix->sort_algo, st,
                                 ix->counts + kCntSortErr));
    // a failed sort leaves the unique count at 0: no later stage runs on it
    SX_TRY(sx::unique_launch(ix->codes_sorted, ix->perm, n, ix->ucodes, ix->rep, ix->uos, nullptr,
                             ix->counts + kCntUnique, ix->ws, st, ix->counts + kCntSortErr));
    SX_TRY(build_trees(ix, n, st));
    ix->n_input = n;
    SX_HIP(hipEventRecord(ix->built_event, sx::as_stream(stream)));
    ix->built = true;
    return SX_OK;
}

int sx_index_build_codes(sx_index* ix, const uint64_t* d_codes, int64_t n, float min_coord, float range,
                         void* stream) {
    if (!ix) return fail(SX_EINVAL, "sx_index_build_codes: null index");
    if (n <= 0 || n > ix->cap) return fail(SX_EINVAL, "sx_index_build_codes: n must be in [1, capacity]");
    if (!d_codes) return fail(SX_EINVAL, "sx_index_build_codes: null codes");
    if (!std::isfinite(min_coord) || !std::isfinite(range) || range < 0.0f)
        return fail(SX_EINVAL, "sx_index_build_codes: the cube must be finite with range >= 0");
    hipStream_t st = sx::as_stream(stream);
    ix->built = false;
    ix->counts_valid = false;
    ix->from_codes = true;
    ix->uoi_valid = false;
    SX_TRY(sx::fill_i32(ix->counts, 0, kCntReruns, st));
    sx::codes_in_kernel<<<sx::grid_for(n, 256), 256, 0, st>>>(d_codes, n, ix->code_bits, ix->ucodes, ix->perm,
                                                               ix->rep, ix->uos, ix->counts + kCntBad);
    SX_LAUNCH_CHECK("codes_in_kernel");
    sx::codes_count_kernel<<<1, 1, 0, st>>>(n, min_coord, range, ix->bounds, ix->counts, kCntBad);
    SX_LAUNCH_CHECK("codes_count_kernel");
    SX_TRY(build_trees(ix, n, st));
    ix->n_input = n;
    SX_HIP(hipEventRecord(ix->built_event, st));
    ix->built = true;
    return SX_OK;
}

int sx_index_get_view(sx_index* ix, sx_index_view* view, void* stream) {
    SX_TRY(check_index(ix));
    if (!view) return fail(SX_EINVAL, "sx_index_get_view: null view");
    SX_TRY(after_build(ix, stream));
    SX_TRY(input_order_map(ix, sx::as_stream(stream)));
    SX_TRY(points12(ix, sx::as_stream(stream)));
    int32_t flags[2] = {0, 0};  // non-finite inputs (rejected codes), sort error
    SX_HIP(hipMemcpyAsync(flags, ix->counts + kCntBad, sizeof(flags), hipMemcpyDeviceToHost, sx::as_stream(stream)));
    SX_HIP(hipStreamSynchronize(sx::as_stream(stream)));
    if (flags[1])
        return fail(SX_EINTERNAL, "sx_index_get_view: the build's sort gave up waiting in a look-back; "
                                  "the index is empty and must be rebuilt");
    if (ix->from_codes && flags[0])
        return fail(SX_EDATA, "sx_index_get_view: " + std::to_string(flags[0]) +
                                  " codes given to sx_index_build_codes are not strictly increasing or do not fit "
                                  "in the index's code bits; the index is empty");
    int64_t nu = 0, no = 0;
    SX_TRY(sx_check_counts(ix->counts + kCntUnique, ix->n_input, ix->counts + kCntOct, ix->oct_cap,
                           ix->counts + kCntBad, &nu, &no, stream));
    SX_HIP(hipMemcpy(ix->host_bounds, ix->bounds, sizeof(float) * 3, hipMemcpyDeviceToHost));
    ix->n_unique = nu;
    ix->n_oct = no;
    ix->counts_valid = true;
    view->n_input = ix->n_input;
    view->n_unique = nu;
    view->n_oct = no;
    view->min_coord = ix->host_bounds[0];
    view->max_coord = ix->host_bounds[1];
    view->code_bits = ix->code_bits;
    view->codes = ix->ucodes;
    view->has_leaf_left = ix->hll;
    view->has_leaf_right = ix->hlr;
    view->prefix_n = ix->pn;
    view->left_child = ix->lc;
    view->parent = ix->parent;
    view->oct_nodes = ix->oct;
    view->points = ix->pts12;
    view->sorted_perm = ix->perm;
    view->rep = ix->rep;
    view->unique_of_input = ix->uoi;
    view->d_counts = ix->counts;
    view->d_bounds = ix->bounds;
    return SX_OK;
}

int sx_estimate_normals_noalloc(sx_index* ix, int64_t q_start, int64_t n_q, int k, int knn_mode,
                                const float* d_radius2, int precision, float* d_normals, void* stream) {
    SX_TRY(check_index(ix));
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_estimate_normals_noalloc: k must be 8, 16, 24 or 32");
    if (n_q < 0 || n_q * k > kNoAllocEntries)
        return fail(SX_EINVAL, "sx_estimate_normals_noalloc: n_q * k must be in [0, 2^24] (" +
                                   std::to_string(n_q) + " x " + std::to_string(k) +
                                   "); split the range into calls of at most 2^24 / k queries");
    SX_TRY(check_queries(ix, q_start, n_q));
    if (n_q == 0) return SX_OK;
    if (!d_normals) return fail(SX_EINVAL, "sx_estimate_normals_noalloc: null normals");
    int32_t* nn = nullptr;
    SX_TRY(noalloc_rows(&nn));
    // every row [0, n_q) is written by the kNN (rows past the unique count
    // as -1), so nothing of the previous call's contents needs clearing
    SX_TRY(sx_knn(ix, q_start, n_q, k, knn_mode, d_radius2, nn, stream));
    return sx_normals(ix, nn, q_start, n_q, k, precision, 0.0f, 0.0f, 0.0f, d_normals, nullptr, stream);
}

int sx_index_to_input_order(sx_index* ix, const void* d_rows, int64_t row_bytes, void* d_out, void* stream) {
    SX_TRY(check_index(ix));
    if (row_bytes <= 0 || row_bytes % 4 != 0)
        return fail(SX_EINVAL, "sx_index_to_input_order: row_bytes must be a positive multiple of 4");
    if (!d_rows || !d_out || (reinterpret_cast<uintptr_t>(d_rows) | reinterpret_cast<uintptr_t>(d_out)) % 4 != 0)
        return fail(SX_EINVAL, "sx_index_to_input_order: rows must be 4-byte aligned device buffers");
    hipStream_t st = sx::as_stream(stream);
    SX_TRY(after_build(ix, stream));
    SX_TRY(input_order_map(ix, st));
    const int64_t words = row_bytes / 4;
    sx::gather_rows_kernel<<<sx::grid_for(ix->n_input * words, 256), 256, 0, st>>>(
        static_cast<const uint32_t*>(d_rows), words, ix->uoi, ix->n_input, static_cast<uint32_t*>(d_out));
    SX_LAUNCH_CHECK("gather_rows_kernel");
    return SX_OK;
}

int sx_index_knn_reruns(sx_index* ix, int64_t* reruns, int reset, void* stream) {
    if (!ix || !reruns) return fail(SX_EINVAL, "sx_index_knn_reruns: null argument");
    int32_t h = 0;
    hipStream_t st = sx::as_stream(stream);
    SX_HIP(hipMemcpyAsync(&h, ix->counts + kCntReruns, 4, hipMemcpyDeviceToHost, st));
    SX_HIP(hipStreamSynchronize(st));
    *reruns = h;
    if (reset) SX_TRY(sx::fill_i32(ix->counts + kCntReruns, 0, 1, st));
    return SX_OK;
}

int sx_index_set_knn_stats(sx_index* ix, int enable) {
    if (!ix) return fail(SX_EINVAL, "sx_index_set_knn_stats: null index");
    ix->knn_stats = enable != 0;
    return SX_OK;
}

int sx_index_knn_candidates(sx_index* ix, int64_t* candidates, int reset, void* stream) {
    if (!ix || !candidates) return fail(SX_EINVAL, "sx_index_knn_candidates: null argument");
    int32_t h = 0;
    hipStream_t st = sx::as_stream(stream);
    SX_HIP(hipMemcpyAsync(&h, ix->counts + kCntCand, 4, hipMemcpyDeviceToHost, st));
    SX_HIP(hipStreamSynchronize(st));
    *candidates = static_cast<uint32_t>(h);
    if (reset) SX_TRY(sx::fill_i32(ix->counts + kCntCand, 0, 1, st));
    return SX_OK;
}

int sx_knn(sx_index* ix, int64_t q_start, int64_t n_q, int k, int mode, const float* d_radius2, int32_t* d_out,
           void* stream) {
    SX_TRY(check_index(ix));
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_knn: k must be 8, 16, 24 or 32");
    if (mode != SX_KNN_PLAIN && mode != SX_KNN_RADIUS) return fail(SX_EINVAL, "sx_knn: bad mode");
    if (mode == SX_KNN_RADIUS && !d_radius2) return fail(SX_EINVAL, "sx_knn: radius mode needs d_radius2");
    SX_TRY(check_queries(ix, q_start, n_q));
    if (n_q > 0 && !d_out) return fail(SX_EINVAL, "sx_knn: null output");
    SX_TRY(after_build(ix, stream));
    const float* r2 = mode == SX_KNN_RADIUS ? d_radius2 : nullptr;
    if (ix->knn_algo == SX_KNN_ALGO_TILES)
        return sx::knn_tiles_launch(ix->pts16, ix->trav, trav_parents(ix), ix->counts, q_start, n_q, k, r2, ix->bucket, d_out,
                                    ix->knn_queue, ix->counts + kCntQueue, ix->counts + kCntReruns,
                                    sx::as_stream(stream), sx::CloudMap(), 0,
                                    ix->knn_stats ? ix->counts + kCntCand : nullptr);
    return sx::knn_launch(ix->pts16, ix->trav, trav_parents(ix), ix->counts, q_start, n_q, nullptr, k, r2, ix->bucket, d_out,
                          ix->counts + kCntReruns, sx::as_stream(stream));
}

int sx_knn_points(sx_index* ix, const float* d_queries, int64_t n_q, int k, int32_t* d_out, void* stream) {
    SX_TRY(check_index(ix));
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_knn_points: k must be 8, 16, 24 or 32");
    if (n_q < 0 || (n_q > 0 && (!d_queries || !d_out))) return fail(SX_EINVAL, "sx_knn_points: bad arguments");
    SX_TRY(after_build(ix, stream));
    return sx::knn_launch(ix->pts16, ix->trav, trav_parents(ix), ix->counts, 0, n_q, d_queries, k, nullptr, ix->bucket, d_out, nullptr,
                          sx::as_stream(stream));
}

int sx_normals(sx_index* ix, const int32_t* d_nn, int64_t q_start, int64_t n_q, int k, int precision, float vx,
               float vy, float vz, float* d_normals, uint8_t* d_flags, void* stream) {
    SX_TRY(check_index(ix));
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_normals: k must be 8, 16, 24 or 32");
    SX_TRY(check_queries(ix, q_start, n_q));
    if (n_q > 0 && (!d_nn || !d_normals)) return fail(SX_EINVAL, "sx_normals: null pointer");
    SX_TRY(after_build(ix, stream));
    return sx::normals_launch(ix->pts16, ix->counts, d_nn, q_start, n_q, k, precision, vx, vy, vz, d_normals, d_flags,
                              sx::as_stream(stream));
}

int sx_estimate_normals(sx_index* ix, int64_t q_start, int64_t n_q, int k, int knn_mode, const float* d_radius2,
                        int precision, int32_t* d_nn, float* d_normals, void* stream) {
    SX_TRY(check_index(ix));
    if (!valid_k(k)) return fail(SX_EINVAL, "sx_estimate_normals: k must be 8, 16, 24 or 32");
    SX_TRY(check_queries(ix, q_start, n_q));
    if (n_q == 0) return SX_OK;
    if (!d_normals) return fail(SX_EINVAL, "sx_estimate_normals: null normals");
    int32_t* nn = d_nn;
    if (!nn) {
        const int64_t need = n_q * k;
        if (need > ix->nn_buf_len) {
            if (ix->nn_buf) SX_HIP(hipFree(ix->nn_buf));
            ix->nn_buf = nullptr;
            ix->nn_buf_len = 0;
            hipError_t e = hipMalloc(&ix->nn_buf, static_cast<size_t>(need) * sizeof(int32_t));
            if (e != hipSuccess) {
                (void)hipGetLastError();
                ix->nn_buf = nullptr;
                return fail(SX_ENOMEM, "sx_estimate_normals: neighbour buffer allocation failed");
            }
            ix->nn_buf_len = need;
        }
        nn = ix->nn_buf;
    }
    SX_TRY(sx_knn(ix, q_start, n_q, k, knn_mode, d_radius2, nn, stream));
    return sx_normals(ix, nn, q_start, n_q, k, precision, 0.0f, 0.0f, 0.0f, d_normals, nullptr, stream);
}

}  // extern "C"
