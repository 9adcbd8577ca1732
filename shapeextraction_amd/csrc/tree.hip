// Duplicate removal, Karras radix tree, bottom-up bounding boxes.
//
// unique:  head flags (code differs from its predecessor) + device scan +
//          compaction in the scan's output step; replaces findDups,
//          cub InclusiveSum and moveDups (src/RadixTree.cu:71-97,258-303).
//          The representative of a run is its first element in stable sort
//          order, i.e. the lowest input index carrying that code.
// karras:  one thread per internal node (Karras 2012), the same direction,
//          exponential + binary range search and ceil-step split search as
//          constructTree (src/RadixTree.cu:99-194); root rule selectable.
// boxes:   the Karras thread of a node also writes its traversal record: the
//          boxes of both children as range min/max queries over per-level
//          block boxes (no atomics or fences between threads).
#include <hip/hip_runtime.h>

#include "sx_internal.hpp"

namespace sx {

constexpr int kMaxBoxLevels = 16;  // 4^15 = 2^30 points

struct UniqueIn {
    const uint64_t* s;
    __device__ int operator()(int64_t i) const { return (i == 0 || s[i] != s[i - 1]) ? 1 : 0; }
};

struct UniqueOut {
    const uint64_t* s;
    const uint32_t* vals;
    uint64_t* uniq;
    uint32_t* rep;
    int32_t* uos;  // unique id of each sorted position
    int32_t* uoi;  // unique id of each input index (needs vals)
    __device__ void operator()(int64_t i, int ex, int v) const {
        if (v) {
            uniq[ex] = s[i];
            if (rep) rep[ex] = vals ? vals[i] : static_cast<uint32_t>(i);
        }
        const int u = ex + v - 1;
        if (uos) uos[i] = u;
        if (uoi && vals) uoi[vals[i]] = u;
    }
};

int unique_launch(const uint64_t* sorted, const uint32_t* vals, int64_t n, uint64_t* uniq, uint32_t* rep,
                  int32_t* uos, int32_t* uoi, int32_t* d_count, void* ws, hipStream_t st) {
    return device_scan(UniqueIn{sorted}, UniqueOut{sorted, vals, uniq, rep, uos, uoi}, nullptr, n, d_count, ws, st);
}

// ---------------------------------------------------------------------------
// Child boxes without inter-thread communication.  Every radix node covers a
// contiguous range of Morton-sorted points, so its box is a range min/max.
// Level L >= 1 holds the box of every aligned block of 4^L points (one
// kernel per level); a range query takes up to 3 elements at each end of the
// range per level - issued together, one memory round trip per level - and
// moves the aligned middle up a level (3.6 elements per node on average on
// an 8M-point scene, about two levels).  min/max are exact, so the boxes equal a
// bottom-up merge, with no atomics or fences between threads - which on a
// GPU of several L2 domains would cost a cache write-back per merge.
// ---------------------------------------------------------------------------
constexpr int kBoxFanLog = 2;
constexpr int kBoxFan = 1 << kBoxFanLog;

struct BoxLevels {
    float4* lo[kMaxBoxLevels];
    float4* hi[kMaxBoxLevels];
    int n_levels;  // levels 1..n_levels-1 are stored; level 0 is the points
};

__device__ __forceinline__ int64_t level_count(int64_t U, int L) {
    int64_t n = U;
    for (int i = 0; i < L; ++i) n = (n + kBoxFan - 1) >> kBoxFanLog;
    return n;
}

__global__ void __launch_bounds__(256) box_level_kernel(const float4* __restrict__ pts, BoxLevels lv, int L,
                                                        const int32_t* d_count, int64_t n_max) {
    const int64_t U = scan_count(d_count, n_max);
    const int64_t n_dst = level_count(U, L), n_src = level_count(U, L - 1);
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n_dst) return;
    float lx = INFINITY, ly = INFINITY, lz = INFINITY, hx = -INFINITY, hy = -INFINITY, hz = -INFINITY;
    const int64_t s0 = j << kBoxFanLog;
    const int64_t s1 = s0 + kBoxFan < n_src ? s0 + kBoxFan : n_src;
    for (int64_t s = s0; s < s1; ++s) {
        float4 l, h;
        if (L == 1) {
            l = h = pts[s];
        } else {
            l = lv.lo[L - 1][s];
            h = lv.hi[L - 1][s];
        }
        lx = fminf(lx, l.x); ly = fminf(ly, l.y); lz = fminf(lz, l.z);
        hx = fmaxf(hx, h.x); hy = fmaxf(hy, h.y); hz = fmaxf(hz, h.z);
    }
    lv.lo[L][j] = make_float4(lx, ly, lz, 0.0f);
    lv.hi[L][j] = make_float4(hx, hy, hz, 0.0f);
}

struct Box3 {
    float lx, ly, lz, hx, hy, hz;
    __device__ __forceinline__ void init() {
        lx = ly = lz = INFINITY;
        hx = hy = hz = -INFINITY;
    }
    __device__ __forceinline__ void set(float4 p) {
        lx = hx = p.x;
        ly = hy = p.y;
        lz = hz = p.z;
    }
    __device__ __forceinline__ void add(float4 l, float4 h) {
        lx = fminf(lx, l.x); ly = fminf(ly, l.y); lz = fminf(lz, l.z);
        hx = fmaxf(hx, h.x); hy = fmaxf(hy, h.y); hz = fmaxf(hz, h.z);
    }
};

// Elements [first, first + cnt) of level L, 0 <= cnt < kBoxFan: the loads
// are unconditional (indices clamped to the run, repeats leave a min/max
// unchanged) so that they are in flight together.
__device__ __forceinline__ void add_run(const float4* __restrict__ pts, const BoxLevels& lv, int L, int64_t first,
                                        int cnt, Box3& bx) {
    if (cnt <= 0) return;
    float4 l[kBoxFan - 1], h[kBoxFan - 1];
    if (L == 0) {
#pragma unroll
        for (int t = 0; t < kBoxFan - 1; ++t) l[t] = h[t] = pts[first + (t < cnt ? t : cnt - 1)];
    } else {
        const float4* lo = lv.lo[L];
        const float4* hi = lv.hi[L];
#pragma unroll
        for (int t = 0; t < kBoxFan - 1; ++t) {
            const int64_t j = first + (t < cnt ? t : cnt - 1);
            l[t] = lo[j];
            h[t] = hi[j];
        }
    }
#pragma unroll
    for (int t = 0; t < kBoxFan - 1; ++t) bx.add(l[t], h[t]);
}

// Box of the points [a, b] (inclusive).
__device__ Box3 range_box(const float4* __restrict__ pts, const BoxLevels& lv, int64_t a, int64_t b) {
    constexpr int64_t m = kBoxFan - 1;
    Box3 bx;
    bx.init();
    int L = 0;
    while (a <= b) {
        if (L == lv.n_levels - 1) {  // a single block remains at the top level
            for (int64_t j = a; j <= b; ++j) add_run(pts, lv, L, j, 1, bx);
            break;
        }
        // unaligned head [a, a + na) and tail (b - nb, b]
        const int64_t len = b - a + 1;
        const int64_t na = ((kBoxFan - (a & m)) & m) < len ? ((kBoxFan - (a & m)) & m) : len;
        const int64_t nb = len - na < ((b + 1) & m) ? len - na : ((b + 1) & m);
        add_run(pts, lv, L, a, static_cast<int>(na), bx);
        add_run(pts, lv, L, b - nb + 1, static_cast<int>(nb), bx);
        a += na;
        b -= nb;
        if (a > b) break;
        a >>= kBoxFanLog;
        b = ((b + 1) >> kBoxFanLog) - 1;
        ++L;
    }
    return bx;
}

int box_levels_for(int64_t n_max) {
    int n = 1;
    for (int64_t c = n_max; c > 1 && n < kMaxBoxLevels; c = (c + kBoxFan - 1) / kBoxFan) ++n;
    return n;
}

size_t box_ws_bytes(int64_t n_max) {
    Carver c(nullptr);
    int64_t cnt = n_max;
    for (int L = 1; L < box_levels_for(n_max); ++L) {
        cnt = (cnt + kBoxFan - 1) / kBoxFan;
        c.take<float4>(static_cast<size_t>(cnt));
        c.take<float4>(static_cast<size_t>(cnt));
    }
    return c.off;
}

int box_levels_launch(const float4* pts, const int32_t* d_count, int64_t n_max, void* ws, BoxLevels* lv,
                      hipStream_t st) {
    *lv = BoxLevels{};
    lv->n_levels = box_levels_for(n_max);
    Carver c(ws);
    int64_t cnt = n_max;
    for (int L = 1; L < lv->n_levels; ++L) {
        cnt = (cnt + kBoxFan - 1) / kBoxFan;
        lv->lo[L] = c.take<float4>(static_cast<size_t>(cnt));
        lv->hi[L] = c.take<float4>(static_cast<size_t>(cnt));
        box_level_kernel<<<grid_for(cnt, 256), 256, 0, st>>>(pts, *lv, L, d_count, n_max);
        SX_LAUNCH_CHECK("box_level_kernel");
    }
    return SX_OK;
}

__device__ __forceinline__ int ceil_log2_i64(int64_t l) {
    return l <= 1 ? 0 : 64 - __clzll(static_cast<long long>(l - 1));
}

// Codes within kKarrasHalo of a block's nodes are staged in LDS: most nodes
// are short (median range 3 points on a scene), so their direction, range
// and split searches never leave the block's window, where a global probe
// would be one dependent memory round trip per search step.
constexpr int kKarrasBlock = 256;
constexpr int kKarrasHalo = 256;

__global__ void __launch_bounds__(kKarrasBlock)
karras_kernel(const uint64_t* __restrict__ codes, const int32_t* d_count, int64_t n_max, int code_bits, int root_mode,
              uint8_t* prefix_n, int32_t* left_child, uint8_t* has_leaf_left, uint8_t* has_leaf_right,
              int32_t* parent, int32_t* node_end, int32_t* leaf_parent, int2* up, TravNode* trav,
              const float4* pts, BoxLevels lv, CloudMap cm) {
    __shared__ uint64_t s_codes[kKarrasBlock + 2 * kKarrasHalo];
    // g: slot; the cloud holding it spans slots [base, base + U) and its
    // internal node i (local) is stored at slot base + i
    const int64_t total = scan_count(d_count, n_max);
    const int64_t blk0 = static_cast<int64_t>(blockIdx.x) * kKarrasBlock;
    if (blk0 >= total) return;  // whole block idle (before the barrier)
    const int64_t w0 = blk0 - kKarrasHalo > 0 ? blk0 - kKarrasHalo : 0;
    const int64_t w1 = blk0 + kKarrasBlock + kKarrasHalo < total ? blk0 + kKarrasBlock + kKarrasHalo : total;
    for (int64_t x = w0 + threadIdx.x; x < w1; x += kKarrasBlock) s_codes[x - w0] = codes[x];
    __syncthreads();
    const int64_t g = blk0 + threadIdx.x;
    if (g >= total) return;
    const int2 cloud = cm.at(g, static_cast<int>(total));
    const int64_t base = cloud.x, U = cloud.y;
    const int64_t N = U - 1;  // internal nodes; leaves are 0..N
    const int64_t i = g - base;
    if (i == 0 && parent) parent[g] = -1;
    if (i == 0 && up) up[g].x = -1;
    if (i >= N) return;
    // code of the cloud's leaf x (local index)
    auto code = [&](int64_t x) -> uint64_t {
        const int64_t gx = base + x;
        return (gx >= w0 && gx < w1) ? s_codes[gx - w0] : codes[gx];
    };
    const uint64_t ci = code(i);
    int d;
    if (i == 0) {
        d = 1;
    } else {
        const int diff = prefix_len(ci, code(i + 1), code_bits) - prefix_len(ci, code(i - 1), code_bits);
        d = (diff > 0) - (diff < 0);
    }
    int64_t l = 0;
    if (i == 0) {
        // Reference root: l = N-1 (src/RadixTree.cu:128-131); with one internal
        // node that length is 0 and the reference's split search is undefined,
        // so both rules use the full range there.
        l = (root_mode == SX_ROOT_REFERENCE && N >= 2) ? N - 1 : N;
    } else {
        const int dmin = prefix_len(ci, code(i - d), code_bits);
        int64_t lmax = 2;
        while (true) {
            const int64_t jj = i + lmax * d;
            if (jj < 0 || jj > N || prefix_len(ci, code(jj), code_bits) <= dmin) break;
            lmax *= 2;
        }
        const int64_t cutoff = (d == -1) ? i : N - i;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (l + t <= cutoff && prefix_len(ci, code(i + (l + t) * d), code_bits) > dmin) l += t;
    }
    const int64_t j = i + l * d;
    const int dnode = prefix_len(ci, code(j), code_bits);
    int64_t s = 0;
    const int64_t scut = (d == -1) ? i - 1 : N - i - 1;
    const int lg = ceil_log2_i64(l);
    for (int e = 1; e <= lg; ++e) {
        const int64_t div = 1ll << e;
        const int64_t t = (l + div - 1) / div;
        if (s + t <= scut && prefix_len(ci, code(i + (s + t) * d), code_bits) > dnode) s += t;
    }
    const int64_t gamma = i + s * d + (d < 0 ? d : 0);
    const int64_t a = i < j ? i : j, b = i < j ? j : i;
    const bool hl = (a == gamma), hr = (b == gamma + 1);
    // global (slot) numbers from here on
    const int64_t G = base + gamma;
    if (prefix_n) prefix_n[g] = static_cast<uint8_t>(dnode);
    if (left_child) left_child[g] = static_cast<int32_t>(G);
    if (has_leaf_left) has_leaf_left[g] = hl ? 1 : 0;
    if (has_leaf_right) has_leaf_right[g] = hr ? 1 : 0;
    if (parent) {
        if (!hl) parent[G] = static_cast<int32_t>(g);
        if (!hr) parent[G + 1] = static_cast<int32_t>(g);
    }
    if (node_end) node_end[g] = static_cast<int32_t>(base + j);
    if (leaf_parent) {
        if (hl) leaf_parent[G] = static_cast<int32_t>(2 * g);
        if (hr) leaf_parent[G + 1] = static_cast<int32_t>(2 * g + 1);
    }
    if (up) {  // .x (parent) of a node is written by its parent's thread, .y here
        up[g].y = dnode;
        if (!hl) up[G].x = static_cast<int32_t>(g);
        if (!hr) up[G + 1].x = static_cast<int32_t>(g);
    }
    if (trav) {
        // the record of node g: child boxes with (first leaf, split, last
        // leaf) in .w, written whole; .w of hi_r (the parent) is written by
        // the parent's thread, so hi_r's box goes out component-wise
        const int64_t A = base + a, B = base + b;
        Box3 bl, br;
        if (hl) {
            const float4 q = pts[G];
            bl.set(q);
        } else {
            bl = range_box(pts, lv, A, G);
        }
        if (hr) {
            const float4 q = pts[G + 1];
            br.set(q);
        } else {
            br = range_box(pts, lv, G + 1, B);
        }
        TravNode* t = trav + g;
        t->lo_l = make_float4(bl.lx, bl.ly, bl.lz, __int_as_float(static_cast<int>(A)));
        t->hi_l = make_float4(bl.hx, bl.hy, bl.hz, __int_as_float(static_cast<int>(G)));
        t->lo_r = make_float4(br.lx, br.ly, br.lz, __int_as_float(static_cast<int>(B)));
        if (i == 0) {
            t->hi_r = make_float4(br.hx, br.hy, br.hz, __int_as_float(-1));
        } else {
            t->hi_r.x = br.hx;
            t->hi_r.y = br.hy;
            t->hi_r.z = br.hz;
        }
        if (!hl) trav[G].hi_r.w = __int_as_float(static_cast<int>(g));
        if (!hr) trav[G + 1].hi_r.w = __int_as_float(static_cast<int>(g));
    }
}

// With trav, pts (the decoded points, slot numbering) and box_ws
// (box_ws_bytes(n_max)) are required: the block-box levels are built first.
int karras_launch(const uint64_t* codes, const int32_t* d_count, int64_t n_max, int code_bits, int root_mode,
                  uint8_t* prefix_n, int32_t* left_child, uint8_t* hll, uint8_t* hlr, int32_t* parent,
                  int32_t* node_end, int32_t* leaf_parent, int2* up, TravNode* trav, const float4* pts,
                  void* box_ws, hipStream_t st, CloudMap cm = CloudMap()) {
    if (n_max <= 0) return SX_OK;
    BoxLevels lv{};
    if (trav) {
        if (!pts || !box_ws) return fail(SX_EINVAL, "karras_launch: traversal records need points and workspace");
        SX_TRY(box_levels_launch(pts, d_count, n_max, box_ws, &lv, st));
    }
    // Nodes that no other node claims (only possible under the reference root
    // rule) keep parent -1.
    if (parent) SX_TRY(fill_i32(parent, -1, n_max, st));
    karras_kernel<<<grid_for(n_max, kKarrasBlock), kKarrasBlock, 0, st>>>(
        codes, d_count, n_max, code_bits, root_mode, prefix_n, left_child, hll, hlr, parent, node_end, leaf_parent, up,
        trav, pts, lv, cm);
    SX_LAUNCH_CHECK("karras_kernel");
    return SX_OK;
}

}  // namespace sx

using sx::fail;

extern "C" {

int sx_unique(const uint64_t* d_sorted, const uint32_t* d_vals, int64_t n, uint64_t* d_unique, uint32_t* d_rep,
              int32_t* d_unique_of_sorted, int32_t* d_count, void* ws, size_t* ws_bytes, void* stream) {
    if (!ws_bytes) return fail(SX_EINVAL, "sx_unique: ws_bytes is null");
    if (n <= 0) return fail(SX_EINVAL, "sx_unique: n must be positive");
    const size_t need = sx::scan_ws_bytes(n);
    if (!ws) {
        *ws_bytes = need;
        return SX_OK;
    }
    if (*ws_bytes < need) return fail(SX_ENOMEM, "sx_unique: workspace too small");
    if (!d_sorted || !d_unique || !d_count) return fail(SX_EINVAL, "sx_unique: null pointer");
    return sx::unique_launch(d_sorted, d_vals, n, d_unique, d_rep, d_unique_of_sorted, nullptr, d_count, ws,
                             sx::as_stream(stream));
}

int sx_build_radix_tree(const uint64_t* d_codes, const int32_t* d_count, int64_t n_max, int code_bits, int root_mode,
                        uint8_t* d_prefix_n, int32_t* d_left_child, uint8_t* d_has_leaf_left,
                        uint8_t* d_has_leaf_right, int32_t* d_parent, int32_t* d_node_end, int32_t* d_leaf_parent,
                        void* stream) {
    if (code_bits != 30 && code_bits != 63) return fail(SX_EINVAL, "sx_build_radix_tree: code_bits must be 30 or 63");
    if (root_mode != SX_ROOT_TEXTBOOK && root_mode != SX_ROOT_REFERENCE)
        return fail(SX_EINVAL, "sx_build_radix_tree: bad root_mode");
    if (n_max <= 0 || !d_codes) return fail(SX_EINVAL, "sx_build_radix_tree: bad arguments");
    return sx::karras_launch(d_codes, d_count, n_max, code_bits, root_mode, d_prefix_n, d_left_child, d_has_leaf_left,
                             d_has_leaf_right, d_parent, d_node_end, d_leaf_parent, nullptr, nullptr, nullptr,
                             nullptr, sx::as_stream(stream));
}

}  // extern "C"
