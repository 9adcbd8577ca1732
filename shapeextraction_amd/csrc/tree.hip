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
// boxes:   the Karras thread of a node also writes its traversal record:
//          tight child boxes of subtrees inside the thread block bottom-up in
//          LDS, Morton-cell boxes for the others (no global atomics or
//          fences between threads).
#include <hip/hip_runtime.h>

#include "sx_internal.hpp"

namespace sx {

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

// sort_err (nullable): the sort's error flag; when it is set the unique count
// is written as 0, so that every later stage (sized by that count) does no
// work on keys that are not sorted.
int unique_launch(const uint64_t* sorted, const uint32_t* vals, int64_t n, uint64_t* uniq, uint32_t* rep,
                  int32_t* uos, int32_t* uoi, int32_t* d_count, void* ws, hipStream_t st,
                  const int32_t* sort_err = nullptr) {
    return device_scan(UniqueIn{sorted}, UniqueOut{sorted, vals, uniq, rep, uos, uoi}, nullptr, n, d_count, ws, st,
                       sort_err);
}

struct Box3 {
    float lx, ly, lz, hx, hy, hz;
    __device__ __forceinline__ void set(float4 p) {
        lx = hx = p.x;
        ly = hy = p.y;
        lz = hz = p.z;
    }
    __device__ __forceinline__ void set(float4 l, float4 h) {
        lx = l.x; ly = l.y; lz = l.z;
        hx = h.x; hy = h.y; hz = h.z;
    }
};

// Box of the Morton cell of `code`'s first p bits: every code with that
// prefix, i.e. every point of a radix node whose prefix length is p.  The
// corners are decoded like the points (dequantize is monotone in the raw
// coordinate), so the box holds every decoded point of the cell.
// The low code_bits - p code bits are free; axis a owns ceil((free - a) / 3) of
// them, the low bits of its raw coordinate, so the corners are the code's raw
// coordinates with those bits cleared and set.
__device__ __forceinline__ Box3 cell_box(uint64_t code, int p, int code_bits, float mn, float range, float fscale) {
    const int free_bits = code_bits - p;
    auto axis_mask = [&](int a) -> uint32_t {
        const int n = free_bits > a ? (free_bits - a + 2) / 3 : 0;
        return (1u << n) - 1u;
    };
    const uint32_t mx = axis_mask(0), my = axis_mask(1), mz = axis_mask(2);
    const uint32_t rx = compact3(code), ry = compact3(code >> 1), rz = compact3(code >> 2);
    Box3 bx;
    bx.lx = dequantize(rx & ~mx, fscale, range, mn);
    bx.ly = dequantize(ry & ~my, fscale, range, mn);
    bx.lz = dequantize(rz & ~mz, fscale, range, mn);
    bx.hx = dequantize(rx | mx, fscale, range, mn);
    bx.hy = dequantize(ry | my, fscale, range, mn);
    bx.hz = dequantize(rz | mz, fscale, range, mn);
    return bx;
}

// For codes of code_bits significant bits: prefix_len(a, b) > d exactly when
// (a ^ b) < longer_than(d): the first differing bit lies below the top d + 1.
__device__ __forceinline__ uint64_t longer_than(int d, int code_bits) {
    return d >= code_bits ? 0ull : (1ull << (code_bits - 1 - d));
}

__device__ __forceinline__ int ceil_log2_i64(int64_t l) {
    return l <= 1 ? 0 : 64 - __clzll(static_cast<long long>(l - 1));
}

// Codes within kKarrasHalo of a block's nodes are staged in LDS: most nodes
// are short (median range 3 points on a scene), so their direction, range
// and split searches never leave the block's window, where a global probe
// would be one dependent memory round trip per search step.
//
// Child boxes: a node whose whole range lies in the block's own keys (most
// nodes) gets its tight box bottom-up in LDS - a level-synchronous pass over
// the block's nodes, each taking the union of its two children once both are
// known, the block's points staged beside them (min/max are exact, so this
// is the bottom-up merge).  A child whose range leaves the block (about 2%
// of the children on an 8M-point scene, the large ones near the top of the
// tree) gets the box of its Morton cell instead: looser, still a bound of
// every point in it, and computed from two codes without touching points.
constexpr int kKarrasBlock = 256;
constexpr int kKarrasHalo = 256;

__global__ void __launch_bounds__(kKarrasBlock)
karras_kernel(const uint64_t* __restrict__ codes, const int32_t* d_count, int64_t n_max, int code_bits, int root_mode,
              uint8_t* prefix_n, int32_t* left_child, uint8_t* has_leaf_left, uint8_t* has_leaf_right,
              int32_t* parent, int32_t* node_end, int32_t* leaf_parent, int2* up, TravNode* trav,
              const float4* pts, const float* __restrict__ bounds, CloudMap cm, int32_t* oct_count) {
    __shared__ uint64_t s_codes[kKarrasBlock + 2 * kKarrasHalo];
    __shared__ float4 s_pts[kKarrasBlock];  // points of the block's own slots
    __shared__ float4 s_lo[kKarrasBlock], s_hi[kKarrasBlock];  // box of each in-block node
    __shared__ int s_done[kKarrasBlock];
    __shared__ int s_pending[3];  // per bottom-up step (rotating): some in-block box still missing
    // g: slot; the cloud holding it spans slots [base, base + U) and its
    // internal node i (local) is stored at slot base + i
    const int64_t total = scan_count(d_count, n_max);
    // XCD-contiguous block order: a block's halo is its neighbours' keys, read
    // from the same L2
    const int64_t blk0 = static_cast<int64_t>(xcd_block()) * kKarrasBlock;
    if (blk0 >= total) return;  // whole block idle (before every barrier)
    const int tid = static_cast<int>(threadIdx.x);
    const int64_t blk1 = blk0 + kKarrasBlock < total ? blk0 + kKarrasBlock : total;  // the block's slots [blk0, blk1)
    const int64_t w0 = blk0 - kKarrasHalo > 0 ? blk0 - kKarrasHalo : 0;
    const int64_t w1 = blk0 + kKarrasBlock + kKarrasHalo < total ? blk0 + kKarrasBlock + kKarrasHalo : total;
    for (int64_t x = w0 + tid; x < w1; x += kKarrasBlock) s_codes[x - w0] = codes[x];
    if (trav) {
        if (blk0 + tid < blk1) s_pts[tid] = pts[blk0 + tid];
        s_done[tid] = 0;
        if (tid < 3) s_pending[tid] = 0;
    }
    __syncthreads();
    const int64_t g = blk0 + tid;
    const bool in_range = g < total;
    const int2 cloud = in_range ? cm.at(g, static_cast<int>(total)) : make_int2(0, 0);
    const int64_t base = cloud.x, U = cloud.y;
    const int64_t N = U - 1;  // internal nodes; leaves are 0..N
    const int64_t i = g - base;
    if (in_range && (i == 0 || i == N) && parent) parent[g] = -1;
    if (in_range && i == 0 && up) up[g].x = -1;
    const bool active = in_range && i < N;
    // slots that hold no internal node (a cloud's last leaf, a batch's
    // padding) own no octree chain
    if (oct_count && in_range && !active) oct_count[g] = 0;
    // code of the cloud's leaf x (local index): from LDS inside [w0, w1), else
    // from HBM.  The LDS read is unconditional (index clamped) and the empty
    // asm keeps it apart from the HBM read: merged they would become one
    // generic load on every search step.
    const uint64_t win = static_cast<uint64_t>(w1 - w0);
    auto code = [&](int64_t x) -> uint64_t {
        const int64_t gx = base + x;
        const uint64_t o = static_cast<uint64_t>(gx - w0);
        const bool in_win = o < win;
        uint64_t v = s_codes[in_win ? static_cast<int>(o) : 0];
        asm volatile("" : "+v"(v));
        if (!in_win) v = codes[gx];
        return v;
    };
    // i + t * d for the direction d = +-1, without a 64-bit multiply
    auto step = [&](int64_t from, int64_t t, int d) -> int64_t { return d > 0 ? from + t : from - t; };
    int64_t A = 0, B = 0, G = 0;  // the node's range [A, B] and split G (slots)
    bool hl = false, hr = false;
    if (active) {
        const uint64_t ci = code(i);
        int d;
        int d_out = -1;  // prefix with the neighbour outside the range on i's side
        if (i == 0) {
            d = 1;
        } else {
            const int p_next = prefix_len(ci, code(i + 1), code_bits), p_prev = prefix_len(ci, code(i - 1), code_bits);
            const int diff = p_next - p_prev;
            d = (diff > 0) - (diff < 0);
            d_out = d > 0 ? p_prev : p_next;
        }
        int64_t l = 0;
        if (i == 0) {
            // Reference root: l = N-1 (src/RadixTree.cu:128-131); with one internal
            // node that length is 0 and the reference's split search is undefined,
            // so both rules use the full range there.
            l = (root_mode == SX_ROOT_REFERENCE && N >= 2) ? N - 1 : N;
        } else {
            const uint64_t tmin = longer_than(prefix_len(ci, code(i - d), code_bits), code_bits);
            int64_t lmax = 2;
            while (true) {
                const int64_t jj = step(i, lmax, d);
                if (jj < 0 || jj > N || (ci ^ code(jj)) >= tmin) break;
                lmax *= 2;
            }
            const int64_t cutoff = (d == -1) ? i : N - i;
            for (int64_t t = lmax / 2; t >= 1; t /= 2)
                if (l + t <= cutoff && (ci ^ code(step(i, l + t, d))) < tmin) l += t;
        }
        const int64_t j = step(i, l, d);
        const uint64_t cj = code(j);
        const int dnode = prefix_len(ci, cj, code_bits);
        int64_t s = 0;
        const int64_t scut = (d == -1) ? i - 1 : N - i - 1;
        const int lg = ceil_log2_i64(l);
        const uint64_t tnode = longer_than(dnode, code_bits);
        for (int e = 1; e <= lg; ++e) {
            const int64_t t = (l + (1ll << e) - 1) >> e;  // ceil(l / 2^e)
            if (s + t <= scut && (ci ^ code(step(i, s + t, d))) < tnode) s += t;
        }
        const int64_t gamma = d > 0 ? i + s : i - s - 1;
        const int64_t a = i < j ? i : j, b = i < j ? j : i;
        hl = (a == gamma);
        hr = (b == gamma + 1);
        // global (slot) numbers from here on
        A = base + a;
        B = base + b;
        G = base + gamma;
        if (prefix_n) prefix_n[g] = static_cast<uint8_t>(dnode);
        if (left_child) left_child[g] = static_cast<int32_t>(G);
        if (has_leaf_left) has_leaf_left[g] = hl ? 1 : 0;
        if (has_leaf_right) has_leaf_right[g] = hr ? 1 : 0;
        if (parent) {
            if (!hl) parent[G] = static_cast<int32_t>(g);
            if (!hr) parent[G + 1] = static_cast<int32_t>(g);
        }
        if (node_end) node_end[g] = static_cast<int32_t>(base + j);
        if (oct_count) {
            // octree edge count depth(node) - depth(parent), depth = prefix / 3
            // (calcEdgeCounts, src/Octree.cu:38-50): the parent's prefix is the
            // longer of the two prefixes across the node's range ends, since
            // the parent adds the sibling on that side; the root counts 1
            int cnt = 1;
            if (i > 0) {
                const int64_t jo = j + d;  // the neighbour outside the range on j's side
                const int d_far = (jo >= 0 && jo <= N) ? prefix_len(cj, code(jo), code_bits) : -1;
                cnt = dnode / 3 - (d_out > d_far ? d_out : d_far) / 3;
            }
            oct_count[g] = cnt;
        }
        if (leaf_parent) {
            if (hl) leaf_parent[G] = static_cast<int32_t>(2 * g);
            if (hr) leaf_parent[G + 1] = static_cast<int32_t>(2 * g + 1);
        }
        if (up) {  // .x (parent) of a node is written by its parent's thread, .y here
            up[g].y = dnode;
            if (!hl) up[G].x = static_cast<int32_t>(g);
            if (!hr) up[G + 1].x = static_cast<int32_t>(g);
        }
    }
    if (!trav) return;  // (uniform: no barriers follow without boxes)

    // ---- boxes of the in-block nodes, bottom-up: a node whose range lies in
    // [blk0, blk1) has both children there (leaves, or in-block nodes).
    const bool local = active && A >= blk0 && B < blk1;
    // One barrier per step.  A box is written before its flag (release) and
    // read after it (acquire), so a box finished in this step may already be
    // used in it.  Pending flags rotate over three slots: slot it % 3 is set
    // before step it's barrier and read after it by every thread alike; it is
    // cleared during step it + 2, when every read of step it has passed the
    // barrier of step it + 1, and set again only after that step's barrier.
    // Every step finishes at least the lowest missing box, so the loop ends
    // within kKarrasBlock steps.
    bool done = !local;
    int slot = 0;
#pragma unroll 1
    for (int it = 0; it < kKarrasBlock; ++it) {
        if (!done) {
            const int cl = static_cast<int>(G - blk0), cr = cl + 1;
            const bool rl = hl || __hip_atomic_load(&s_done[cl], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
            const bool rr = hr || __hip_atomic_load(&s_done[cr], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
            if (rl && rr) {
                const float4 l1 = hl ? s_pts[cl] : s_lo[cl], h1 = hl ? s_pts[cl] : s_hi[cl];
                const float4 l2 = hr ? s_pts[cr] : s_lo[cr], h2 = hr ? s_pts[cr] : s_hi[cr];
                s_lo[tid] = make_float4(fminf(l1.x, l2.x), fminf(l1.y, l2.y), fminf(l1.z, l2.z), 0.0f);
                s_hi[tid] = make_float4(fmaxf(h1.x, h2.x), fmaxf(h1.y, h2.y), fmaxf(h1.z, h2.z), 0.0f);
                __hip_atomic_store(&s_done[tid], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                done = true;
            }
        }
        const int next = slot == 2 ? 0 : slot + 1;
        if (!done) s_pending[slot] = 1;
        if (tid == 0) s_pending[next] = 0;
        __syncthreads();
        if (s_pending[slot] == 0) break;
        slot = next;
    }
    if (!active) return;
    // the record of node g: child boxes with (first leaf, split, last leaf)
    // in .w, written whole by this thread (parents are in their own array)
    const bool far_l = !hl && !(A >= blk0 && G < blk1), far_r = !hr && !(G + 1 >= blk0 && B < blk1);
    float mn = 0.0f, range = 0.0f;
    if (far_l || far_r) {
        const float* bd = cm.tile_cloud ? bounds + 4 * cm.tile_cloud[g / kCloudAlign] : bounds;
        mn = bd[0];
        range = bd[2];
    }
    const float fscale = static_cast<float>(axis_scale(code_bits));
    Box3 bl, br;
    // a leaf child's point: from LDS in the block, else from HBM (the empty asm
    // keeps the two loads apart; merged they would become one generic load)
    auto leaf_point = [&](int64_t x) -> float4 {
        const int o = static_cast<int>(x - blk0);
        float4 p = s_pts[o < 0 ? 0 : (o >= kKarrasBlock ? kKarrasBlock - 1 : o)];
        asm volatile("" : "+v"(p.x), "+v"(p.y), "+v"(p.z));
        if (x < blk0 || x >= blk1) p = pts[x];
        return p;
    };
    if (hl) {
        bl.set(leaf_point(G));
    } else if (!far_l) {
        bl.set(s_lo[G - blk0], s_hi[G - blk0]);
    } else {
        const uint64_t ca = code(A - base);
        bl = cell_box(ca, prefix_len(ca, code(G - base), code_bits), code_bits, mn, range, fscale);
    }
    if (hr) {
        br.set(leaf_point(G + 1));
    } else if (!far_r) {
        br.set(s_lo[G + 1 - blk0], s_hi[G + 1 - blk0]);
    } else {
        const uint64_t cb = code(B - base);
        br = cell_box(cb, prefix_len(code(G + 1 - base), cb, code_bits), code_bits, mn, range, fscale);
    }
    TravNode* t = trav + g;
    t->lo_l = make_float4(bl.lx, bl.ly, bl.lz, __int_as_float(static_cast<int>(A)));
    t->hi_l = make_float4(bl.hx, bl.hy, bl.hz, __int_as_float(static_cast<int>(G)));
    t->lo_r = make_float4(br.lx, br.ly, br.lz, __int_as_float(static_cast<int>(B)));
    t->hi_r = make_float4(br.hx, br.hy, br.hz, 0.0f);
}

// With trav, pts (the decoded points, slot numbering) and bounds (the cube:
// min at [0], range at [2]; per cloud at stride 4 with a cloud map) are
// required.
int karras_launch(const uint64_t* codes, const int32_t* d_count, int64_t n_max, int code_bits, int root_mode,
                  uint8_t* prefix_n, int32_t* left_child, uint8_t* hll, uint8_t* hlr, int32_t* parent,
                  int32_t* node_end, int32_t* leaf_parent, int2* up, TravNode* trav, const float4* pts,
                  const float* bounds, hipStream_t st, CloudMap cm = CloudMap(), int32_t* oct_count = nullptr,
                  bool claimed_parents = false) {
    if (n_max <= 0) return SX_OK;
    if (trav && (!pts || !bounds)) return fail(SX_EINVAL, "karras_launch: traversal records need points and bounds");
    // Nodes that no other node claims (only possible under the reference root
    // rule) keep parent -1, as do a batch's padding slots and the slots past
    // the unique count.  claimed_parents: the caller reads only the slots below
    // the unique count of a textbook tree, where every internal node is
    // claimed and the kernel writes -1 at the root and past the last node.
    const bool need_fill = !(claimed_parents && root_mode == SX_ROOT_TEXTBOOK && !cm.tile_cloud);
    if (parent && need_fill) SX_TRY(fill_i32(parent, -1, n_max, st));
    karras_kernel<<<grid_for(n_max, kKarrasBlock), kKarrasBlock, 0, st>>>(
        codes, d_count, n_max, code_bits, root_mode, prefix_n, left_child, hll, hlr, parent, node_end, leaf_parent, up,
        trav, pts, bounds, cm, oct_count);
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
