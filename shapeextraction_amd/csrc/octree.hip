// Octree from the radix tree, with the reference's node numbering and values.
//
// counts+offsets: edge count of radix node i = depth(i) - depth(parent(i))
//   with depth = prefixN/3, count(root) = 1 (calcEdgeCounts,
//   src/Octree.cu:38-50 and :218); offsets = exclusive scan (cub InclusiveSum
//   into offsets+1, src/Octree.cu:228-247), computed by the device scan with
//   the total landing in the octree node count.
// nodes: radix node i owns a chain of count(i) octree nodes at
//   offsets[i] .. offsets[i]+count(i)-1, deepest first, each the child of the
//   next; the top of the chain hangs under the bottom node of the nearest
//   ancestor with a non-zero count (makeNodes, src/Octree.cu:114-161), and
//   point leaves sit in that bottom node at the slot of their next 3-bit digit
//   (linkLeafNodes, src/Octree.cu:65-112).  The root (node 0) gets
//   cell_size = range (src/Octree.cu:257-261).
// Nodes are gathered, not scattered: one thread writes each node whole.
#include <hip/hip_runtime.h>

#include "sx_internal.hpp"

namespace sx {

struct OctCountIn {
    const uint8_t* pn;
    const int32_t* parent;
    const int32_t* d_count;
    __device__ int operator()(int64_t i) const {
        const int64_t M = static_cast<int64_t>(*d_count) - 1;
        if (i >= M) return 0;
        if (i == 0) return 1;
        const int p = parent[i];
        return p < 0 ? 0 : pn[i] / 3 - pn[p] / 3;  // p < 0: node unreachable from the root
    }
};

// Edge counts already written by the radix-tree kernel (textbook root).
struct OctCountGiven {
    const int32_t* counts;
    const int32_t* d_count;
    __device__ int operator()(int64_t i) const { return i < static_cast<int64_t>(*d_count) - 1 ? counts[i] : 0; }
};

struct OctOffsetOut {
    int32_t* offsets;
    __device__ void operator()(int64_t i, int ex, int) const { offsets[i] = ex; }
};

struct OctCountOut {
    int32_t* offsets;
    int32_t* counts;
    __device__ void operator()(int64_t i, int ex, int v) const {
        offsets[i] = ex;
        counts[i] = v;
    }
};

__device__ __forceinline__ float3 cell_corner(uint64_t prefix, int level, int code_bits, float mn, float range,
                                              float fscale) {
    const uint64_t c = (level == 0) ? 0ull : (prefix << (code_bits - 3 * level));
    return make_float3(dequantize(compact3(c), fscale, range, mn), dequantize(compact3(c >> 1), fscale, range, mn),
                       dequantize(compact3(c >> 2), fscale, range, mn));
}

// The same corner from the raw axis coordinates of any code inside the cell:
// the level-L prefix keeps the top L bits of each axis (B = code_bits / 3 bits
// per axis), so clearing the low B - L raw bits gives exactly the raw
// coordinates cell_corner decodes.
struct RawXYZ {
    uint32_t x, y, z;
};
__device__ __forceinline__ RawXYZ raw_xyz(uint64_t code) { return {compact3(code), compact3(code >> 1), compact3(code >> 2)}; }
__device__ __forceinline__ float3 cell_corner_raw(RawXYZ r, int level, int code_bits, float mn, float range,
                                                  float fscale) {
    const uint32_t m = ~0u << (code_bits / 3 - level);
    return make_float3(dequantize(r.x & m, fscale, range, mn), dequantize(r.y & m, fscale, range, mn),
                       dequantize(r.z & m, fscale, range, mn));
}

__device__ __forceinline__ uint64_t code_prefix(uint64_t code, int level, int code_bits) {
    const int sh = code_bits - 3 * level;
    return sh >= 64 ? 0ull : (code >> sh);
}

// Every octree node is written whole by one thread (no atomics, no zeroing).
// Radix node i with count c > 0 owns nodes offsets[i] .. offsets[i]+c-1 at
// depths D, D-1, ..., D-c+1 (D = prefixN[i]/3).  A node above the bottom of
// the chain holds exactly one child, the chain node below it: its cell
// contains only i's points.  The bottom node (depth D) gathers its children
// from i's two radix children: a leaf lands in the slot of its next digit; an
// internal child with a non-zero count contributes the top of its own chain;
// an internal child with count 0 lies at the same depth, so its children are
// gathered in turn (at most 8 slots, so at most 7 such nodes).  This is the
// node set and numbering makeNodes + linkLeafNodes produce
// (src/Octree.cu:65-161), with unused child slots left 0 as after the
// reference's cudaMemset (src/Octree.cu:253).
//
// The gather walks two or three binary levels below a node, each level one
// dependent read of the radix arrays; the descendants of a short node lie
// within a few hundred indices, so the arrays of the block's nodes +-
// kGatherHalo are staged in LDS first and only far descendants are read
// from HBM.
constexpr int kGatherBlock = 256;
constexpr int kGatherHalo = 256;
constexpr int kGatherWin = kGatherBlock + 2 * kGatherHalo;
__device__ __forceinline__ void store_node(sx_otnode* nodes, int64_t cap, int o, const int32_t ch[8], float3 corner,
                                           float cell, int node_mask, int leaf_mask) {
    if (o < 0 || o >= cap) return;
    int2* w = reinterpret_cast<int2*>(nodes + o);  // 56-byte stride: 8-byte aligned
    w[0] = make_int2(ch[0], ch[1]);
    w[1] = make_int2(ch[2], ch[3]);
    w[2] = make_int2(ch[4], ch[5]);
    w[3] = make_int2(ch[6], ch[7]);
    w[4] = make_int2(__float_as_int(corner.x), __float_as_int(corner.y));
    w[5] = make_int2(__float_as_int(corner.z), __float_as_int(cell));
    w[6] = make_int2(node_mask, leaf_mask);
}

// The nodes of a block's radix nodes form one contiguous run of the node
// array (the exclusive scan of the chain lengths); when the run fits in the
// block's LDS (reused once the gather is done) it is staged there and written
// out as contiguous 8-byte words, so a wave's stores cover whole cache lines
// instead of seven scattered 8-byte pieces per node.  Larger runs are written
// node by node.
constexpr int kGatherLds = kGatherWin * (8 + 4 + 4 + 4 + 1);
constexpr int kStageNodes = kGatherLds / static_cast<int>(sizeof(sx_otnode));
static_assert(sizeof(sx_otnode) == 56, "octree nodes are seven 8-byte words");

__device__ __forceinline__ void stage_node(int2* w, const int32_t ch[8], float3 corner, float cell, int node_mask,
                                           int leaf_mask) {
    w[0] = make_int2(ch[0], ch[1]);
    w[1] = make_int2(ch[2], ch[3]);
    w[2] = make_int2(ch[4], ch[5]);
    w[3] = make_int2(ch[6], ch[7]);
    w[4] = make_int2(__float_as_int(corner.x), __float_as_int(corner.y));
    w[5] = make_int2(__float_as_int(corner.z), __float_as_int(cell));
    w[6] = make_int2(node_mask, leaf_mask);
}

// kBatch: the slots of many clouds (cloud map cm; cloud c's radix node i at
// slot base[c] + i, counts 0 on every slot that is not an internal node).
// The scan of the counts runs over all slots, so cloud c's nodes are one run
// starting at oct_base[c]; child links are written relative to the cloud
// (octree children minus oct_base[c], leaves minus base[c]), so each cloud's
// run is byte for byte the octree of a standalone build of that cloud.
template <bool kBatch>
__global__ void __launch_bounds__(kGatherBlock) oct_gather_kernel(const uint64_t* __restrict__ codes, const int32_t* d_count,
                                                         int64_t n_max, int code_bits, const float* __restrict__ bounds,
                                                         const uint8_t* __restrict__ pn,
                                                         const int32_t* __restrict__ left_child,
                                                         const uint8_t* __restrict__ hll,
                                                         const uint8_t* __restrict__ hlr,
                                                         const int32_t* __restrict__ offsets,
                                                         const int32_t* __restrict__ counts, sx_otnode* nodes,
                                                         int64_t cap, int32_t* leaf_oct, CloudMap cm,
                                                         const int32_t* __restrict__ oct_base) {
    // the radix arrays of the block's window, then (same memory) the staged nodes
    __shared__ __align__(16) unsigned char s_raw[kGatherLds];
    uint64_t* s_code = reinterpret_cast<uint64_t*>(s_raw);
    int32_t* s_lc = reinterpret_cast<int32_t*>(s_code + kGatherWin);
    int32_t* s_cnt = s_lc + kGatherWin;
    int32_t* s_off = s_cnt + kGatherWin;
    uint8_t* s_leaf = reinterpret_cast<uint8_t*>(s_off + kGatherWin);  // bit 0: left child a leaf, bit 1: right
    const int64_t U = scan_count(d_count, n_max);
    const int64_t M = kBatch ? U : U - 1;  // slots that may hold a radix node
    // XCD-contiguous block order: a block's halo is its neighbours' nodes, read
    // from the same L2
    const int64_t blk0 = static_cast<int64_t>(xcd_block()) * kGatherBlock;
    if (blk0 >= M) return;  // whole block idle (before the barriers)
    const int64_t w0 = blk0 - kGatherHalo > 0 ? blk0 - kGatherHalo : 0;
    const int64_t w1 = blk0 + kGatherBlock + kGatherHalo < U ? blk0 + kGatherBlock + kGatherHalo : U;
    for (int64_t x = w0 + threadIdx.x; x < w1; x += kGatherBlock) {
        const int o = static_cast<int>(x - w0);
        s_code[o] = codes[x];
        if (x < M) {
            s_lc[o] = left_child[x];
            s_cnt[o] = counts[x];
            s_off[o] = offsets[x];
            s_leaf[o] = static_cast<uint8_t>((hll[x] != 0 ? 1 : 0) | (hlr[x] != 0 ? 2 : 0));
        }
    }
    __syncthreads();
    // radix arrays at x: from LDS inside [w0, w1) (node arrays below M), else
    // HBM.  The LDS read is unconditional (index clamped) and the empty asm
    // keeps it apart from the HBM read: merged they would become one generic
    // load on every access.
    auto in_win = [&](int64_t x) { return x >= w0 && x < w1; };
    const int win = static_cast<int>(w1 - w0);
    auto clampw = [&](int64_t x) {
        const int64_t o = x - w0;
        return static_cast<int>(o < 0 ? 0 : (o >= win ? win - 1 : o));
    };
    auto code_at = [&](int64_t x) {
        uint64_t v = s_code[clampw(x)];
        asm volatile("" : "+v"(v));
        if (!in_win(x)) v = codes[x];
        return v;
    };
    auto cnt_at = [&](int64_t x) {
        int v = s_cnt[clampw(x)];
        asm volatile("" : "+v"(v));
        if (!in_win(x)) v = counts[x];
        return v;
    };
    auto off_at = [&](int64_t x) {
        int v = s_off[clampw(x)];
        asm volatile("" : "+v"(v));
        if (!in_win(x)) v = offsets[x];
        return v;
    };
    // the block's run of nodes [o_begin, o_begin + n_run)
    const int64_t last = (blk0 + kGatherBlock < M ? blk0 + kGatherBlock : M) - 1;
    const int o_begin = s_off[blk0 - w0];
    const int n_run = s_off[last - w0] + s_cnt[last - w0] - o_begin;
    const bool staged = n_run <= kStageNodes;  // uniform

    const int64_t i = blk0 + threadIdx.x;
    const int c = i < M ? cnt_at(i) : 0;
    // the thread's cloud: first slot, first octree node, cube
    int64_t s0 = 0;
    int o0 = 0;
    const float* bd = bounds;
    if (kBatch && c > 0) {
        const int cid = cm.tile_cloud[i / kCloudAlign];
        s0 = cm.base[cid];
        o0 = oct_base[cid];
        bd = bounds + 4 * cid;
    }
    const float mn = bd[0], range = bd[2];
    const float fscale = static_cast<float>(axis_scale(code_bits));
    const int root_level = c > 0 ? pn[s0] / 3 : 0;
    int D = 0, base = 0;
    uint64_t ci = 0;
    int32_t ch[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) ch[s] = 0;
    int node_mask = 0, leaf_mask = 0;
    if (c > 0) {
        D = pn[i] / 3;
        base = off_at(i);
        ci = code_at(i);
        // bottom node: gather children through count-0 radix descendants
        int stack[8];
        int sp = 0;
        int cur = static_cast<int>(i);
        while (true) {
            int g;
            uint32_t leafs;
            g = s_lc[clampw(cur)];
            leafs = s_leaf[clampw(cur)];
            asm volatile("" : "+v"(g), "+v"(leafs));
            if (!in_win(cur)) {  // internal nodes are below M, so inside the window they are staged
                g = left_child[cur];
                leafs = (hll[cur] != 0 ? 1u : 0u) | (hlr[cur] != 0 ? 2u : 0u);
            }
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                const int child = g + side;
                if ((leafs >> side) & 1u) {
                    const int slot = static_cast<int>(code_prefix(code_at(child), D + 1, code_bits) & 7);
                    ch[slot] = static_cast<int32_t>(child - s0);
                    leaf_mask |= 1 << slot;
                    if (leaf_oct) leaf_oct[child] = base;
                    continue;
                }
                const int cc = cnt_at(child);
                if (cc > 0) {
                    const int slot = static_cast<int>(code_prefix(code_at(child), D + 1, code_bits) & 7);
                    ch[slot] = off_at(child) + cc - 1 - o0;
                    node_mask |= 1 << slot;
                } else if (sp < 8) {
                    stack[sp++] = child;
                }
            }
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    float3 corner = make_float3(0.f, 0.f, 0.f);
    float cell = 0.0f;
    const RawXYZ ri = raw_xyz(ci);
    if (c > 0) {
        if (i == s0) {
            corner = cell_corner(code_prefix(codes[s0], root_level, code_bits), root_level, code_bits, mn, range, fscale);
            cell = range;
        } else {
            corner = cell_corner_raw(ri, D, code_bits, mn, range, fscale);
            cell = ldexpf(range, root_level - D);
        }
    }
    // Chain nodes above the bottom hold one child each, the chain node below:
    // node base + j sits at depth D - j.
    auto chain_node = [&](int j, int32_t (&cj)[8], float3& cc, float& cs, int& cm) {
        const int L = D - j;
#pragma unroll
        for (int s = 0; s < 8; ++s) cj[s] = 0;
        const int slot = static_cast<int>(code_prefix(ci, L + 1, code_bits) & 7);
        cj[slot] = base + j - 1 - o0;
        cc = cell_corner_raw(ri, L, code_bits, mn, range, fscale);
        cs = ldexpf(range, root_level - L);
        cm = 1 << slot;
    };
    if (!staged) {
        if (c > 0) {
            int32_t cj[8];
            float3 cc;
            float cs;
            int cm;
            for (int j = 1; j < c; ++j) {
                chain_node(j, cj, cc, cs, cm);
                store_node(nodes, cap, base + j, cj, cc, cs, cm, 0);
            }
            store_node(nodes, cap, base, ch, corner, cell, node_mask, leaf_mask);
        }
        return;  // uniform: the whole block takes this path
    }
    __syncthreads();  // every read of the staged radix arrays is done
    int2* stage = reinterpret_cast<int2*>(s_raw);
    if (c > 0) {
        int32_t cj[8];
        float3 cc;
        float cs;
        int cm;
        for (int j = 1; j < c; ++j) {
            chain_node(j, cj, cc, cs, cm);
            stage_node(stage + 7 * (base + j - o_begin), cj, cc, cs, cm, 0);
        }
        stage_node(stage + 7 * (base - o_begin), ch, corner, cell, node_mask, leaf_mask);
    }
    __syncthreads();
    int2* out = reinterpret_cast<int2*>(nodes) + 7 * static_cast<int64_t>(o_begin);
    const int64_t words = 7 * static_cast<int64_t>(n_run);
    const int64_t cap_words = 7 * (cap - o_begin);
    for (int e = static_cast<int>(threadIdx.x); e < words; e += kGatherBlock)
        if (e < cap_words) out[e] = stage[e];
}

size_t octree_ws_bytes(int64_t n_max) {
    Carver c(nullptr);
    c.take<int32_t>(static_cast<size_t>(n_max));  // offsets
    c.take<int32_t>(static_cast<size_t>(n_max));  // counts
    c.take<char>(scan_ws_bytes(n_max));
    return c.off;
}

// offsets/counts (chain start and length of each radix node) are carved
// from ws unless the caller keeps them (keep_offsets, keep_counts); with
// counts_given the radix-tree kernel has already written the counts into
// keep_counts, and only their scan remains.
int octree_launch(const uint64_t* codes, const int32_t* d_count, int64_t n_max, int code_bits, const float* bounds,
                  const uint8_t* pn, const int32_t* left_child, const uint8_t* hll, const uint8_t* hlr,
                  const int32_t* parent, sx_otnode* nodes, int64_t cap, int32_t* d_oct_count, int32_t* leaf_oct,
                  void* ws, hipStream_t st, int32_t* keep_offsets = nullptr, int32_t* keep_counts = nullptr,
                  bool counts_given = false) {
    if (n_max <= 0) return SX_OK;
    Carver c(ws);
    int32_t* offsets = c.take<int32_t>(static_cast<size_t>(n_max));
    int32_t* counts = c.take<int32_t>(static_cast<size_t>(n_max));
    void* scan_ws = c.take<char>(scan_ws_bytes(n_max));
    if (keep_offsets) offsets = keep_offsets;
    if (keep_counts) counts = keep_counts;
    if (counts_given && keep_counts)
        SX_TRY(device_scan(OctCountGiven{keep_counts, d_count}, OctOffsetOut{offsets}, d_count, n_max, d_oct_count,
                           scan_ws, st));
    else
        SX_TRY(device_scan(OctCountIn{pn, parent, d_count}, OctCountOut{offsets, counts}, d_count, n_max,
                           d_oct_count, scan_ws, st));
    oct_gather_kernel<false><<<grid_for(n_max, kGatherBlock), kGatherBlock, 0, st>>>(
        codes, d_count, n_max, code_bits, bounds, pn, left_child, hll, hlr, offsets, counts, nodes, cap, leaf_oct,
        CloudMap(), nullptr);
    SX_LAUNCH_CHECK("oct_gather_kernel");
    return SX_OK;
}

// Edge counts of a batch's slots, written by the radix-tree kernel (0 on every
// slot that is not an internal node, padding included).
struct OctCountSlots {
    const int32_t* counts;
    const int32_t* d_slots;
    __device__ int operator()(int64_t i) const { return i < static_cast<int64_t>(*d_slots) ? counts[i] : 0; }
};

// First octree node and node count of every cloud: the scanned offset at its
// first slot, up to the offset at its first slot past its unique points
// (padding slots count 0), or the batch total past the last slot.
__global__ void __launch_bounds__(256) cloud_oct_kernel(const int32_t* __restrict__ offsets,
                                                        const int32_t* __restrict__ d_slots,
                                                        const int32_t* __restrict__ d_total, CloudMap cm,
                                                        int n_clouds, int32_t* __restrict__ oct_base,
                                                        int32_t* __restrict__ oct_size) {
    const int c = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= n_clouds) return;
    const int64_t slots = *d_slots, total = *d_total;
    const int64_t s0 = cm.base[c], s1 = s0 + cm.size[c];
    const int64_t o0 = s0 < slots ? offsets[s0] : total;
    const int64_t o1 = s1 < slots ? offsets[s1] : total;
    oct_base[c] = static_cast<int32_t>(o0);
    oct_size[c] = static_cast<int32_t>(o1 - o0);
}

// The octrees of a batch's clouds in one pass over the slots: the counts
// (given by the radix-tree kernel) are scanned over every slot, then each
// cloud's node run is located and gathered with cloud-relative links.
// offsets: the chain start of each slot (kept for the feature links).
int batch_octree_launch(const uint64_t* codes, const int32_t* d_slots, int64_t slot_cap, int code_bits,
                        const float* bounds, const uint8_t* pn, const int32_t* left_child, const uint8_t* hll,
                        const uint8_t* hlr, const int32_t* counts, int32_t* offsets, CloudMap cm, int n_clouds,
                        int32_t* oct_base, int32_t* oct_size, sx_otnode* nodes, int64_t cap, int32_t* d_oct_count,
                        void* scan_ws, hipStream_t st) {
    if (slot_cap <= 0 || n_clouds <= 0) return SX_OK;
    SX_TRY(device_scan(OctCountSlots{counts, d_slots}, OctOffsetOut{offsets}, d_slots, slot_cap, d_oct_count, scan_ws,
                       st));
    cloud_oct_kernel<<<grid_for(n_clouds, 256), 256, 0, st>>>(offsets, d_slots, d_oct_count, cm, n_clouds, oct_base,
                                                              oct_size);
    SX_LAUNCH_CHECK("cloud_oct_kernel");
    oct_gather_kernel<true><<<grid_for(slot_cap, kGatherBlock), kGatherBlock, 0, st>>>(
        codes, d_slots, slot_cap, code_bits, bounds, pn, left_child, hll, hlr, offsets, counts, nodes, cap, nullptr,
        cm, oct_base);
    SX_LAUNCH_CHECK("oct_gather_kernel");
    return SX_OK;
}

// The octree links the learned-radius features read, for an octree built on
// the traversal tree itself (textbook root), from the arrays the build keeps:
//   leaf_oct[u]: the octree node holding point u - the bottom chain node
//     (offsets[q]) of the nearest radix ancestor q of u with a non-zero
//     chain, as the gather places leaves;
//   aux[o] = (parent octree node, points in the cell) for every chain node:
//     a chain node's parent is the next one up, the top's parent is the
//     bottom node of the nearest ancestor with a chain (-1 at the root), and
//     every node of radix node i's chain holds exactly i's leaf range.
// The walks are short: a radix node without a chain lies at its parent's
// octree depth, and prefix lengths grow by at least one bit per level.
// root: the slot of the tree's root (its count is 1, so every walk stops there).
__device__ __forceinline__ int chain_holder(const int32_t* __restrict__ parent, const int32_t* __restrict__ counts,
                                            int p, int count_p, int root) {
#pragma unroll 1
    for (int s = 0; s < 64 && p > root && count_p == 0; ++s) {
        p = parent[p];
        count_p = p >= 0 ? counts[p] : 1;
    }
    return p < 0 ? root : p;
}

// One thread per point t (leaf link) and radix node t (its chain's links);
// the loads of each of the three steps are independent, so a thread waits for three
// memory round trips when no walk is needed (most threads).  With a cloud
// map, t is a slot and each cloud's tree is walked from its own root (the
// links stay in slot and batch-octree numbering).
__global__ void __launch_bounds__(256) feature_links_kernel(const int32_t* d_count, int64_t n_max,
                                                            const int32_t* __restrict__ parent,
                                                            const int32_t* __restrict__ leaf_parent,
                                                            const int32_t* __restrict__ node_end,
                                                            const int32_t* __restrict__ offsets,
                                                            const int32_t* __restrict__ counts, int64_t cap,
                                                            int32_t* __restrict__ leaf_oct, int2* __restrict__ aux,
                                                            CloudMap cm) {
    const int64_t total = scan_count(d_count, n_max);
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int2 cloud = cm.at(t, static_cast<int>(total));
    const int64_t U = cloud.y, li = t - cloud.x;
    if (U < 2 || li >= U) return;
    const int i = static_cast<int>(t), root = cloud.x;
    const bool node = li < U - 1;
    // first loads (independent of each other)
    const int lp = leaf_parent[i] >> 1;
    const int c = node ? counts[i] : 0;
    const int pa = node && li > 0 ? parent[i] : root;
    const int base = node ? offsets[i] : 0;
    const int e = node ? node_end[i] : i;
    // second loads
    const int c_lp = counts[lp];
    const int c_pa = counts[pa];
    const int hl = chain_holder(parent, counts, lp, c_lp, root);
    const int hp = chain_holder(parent, counts, pa, c_pa, root);
    // third loads
    leaf_oct[i] = offsets[hl];
    if (c <= 0) return;
    const int top_parent = li == 0 ? -1 : offsets[hp];
    const int sz = (e > i ? e - i : i - e) + 1;
    for (int j = 0; j < c; ++j) {
        if (base + j >= cap) break;
        aux[base + j] = make_int2(j + 1 < c ? base + j + 1 : top_parent, sz);
    }
}

int feature_links_launch(const int32_t* d_count, int64_t n_max, const int32_t* parent, const int32_t* leaf_parent,
                         const int32_t* node_end, const int32_t* offsets, const int32_t* counts, int64_t cap,
                         int32_t* leaf_oct, int2* aux, hipStream_t st, CloudMap cm) {
    if (n_max <= 0) return SX_OK;
    feature_links_kernel<<<grid_for(n_max, 256), 256, 0, st>>>(d_count, n_max, parent, leaf_parent, node_end, offsets,
                                                               counts, cap, leaf_oct, aux, cm);
    SX_LAUNCH_CHECK("feature_links_kernel");
    return SX_OK;
}

}  // namespace sx

using sx::fail;

extern "C" {

int sx_build_octree(const uint64_t* d_codes, const int32_t* d_count, int64_t n_max, int code_bits,
                    const float* d_bounds, const uint8_t* d_prefix_n, const int32_t* d_left_child,
                    const uint8_t* d_has_leaf_left, const uint8_t* d_has_leaf_right, const int32_t* d_parent,
                    sx_otnode* d_nodes, int64_t oct_capacity, int32_t* d_oct_count, void* ws, size_t* ws_bytes,
                    void* stream) {
    if (!ws_bytes) return fail(SX_EINVAL, "sx_build_octree: ws_bytes is null");
    if (n_max <= 0) return fail(SX_EINVAL, "sx_build_octree: n_max must be positive");
    const size_t need = sx::octree_ws_bytes(n_max);
    if (!ws) {
        *ws_bytes = need;
        return SX_OK;
    }
    if (*ws_bytes < need) return fail(SX_ENOMEM, "sx_build_octree: workspace too small");
    if (code_bits != 30 && code_bits != 63) return fail(SX_EINVAL, "sx_build_octree: code_bits must be 30 or 63");
    if (!d_codes || !d_count || !d_bounds || !d_prefix_n || !d_left_child || !d_has_leaf_left || !d_has_leaf_right ||
        !d_parent || !d_nodes || !d_oct_count || oct_capacity <= 0)
        return fail(SX_EINVAL, "sx_build_octree: bad arguments");
    return sx::octree_launch(d_codes, d_count, n_max, code_bits, d_bounds, d_prefix_n, d_left_child, d_has_leaf_left,
                             d_has_leaf_right, d_parent, d_nodes, oct_capacity, d_oct_count, nullptr, ws,
                             sx::as_stream(stream));
}

int sx_check_counts(const int32_t* d_count, int64_t count_capacity, const int32_t* d_oct_count, int64_t oct_capacity,
                    const int32_t* d_bad, int64_t* count_out, int64_t* oct_count_out, void* stream) {
    hipStream_t st = sx::as_stream(stream);
    int32_t h[3] = {0, 0, 0};
    if (d_count) SX_HIP(hipMemcpyAsync(&h[0], d_count, 4, hipMemcpyDeviceToHost, st));
    if (d_oct_count) SX_HIP(hipMemcpyAsync(&h[1], d_oct_count, 4, hipMemcpyDeviceToHost, st));
    if (d_bad) SX_HIP(hipMemcpyAsync(&h[2], d_bad, 4, hipMemcpyDeviceToHost, st));
    SX_HIP(hipStreamSynchronize(st));
    if (count_out) *count_out = h[0];
    if (oct_count_out) *oct_count_out = h[1];
    if (d_bad && h[2] != 0) return fail(SX_EDATA, std::to_string(h[2]) + " non-finite coordinates in the input");
    if (d_count && (h[0] < 0 || h[0] > count_capacity)) return fail(SX_ERANGE, "point count exceeds capacity");
    if (d_oct_count && (h[1] < 0 || h[1] > oct_capacity))
        return fail(SX_ERANGE, "octree needs " + std::to_string(h[1]) + " nodes, capacity is " +
                                   std::to_string(oct_capacity));
    return SX_OK;
}

}  // extern "C"
