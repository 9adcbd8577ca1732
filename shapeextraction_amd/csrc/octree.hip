// Octree from the radix tree, with the reference's node numbering and values.
//
// counts+offsets: edge count of radix node i = depth(i) - depth(parent(i))
//   with depth = prefixN/3, count(root) = 1 (calcEdgeCounts,
//   src/Octree.cu:38-50 and :218); offsets = exclusive scan (cub InclusiveSum
//   into offsets+1, src/Octree.cu:228-247), computed by the device scan with
//   the total landing in the octree node count.
// nodes: radix node i emits a chain of count(i) octree nodes at
//   offsets[i] .. offsets[i]+count(i)-1, deepest first, each the child of the
//   next; the top of the chain hangs under the bottom node of the nearest
//   ancestor with a non-zero count (makeNodes, src/Octree.cu:114-161).  The
//   root (node 0) gets cell_size = range (src/Octree.cu:257-261).
// leaves: point leaves go into the bottom node of the nearest contributing
//   ancestor, slot = the leaf's next 3-bit digit (linkLeafNodes,
//   src/Octree.cu:65-112).
// Masks are set with atomicOr (several radix nodes write one octree node);
// every slot is written by exactly one thread.
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

__device__ __forceinline__ uint64_t code_prefix(uint64_t code, int level, int code_bits) {
    const int sh = code_bits - 3 * level;
    return sh >= 64 ? 0ull : (code >> sh);
}

__global__ void __launch_bounds__(256) oct_zero_kernel(sx_otnode* nodes, const int32_t* d_oct_count, int64_t cap) {
    int64_t O = *d_oct_count;
    if (O > cap) O = cap;
    const int64_t words = O * static_cast<int64_t>(sizeof(sx_otnode) / 4);
    int32_t* w = reinterpret_cast<int32_t*>(nodes);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) w[i] = 0;
}

__device__ __forceinline__ void set_child(sx_otnode* nodes, int64_t cap, int node, int child, int slot) {
    if (node < 0 || node >= cap) return;
    nodes[node].children[slot] = child;
    atomicOr(&nodes[node].child_node_mask, 1 << slot);
}

__device__ __forceinline__ void set_leaf(sx_otnode* nodes, int64_t cap, int node, int leaf, int slot) {
    if (node < 0 || node >= cap) return;
    nodes[node].children[slot] = leaf;
    atomicOr(&nodes[node].child_leaf_mask, 1 << slot);
}

__device__ __forceinline__ void set_cell(sx_otnode* nodes, int64_t cap, int node, float3 corner, float cell) {
    if (node < 0 || node >= cap) return;
    nodes[node].corner[0] = corner.x;
    nodes[node].corner[1] = corner.y;
    nodes[node].corner[2] = corner.z;
    nodes[node].cell_size = cell;
}

__global__ void __launch_bounds__(256) oct_build_kernel(const uint64_t* __restrict__ codes, const int32_t* d_count,
                                                        int64_t n_max, int code_bits, const float* __restrict__ bounds,
                                                        const uint8_t* __restrict__ pn,
                                                        const int32_t* __restrict__ left_child,
                                                        const uint8_t* __restrict__ hll,
                                                        const uint8_t* __restrict__ hlr,
                                                        const int32_t* __restrict__ parent,
                                                        const int32_t* __restrict__ offsets,
                                                        const int32_t* __restrict__ counts, sx_otnode* nodes,
                                                        int64_t cap, int32_t* leaf_oct) {
    const int64_t U = scan_count(d_count, n_max);
    const int64_t M = U - 1;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const float mn = bounds[0], range = bounds[2];
    const float fscale = static_cast<float>(axis_scale(code_bits));
    const int root_level = pn[0] / 3;
    if (i == 0) {
        const uint64_t rp = code_prefix(codes[0], root_level, code_bits);
        set_cell(nodes, cap, 0, cell_corner(rp, root_level, code_bits, mn, range, fscale), range);
    } else {
        int oct = offsets[i];
        const int n_new = counts[i];
        const uint64_t ci = codes[i];
        const int depth = pn[i] / 3;
        for (int jj = 0; jj < n_new - 1; ++jj) {
            const int level = depth - jj;
            const uint64_t pref = code_prefix(ci, level, code_bits);
            set_child(nodes, cap, oct + 1, oct, static_cast<int>(pref & 7));
            set_cell(nodes, cap, oct, cell_corner(pref, level, code_bits, mn, range, fscale),
                     __fdiv_rn(range, static_cast<float>(1 << (level - root_level))));
            oct += 1;
        }
        if (n_new > 0) {
            int rp = parent[i];
            while (rp >= 0 && counts[rp] == 0) rp = parent[rp];
            const int top = depth - n_new + 1;
            const uint64_t pref = code_prefix(ci, top, code_bits);
            if (rp >= 0) set_child(nodes, cap, offsets[rp], oct, static_cast<int>(pref & 7));
            set_cell(nodes, cap, oct, cell_corner(pref, top, code_bits, mn, range, fscale),
                     __fdiv_rn(range, static_cast<float>(1 << (top - root_level))));
        }
    }
    if (hll[i] || hlr[i]) {
        int r = static_cast<int>(i);
        while (r >= 0 && counts[r] == 0) r = parent[r];
        const int bottom = r >= 0 ? offsets[r] : -1;
        const int level = pn[i] / 3 + 1;
        if (hll[i]) {
            const int leaf = left_child[i];
            set_leaf(nodes, cap, bottom, leaf, static_cast<int>(code_prefix(codes[leaf], level, code_bits) & 7));
            if (leaf_oct) leaf_oct[leaf] = bottom;
        }
        if (hlr[i]) {
            const int leaf = left_child[i] + 1;
            set_leaf(nodes, cap, bottom, leaf, static_cast<int>(code_prefix(codes[leaf], level, code_bits) & 7));
            if (leaf_oct) leaf_oct[leaf] = bottom;
        }
    }
}

size_t octree_ws_bytes(int64_t n_max) {
    Carver c(nullptr);
    c.take<int32_t>(static_cast<size_t>(n_max));  // offsets
    c.take<int32_t>(static_cast<size_t>(n_max));  // counts
    c.take<char>(scan_ws_bytes(n_max));
    return c.off;
}

// offsets/counts are carved from ws unless given (the index keeps them).
int octree_launch(const uint64_t* codes, const int32_t* d_count, int64_t n_max, int code_bits, const float* bounds,
                  const uint8_t* pn, const int32_t* left_child, const uint8_t* hll, const uint8_t* hlr,
                  const int32_t* parent, sx_otnode* nodes, int64_t cap, int32_t* d_oct_count, int32_t* leaf_oct,
                  void* ws, hipStream_t st) {
    if (n_max <= 0) return SX_OK;
    Carver c(ws);
    int32_t* offsets = c.take<int32_t>(static_cast<size_t>(n_max));
    int32_t* counts = c.take<int32_t>(static_cast<size_t>(n_max));
    void* scan_ws = c.take<char>(scan_ws_bytes(n_max));
    SX_TRY(device_scan(OctCountIn{pn, parent, d_count}, OctCountOut{offsets, counts}, d_count, n_max, d_oct_count,
                       scan_ws, st));
    unsigned zg = grid_for(cap * static_cast<int64_t>(sizeof(sx_otnode) / 4), 256);
    if (zg > 8192) zg = 8192;
    oct_zero_kernel<<<zg, 256, 0, st>>>(nodes, d_oct_count, cap);
    SX_LAUNCH_CHECK("oct_zero_kernel");
    oct_build_kernel<<<grid_for(n_max, 256), 256, 0, st>>>(codes, d_count, n_max, code_bits, bounds, pn, left_child,
                                                           hll, hlr, parent, offsets, counts, nodes, cap, leaf_oct);
    SX_LAUNCH_CHECK("oct_build_kernel");
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
