// Exact k-nearest-neighbour search, one query per lane.
//
// Replaces knnSearchKernel<k> and its two per-thread min-max heaps in local
// memory (src/Octree.hpp:136-193, src/dequeue.hpp).  The reference search
// never updates its distance bound (src/Octree.hpp:152-155) and evicts from a
// 32-entry search queue when full (src/dequeue.hpp:60-68), so it is not exact;
// this kernel is.
//
// Structure searched: the Karras radix tree of the sorted unique points with
// a tight box per child (TravNode, 64 B per internal node).  A child whose
// leaf range holds at most `bucket` points is scanned linearly (its points are
// contiguous in Morton order) instead of descended.
//
// Traversal: nearest child first, depth first, stackless.  Returning from a
// child to its parent re-derives which child was nearer (the same
// comparison as on the way down), so the far child is visited after the near
// one without any stack in scratch or LDS; the parent index sits in the
// record.  A subtree is skipped when its box distance exceeds the current
// bound, min(R^2, k-th best key); ties on distance survive because pruning is
// strict and box distances are exact lower bounds (see box_dist2).
//
// Result list: K (squared distance, index) pairs packed as 64-bit keys
// (float bits << 32 | index; squared distances are non-negative so the bit
// order is the numeric order), kept sorted in registers with a branch-free
// shift insert.  Ties are therefore broken by point index.
//
// Radius seeding (SX_KNN_RADIUS): the bound starts at R^2.  If the final
// k-th distance is <= R^2, every pruned point was strictly farther than it,
// so the list is exact; otherwise the query is searched again unseeded.
#include <hip/hip_runtime.h>

#include "sx_internal.hpp"

namespace sx {

template <int K>
struct TopK {
    uint64_t e[K];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int s = 0; s < K; ++s) e[s] = ~0ull;
    }
    __device__ __forceinline__ uint32_t kth_bits() const { return static_cast<uint32_t>(e[K - 1] >> 32); }
    __device__ __forceinline__ void insert(uint64_t key) {
#pragma unroll
        for (int s = K - 1; s > 0; --s) {
            const uint64_t prev = e[s - 1];
            const uint64_t cur = e[s];
            e[s] = key < prev ? prev : (key < cur ? key : cur);
        }
        e[0] = key < e[0] ? key : e[0];
    }
};

template <int K>
__device__ __forceinline__ void scan_range(const float4* __restrict__ pts, int lo, int hi, float qx, float qy,
                                           float qz, TopK<K>& top) {
    for (int idx = lo; idx <= hi; ++idx) {
        const float4 p = pts[idx];
        const float d = dist2(qx, qy, qz, p.x, p.y, p.z);
        const uint64_t key = (static_cast<uint64_t>(__float_as_uint(d)) << 32) | static_cast<uint32_t>(idx);
        if (key < top.e[K - 1]) top.insert(key);
    }
}

template <int K>
__device__ void knn_search(const float4* __restrict__ pts, const TravNode* __restrict__ nodes, int U, float qx,
                           float qy, float qz, uint32_t r2bits, int bucket, TopK<K>& top) {
    if (U <= bucket) {
        scan_range<K>(pts, 0, U - 1, qx, qy, qz, top);
        return;
    }
    const int* node_words = reinterpret_cast<const int*>(nodes);
    int cur = 0;
    int from = -1;
    bool down = true;
    while (true) {
        const uint32_t kb = top.kth_bits();
        const uint32_t bound = kb < r2bits ? kb : r2bits;
        if (down) {
            const TravNode n = nodes[cur];
            const int a = w_int(n.lo_l.w), g = w_int(n.hi_l.w), b = w_int(n.lo_r.w);
            const float dl = box_dist2(qx, qy, qz, n.lo_l, n.hi_l);
            const float dr = box_dist2(qx, qy, qz, n.lo_r, n.hi_r);
            const bool lf = dl <= dr;
            const uint32_t nd = __float_as_uint(lf ? dl : dr);
            if (nd > bound) {
                from = cur;
                down = false;
                continue;
            }
            const int nlo = lf ? a : g + 1, nhi = lf ? g : b;
            if (nhi - nlo < bucket) {
                scan_range<K>(pts, nlo, nhi, qx, qy, qz, top);
                const uint32_t kb2 = top.kth_bits();
                const uint32_t bound2 = kb2 < r2bits ? kb2 : r2bits;
                const uint32_t fd = __float_as_uint(lf ? dr : dl);
                if (fd <= bound2) {
                    const int flo = lf ? g + 1 : a, fhi = lf ? b : g;
                    if (fhi - flo < bucket) {
                        scan_range<K>(pts, flo, fhi, qx, qy, qz, top);
                    } else {
                        cur = lf ? g + 1 : g;
                        continue;
                    }
                }
                from = cur;
                down = false;
                continue;
            }
            cur = lf ? g : g + 1;
        } else {
            if (from == 0) break;
            const int p = node_words[static_cast<int64_t>(from) * 16 + 15];  // parent, in hi_r.w
            const TravNode n = nodes[p];
            const int a = w_int(n.lo_l.w), g = w_int(n.hi_l.w), b = w_int(n.lo_r.w);
            const float dl = box_dist2(qx, qy, qz, n.lo_l, n.hi_l);
            const float dr = box_dist2(qx, qy, qz, n.lo_r, n.hi_r);
            const bool lf = dl <= dr;
            const bool from_left = (from == g);
            if (from_left == lf) {
                const uint32_t fd = __float_as_uint(lf ? dr : dl);
                if (fd <= bound) {
                    const int flo = lf ? g + 1 : a, fhi = lf ? b : g;
                    if (fhi - flo < bucket) {
                        scan_range<K>(pts, flo, fhi, qx, qy, qz, top);
                    } else {
                        cur = lf ? g + 1 : g;
                        down = true;
                        continue;
                    }
                }
            }
            from = p;
        }
    }
}

constexpr int kKnnBlock = 256;

template <int K>
__global__ void __launch_bounds__(kKnnBlock) knn_kernel(const float4* __restrict__ pts,
                                                        const TravNode* __restrict__ nodes,
                                                        const int32_t* __restrict__ d_counts, int64_t q_start,
                                                        int64_t n_q, const float* __restrict__ radius2, int bucket,
                                                        int32_t* __restrict__ out, int32_t* __restrict__ reruns) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_q) return;
    const int U = d_counts[0];
    const int64_t qi = q_start + t;
    if (qi >= U || U <= 0) return;
    const float4 q = pts[qi];
    TopK<K> top;
    top.init();
    uint32_t r2bits = 0xFFFFFFFFu;
    if (radius2) {
        const float r2 = radius2[t];
        r2bits = (r2 >= 0.0f) ? __float_as_uint(r2) : 0xFFFFFFFFu;
    }
    knn_search<K>(pts, nodes, U, q.x, q.y, q.z, r2bits, bucket, top);
    if (r2bits != 0xFFFFFFFFu && top.kth_bits() > r2bits) {
        // Fewer than K points within R: the seeded result is not proven exact.
        if (reruns) atomicAdd(reruns, 1);
        top.init();
        knn_search<K>(pts, nodes, U, q.x, q.y, q.z, 0xFFFFFFFFu, bucket, top);
    }
    int32_t* row = out + t * K;
    const int32_t first = static_cast<int32_t>(top.e[0] & 0xFFFFFFFFull);
#pragma unroll
    for (int s = 0; s < K; s += 4) {
        int4 v;
        v.x = top.e[s] == ~0ull ? first : static_cast<int32_t>(top.e[s] & 0xFFFFFFFFull);
        v.y = top.e[s + 1] == ~0ull ? first : static_cast<int32_t>(top.e[s + 1] & 0xFFFFFFFFull);
        v.z = top.e[s + 2] == ~0ull ? first : static_cast<int32_t>(top.e[s + 2] & 0xFFFFFFFFull);
        v.w = top.e[s + 3] == ~0ull ? first : static_cast<int32_t>(top.e[s + 3] & 0xFFFFFFFFull);
        *reinterpret_cast<int4*>(row + s) = v;
    }
}

int knn_launch(const float4* pts, const TravNode* nodes, const int32_t* d_counts, int64_t q_start, int64_t n_q, int k,
               const float* radius2, int bucket, int32_t* out, int32_t* reruns, hipStream_t st) {
    if (n_q <= 0) return SX_OK;
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0) return fail(SX_EINVAL, "knn: output must be 16-byte aligned");
    const unsigned g = grid_for(n_q, kKnnBlock);
    switch (k) {
        case 8: knn_kernel<8><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, radius2, bucket, out, reruns); break;
        case 16: knn_kernel<16><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, radius2, bucket, out, reruns); break;
        case 24: knn_kernel<24><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, radius2, bucket, out, reruns); break;
        case 32: knn_kernel<32><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, radius2, bucket, out, reruns); break;
        default: return fail(SX_EINVAL, "knn: k must be 8, 16, 24 or 32");
    }
    SX_LAUNCH_CHECK("knn_kernel");
    return SX_OK;
}

}  // namespace sx
