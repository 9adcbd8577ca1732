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

// One query's exact search.  State machine with one record load per step:
//   down at `node`: pick the nearer child; a bucket child is scanned (next
//     step) and counts as returned-from; an internal child is entered.
//   up at `p` coming back from child `from`: if `from` was the nearer child,
//     the farther one is visited (scanned or entered) when not pruned;
//     otherwise p is finished and the walk returns to p's parent.
// `from` is compared with the split gamma of p: the left child's index is
// gamma and the right child's gamma+1, for leaves and internal nodes alike.
template <int K>
__device__ void knn_search(const float4* __restrict__ pts, const TravNode* __restrict__ nodes, int U, float qx,
                           float qy, float qz, uint32_t r2bits, int bucket, TopK<K>& top) {
    bool scan = U <= bucket;
    bool done = scan;
    int slo = 0, shi = U - 1;
    int node = 0, p = -1, from = -1;
    bool down = true;
    while (true) {
        if (scan) {
            scan_range<K>(pts, slo, shi, qx, qy, qz, top);
            scan = false;
            if (done) break;
        }
        const uint32_t kb = top.kth_bits();
        const uint32_t bound = kb < r2bits ? kb : r2bits;
        const TravNode n = nodes[down ? node : p];
        const int a = w_int(n.lo_l.w), g = w_int(n.hi_l.w), b = w_int(n.lo_r.w), parent = w_int(n.hi_r.w);
        const float dl = box_dist2(qx, qy, qz, n.lo_l, n.hi_l);
        const float dr = box_dist2(qx, qy, qz, n.lo_r, n.hi_r);
        const bool lf = dl <= dr;
        if (down) {
            if (__float_as_uint(lf ? dl : dr) > bound) {  // both children pruned
                from = node;
                p = parent;
                down = false;
                if (p < 0) break;
                continue;
            }
            const int near = lf ? g : g + 1;
            const int nlo = lf ? a : g + 1, nhi = lf ? g : b;
            if (nhi - nlo < bucket) {
                scan = true;
                slo = nlo;
                shi = nhi;
                from = near;
                p = node;
                down = false;
            } else {
                node = near;
            }
        } else {
            if ((from == g) == lf) {  // returned from the nearer child: the farther one is next
                if (__float_as_uint(lf ? dr : dl) <= bound) {
                    const int far = lf ? g + 1 : g;
                    const int flo = lf ? g + 1 : a, fhi = lf ? b : g;
                    if (fhi - flo < bucket) {
                        scan = true;
                        slo = flo;
                        shi = fhi;
                        from = far;
                    } else {
                        node = far;
                        down = true;
                    }
                    continue;
                }
            }
            if (parent < 0) break;  // p (the root) is finished
            from = p;
            p = parent;
        }
    }
}

constexpr int kKnnBlock = 256;

template <int K, bool kExternal>
__global__ void __launch_bounds__(kKnnBlock) knn_kernel(const float4* __restrict__ pts,
                                                        const TravNode* __restrict__ nodes,
                                                        const int32_t* __restrict__ d_counts, int64_t q_start,
                                                        int64_t n_q, const float* __restrict__ queries,
                                                        const float* __restrict__ radius2, int bucket,
                                                        int32_t* __restrict__ out, int32_t* __restrict__ reruns) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_q) return;
    const int U = d_counts[0];
    if (U <= 0) return;
    float qx, qy, qz;
    if (kExternal) {
        qx = queries[3 * t];
        qy = queries[3 * t + 1];
        qz = queries[3 * t + 2];
    } else {
        const int64_t qi = q_start + t;
        if (qi >= U) return;
        const float4 q = pts[qi];
        qx = q.x;
        qy = q.y;
        qz = q.z;
    }
    TopK<K> top;
    top.init();
    uint32_t r2bits = 0xFFFFFFFFu;
    if (radius2) {
        const float r2 = radius2[t];
        r2bits = (r2 >= 0.0f) ? __float_as_uint(r2) : 0xFFFFFFFFu;
    }
    // A seeded search that finds fewer than K points within R is not proven
    // exact: it is repeated unseeded (one loop so the search is inlined once).
    for (int attempt = 0; attempt < 2; ++attempt) {
        knn_search<K>(pts, nodes, U, qx, qy, qz, r2bits, bucket, top);
        if (r2bits == 0xFFFFFFFFu || top.kth_bits() <= r2bits) break;
        if (reruns) atomicAdd(reruns, 1);
        top.init();
        r2bits = 0xFFFFFFFFu;
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

template <bool kExternal>
int knn_dispatch(const float4* pts, const TravNode* nodes, const int32_t* d_counts, int64_t q_start, int64_t n_q,
                 const float* queries, int k, const float* radius2, int bucket, int32_t* out, int32_t* reruns,
                 hipStream_t st) {
    const unsigned g = grid_for(n_q, kKnnBlock);
    switch (k) {
        case 8: knn_kernel<8, kExternal><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, queries, radius2, bucket, out, reruns); break;
        case 16: knn_kernel<16, kExternal><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, queries, radius2, bucket, out, reruns); break;
        case 24: knn_kernel<24, kExternal><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, queries, radius2, bucket, out, reruns); break;
        case 32: knn_kernel<32, kExternal><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, queries, radius2, bucket, out, reruns); break;
        default: return fail(SX_EINVAL, "knn: k must be 8, 16, 24 or 32");
    }
    SX_LAUNCH_CHECK("knn_kernel");
    return SX_OK;
}

// queries == nullptr: the index points [q_start, q_start + n_q) are the
// queries; otherwise n_q external points (12-byte AoS).
int knn_launch(const float4* pts, const TravNode* nodes, const int32_t* d_counts, int64_t q_start, int64_t n_q,
               const float* queries, int k, const float* radius2, int bucket, int32_t* out, int32_t* reruns,
               hipStream_t st) {
    if (n_q <= 0) return SX_OK;
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0) return fail(SX_EINVAL, "knn: output must be 16-byte aligned");
    if (queries)
        return knn_dispatch<true>(pts, nodes, d_counts, q_start, n_q, queries, k, radius2, bucket, out, reruns, st);
    return knn_dispatch<false>(pts, nodes, d_counts, q_start, n_q, nullptr, k, radius2, bucket, out, reruns, st);
}

}  // namespace sx
