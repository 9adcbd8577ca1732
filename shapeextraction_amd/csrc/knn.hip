// Exact k-nearest-neighbour search.  Three kernels share one result format:
//   knn_kernel       one query per lane, stackless depth-first descent (below);
//   knn_tile_kernel  64 Morton-consecutive queries per wave share one range
//                    query and one candidate stream (the default for index
//                    queries);
//   knn_group_kernel one 16-lane group per query, for the queries a tile
//                    defers.
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
// one without any stack in scratch or LDS; parents come from the radix
// tree's parent array.  A subtree is skipped when its box distance exceeds the current
// bound, min(R^2, k-th best key); ties on distance survive because pruning is
// strict and box distances are exact lower bounds (see box_dist2).
//
// Result list: K (squared distance, index) pairs packed as 64-bit keys
// (float bits << 32 | index; squared distances are non-negative so the bit
// order is the numeric order), kept sorted in registers with a branch-free
// shift insert.  Ties are therefore broken by point index.
//
// Read as an IEEE double, such a key is a non-negative number and never a
// NaN: its high word is the bits of a non-negative float distance, below
// 0x7FF00000 even for +inf.  The numeric order of those doubles is the
// unsigned order of the keys (denormal doubles included: FP64 denormals are
// not flushed), so a 64-bit min or max is one v_min_f64 / v_max_f64 instead
// of a 64-bit compare and four selects.  The empty key is kNoKey (+inf as a
// double), above every key of a real point.
//
// Radius seeding (SX_KNN_RADIUS): the bound starts at R^2.  If the final
// k-th distance is <= R^2, every pruned point was strictly farther than it,
// so the list is exact; otherwise the query is searched again unseeded.
#include <hip/hip_runtime.h>

#include "sx_internal.hpp"

namespace sx {

constexpr uint64_t kNoKey = 0x7FF0000000000000ull;

// Written as asm: the compiler's double min/max first quiets both operands
// (one more v_max_f64 each) unless it can prove they are not signalling NaNs,
// which it cannot for list entries carried across loop iterations.
__device__ __forceinline__ uint64_t key_min(uint64_t a, uint64_t b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(__builtin_bit_cast(double, a)), "v"(__builtin_bit_cast(double, b)));
    return __builtin_bit_cast(uint64_t, r);
}
__device__ __forceinline__ uint64_t key_max(uint64_t a, uint64_t b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(__builtin_bit_cast(double, a)), "v"(__builtin_bit_cast(double, b)));
    return __builtin_bit_cast(uint64_t, r);
}

template <int K>
struct TopK {
    uint64_t e[K];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int s = 0; s < K; ++s) e[s] = kNoKey;
    }
    __device__ __forceinline__ uint32_t kth_bits() const { return static_cast<uint32_t>(e[K - 1] >> 32); }
    // Slot s of the sorted list with key added is max(e[s-1], min(key, e[s]))
    // (e[s-1] where the key ranks below s, else the smaller of key and e[s]);
    // a key not below e[K-1] leaves the list as it is.
    __device__ __forceinline__ void insert(uint64_t key) {
#pragma unroll
        for (int s = K - 1; s > 0; --s) e[s] = key_max(e[s - 1], key_min(key, e[s]));
        e[0] = key_min(key, e[0]);
    }
    // Shift slots [lo, hi): the insert's formula per slot, highest first (a
    // slot reads its lower neighbour before that neighbour is shifted).
    template <int lo, int hi>
    __device__ __forceinline__ void shift_slots(uint64_t key) {
#pragma unroll
        for (int s = hi - 1; s >= lo; --s) e[s] = s > 0 ? key_max(e[s > 0 ? s - 1 : 0], key_min(key, e[s])) : key_min(key, e[s]);
    }
    // insert() for the active lanes of a wave, one quarter of the list at a
    // time from the top: a lower quarter is shifted only when some active
    // lane's key ranks in or below it (uniform branches), so a wave whose
    // inserted keys all rank near the end of their lists shifts a quarter.
    __device__ __forceinline__ void insert_wave(uint64_t key) {
        constexpr int Q = K / 4;
        const bool q2 = __ballot(key < e[3 * Q - 1]) != 0;
        shift_slots<3 * Q, K>(key);
        if (q2) {
            const bool q1 = __ballot(key < e[2 * Q - 1]) != 0;
            shift_slots<2 * Q, 3 * Q>(key);
            if (q1) {
                const bool q0 = __ballot(key < e[Q - 1]) != 0;
                shift_slots<Q, 2 * Q>(key);
                if (q0) shift_slots<0, Q>(key);
            }
        }
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

// Row of K neighbour indices relative to the cloud's first point `base`;
// missing entries (fewer than K points) repeat the nearest.
template <int K>
__device__ __forceinline__ void write_row(int32_t* row, const TopK<K>& top, int base) {
    const int32_t first = top.e[0] == kNoKey ? -1 : static_cast<int32_t>(top.e[0] & 0xFFFFFFFFull) - base;
#pragma unroll
    for (int s = 0; s < K; s += 4) {
        int4 v;
        v.x = top.e[s] == kNoKey ? first : static_cast<int32_t>(top.e[s] & 0xFFFFFFFFull) - base;
        v.y = top.e[s + 1] == kNoKey ? first : static_cast<int32_t>(top.e[s + 1] & 0xFFFFFFFFull) - base;
        v.z = top.e[s + 2] == kNoKey ? first : static_cast<int32_t>(top.e[s + 2] & 0xFFFFFFFFull) - base;
        v.w = top.e[s + 3] == kNoKey ? first : static_cast<int32_t>(top.e[s + 3] & 0xFFFFFFFFull) - base;
        *reinterpret_cast<int4*>(row + s) = v;
    }
}

// Row of a query position at or beyond the unique count: every entry -1.
template <int K>
__device__ __forceinline__ void write_missing_row(int32_t* row) {
#pragma unroll
    for (int s = 0; s < K; s += 4) *reinterpret_cast<int4*>(row + s) = make_int4(-1, -1, -1, -1);
}

// One query's exact search.  State machine with one record load per step:
//   down at `node`: pick the nearer child; a bucket child is scanned (next
//     step) and counts as returned-from; an internal child is entered.
//   up at `p` coming back from child `from`: if `from` was the nearer child,
//     the farther one is visited (scanned or entered) when not pruned;
//     otherwise p is finished and the walk returns to p's parent.
// `from` is compared with the split gamma of p: the left child's index is
// gamma and the right child's gamma+1, for leaves and internal nodes alike.
// The cloud spans points (and radix nodes) [base, base + U); its root is node base.
template <int K>
__device__ void knn_search(const float4* __restrict__ pts, const TravNode* __restrict__ nodes, ParentRef par, int base, int U,
                           float qx, float qy, float qz, uint32_t r2bits, int bucket, TopK<K>& top) {
    bool scan = U <= bucket;
    bool done = scan;
    int slo = base, shi = base + U - 1;
    int node = base, p = -1, from = -1;
    bool down = true;
    while (true) {
        if (scan) {
            scan_range<K>(pts, slo, shi, qx, qy, qz, top);
            scan = false;
            if (done) break;
        }
        const uint32_t kb = top.kth_bits();
        const uint32_t bound = kb < r2bits ? kb : r2bits;
        const int cur = down ? node : p;
        const TravNode n = nodes[cur];
        const int parent = par.at(cur);
        const int a = w_int(n.lo_l.w), g = w_int(n.hi_l.w), b = w_int(n.lo_r.w);
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
                                                        const TravNode* __restrict__ nodes, ParentRef par,
                                                        const int32_t* __restrict__ d_counts, int64_t q_start,
                                                        int64_t n_q, const float* __restrict__ queries,
                                                        const float* __restrict__ radius2, int bucket,
                                                        int32_t* __restrict__ out, int32_t* __restrict__ reruns,
                                                        CloudMap cm) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_q) return;
    const int total = d_counts[0];
    if (total <= 0) return;
    int base = 0, U = total;
    float qx, qy, qz;
    if (kExternal) {
        qx = queries[3 * t];
        qy = queries[3 * t + 1];
        qz = queries[3 * t + 2];
    } else {
        const int64_t qi = q_start + t;
        if (qi >= total) {  // past the unique count (range checked against the input count)
            write_missing_row<K>(out + t * K);
            return;
        }
        const int2 cloud = cm.at(qi, total);
        base = cloud.x;
        U = cloud.y;
        if (qi - base >= U) return;  // padding slot of a batched cloud
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
        knn_search<K>(pts, nodes, par, base, U, qx, qy, qz, r2bits, bucket, top);
        if (r2bits == 0xFFFFFFFFu || top.kth_bits() <= r2bits) break;
        if (reruns) atomicAdd(reruns, 1);
        top.init();
        r2bits = 0xFFFFFFFFu;
    }
    write_row<K>(out + t * K, top, base);
}

template <bool kExternal>
int knn_dispatch(const float4* pts, const TravNode* nodes, ParentRef par, const int32_t* d_counts, int64_t q_start, int64_t n_q,
                 const float* queries, int k, const float* radius2, int bucket, int32_t* out, int32_t* reruns,
                 hipStream_t st, CloudMap cm) {
    const unsigned g = grid_for(n_q, kKnnBlock);
    switch (k) {
        case 8: knn_kernel<8, kExternal><<<g, kKnnBlock, 0, st>>>(pts, nodes, par, d_counts, q_start, n_q, queries, radius2, bucket, out, reruns, cm); break;
        case 16: knn_kernel<16, kExternal><<<g, kKnnBlock, 0, st>>>(pts, nodes, par, d_counts, q_start, n_q, queries, radius2, bucket, out, reruns, cm); break;
        case 24: knn_kernel<24, kExternal><<<g, kKnnBlock, 0, st>>>(pts, nodes, par, d_counts, q_start, n_q, queries, radius2, bucket, out, reruns, cm); break;
        case 32: knn_kernel<32, kExternal><<<g, kKnnBlock, 0, st>>>(pts, nodes, par, d_counts, q_start, n_q, queries, radius2, bucket, out, reruns, cm); break;
        default: return fail(SX_EINVAL, "knn: k must be 8, 16, 24 or 32");
    }
    SX_LAUNCH_CHECK("knn_kernel");
    return SX_OK;
}

// queries == nullptr: the index points [q_start, q_start + n_q) are the
// queries; otherwise n_q external points (12-byte AoS).
int knn_launch(const float4* pts, const TravNode* nodes, ParentRef par, const int32_t* d_counts, int64_t q_start,
               int64_t n_q, const float* queries, int k, const float* radius2, int bucket, int32_t* out, int32_t* reruns,
               hipStream_t st, CloudMap cm = CloudMap()) {
    if (n_q <= 0) return SX_OK;
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0) return fail(SX_EINVAL, "knn: output must be 16-byte aligned");
    if (queries)
        return knn_dispatch<true>(pts, nodes, par, d_counts, q_start, n_q, queries, k, radius2, bucket, out, reruns, st,
                                  CloudMap());
    return knn_dispatch<false>(pts, nodes, par, d_counts, q_start, n_q, nullptr, k, radius2, bucket, out, reruns, st,
                               cm);
}


// ---------------------------------------------------------------------------
// Wave-cooperative tiles.  A wave owns 64 consecutive queries (Morton
// neighbours, so spatially compact).  Each lane's list is seeded with the
// index points within +-kWindow<K> positions of the tile, and its squared
// radius r2 is:
//   plain:   the k-th best squared distance in that window (a true upper
//            bound of its k-th neighbour distance);
//   learned: the same, capped by kLearnedCap x the predicted r2 where that is
//            smaller - the window bound is loose exactly where the Morton
//            curve jumps, and a capped lane must prove its list at the end.
// Lanes whose radius is at most twice the tile's mean join the tile search:
// the wave walks the tree once, breadth first, for the union box E of their
// balls, collecting the leaf ranges of every subtree whose box meets E.  E is
// widened with directed rounding so that every point with dist2 <= r2 of a
// joining lane lies inside it (same monotonicity argument as box_dist2).
// Candidate points are then loaded 64 at a time, one per lane (coalesced);
// those inside E are broadcast lane by lane while every lane updates its own
// list.  A lane's list is exact when (plain) it joined the search, or
// (learned) it also found at least K points within its radius.  Other lanes
// - and whole tiles whose frontier or range list overflows LDS - are appended
// to a queue that knn_group_kernel searches, one 16-lane group per query.
// ---------------------------------------------------------------------------
constexpr int kTileWaves = kKnnBlock / kWave;
constexpr int kTileStack = 256;                          // subtrees waiting to be tested against E
constexpr int kRangeFlush = 192;                         // stream the ranges once this many are pending
constexpr int kTileRanges = kRangeFlush - 1 + 2 * kWave;  // one expansion adds at most 2 per lane
// Plain window: the tile's points +-kWindow<K> in Morton order.  Measured:
// K = 16 is flat from 16 to 64 (8M scene); K = 32 gains from 96 (16M
// mixture: 24.7 -> 22.2 ms for the kNN) since 128 points hold too few
// 32-neighbourhoods.
template <int K>
constexpr int kWindow = K >= 32 ? 96 : 32;
// Points of the staged window (the tile and kWindow points either side) and
// its LDS slots (a power of two: the slot number fills the low bits of a
// coarse key).
template <int K>
constexpr int kWinPoints = kWave + 2 * kWindow<K>;
constexpr int pow2_ceil(int v) { return v <= 1 ? 1 : 2 * pow2_ceil((v + 1) / 2); }
template <int K>
constexpr int kWinSlots = pow2_ceil(kWinPoints<K>);
// Candidates one segment may stream before its lanes are deferred.  The
// check runs after each flush, so a segment that overruns has already
// streamed about this many candidates to every lane; the deferred-query
// kernel searches such queries for less.  Measured on the 8M scene, k = 16:
// plain kNN 4.84 / 4.65 / 4.38 / 4.17 / 4.25 / 4.58 ms at 8192 / 6144 / 4096
// / 2048 / 1536 / 1024 (flat from 1792 to 2048).
template <int K>
constexpr int kSegBudget = 2048 * K / 16;
constexpr float kSegGap = 5.0f;                          // segment break: gap > kSegGap x mean radius (5 / 6 / 8 / 12: scene8m kNN 2.58 / 2.59 / 2.66 / 2.74 ms)
constexpr float kRadiusStretch = 2.25f;                  // learned: accept up to 1.5x the predicted radius
constexpr float kLearnedCap = 1.21f;                     // learned: radius capped at 1.1x the predicted one

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
// An empty asm that "modifies" v: values derived from the result are not
// hoisted out of the enclosing loop (keeps per-iteration temporaries out of
// the registers that stay live across the loop).
__device__ __forceinline__ float opaque(float v) {
    asm volatile("" : "+v"(v));
    return v;
}
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}
__device__ __forceinline__ float wave_min_at(float v, int lane) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
        v = fminf(v, __int_as_float(__builtin_amdgcn_ds_bpermute((lane ^ o) << 2, __float_as_int(v))));
    return v;
}
__device__ __forceinline__ float wave_max_at(float v, int lane) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
        v = fmaxf(v, __int_as_float(__builtin_amdgcn_ds_bpermute((lane ^ o) << 2, __float_as_int(v))));
    return v;
}
__device__ __forceinline__ bool box_meets(float4 lo, float4 hi, float ex0, float ey0, float ez0, float ex1, float ey1,
                                          float ez1) {
    return hi.x >= ex0 && lo.x <= ex1 && hi.y >= ey0 && lo.y <= ey1 && hi.z >= ez0 && lo.z <= ez1;
}
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uniform_f(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ __forceinline__ float read_lane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ uint64_t read_lane64(uint64_t v, int src) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), src));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), src));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(v), src, kWave));
    const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(v >> 32), src, kWave));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl64_xor(uint64_t v, int mask) {
    const uint32_t lo = static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), mask, kWave));
    const uint32_t hi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(v >> 32), mask, kWave));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
// Ascending bitonic sort of one 64-bit key per lane across the wave.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t v) {
    const int lane = static_cast<int>(lane_id());
#pragma unroll
    for (int k = 2; k <= kWave; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t o = shfl64_xor(v, j);
            const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
            v = keep_min ? (v < o ? v : o) : (v < o ? o : v);
        }
    }
    return v;
}
// Number of set bits of m below this lane (popcount(m & lanes_below)).
__device__ __forceinline__ int rank_below(uint64_t m) {
    return static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                      __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
}
__device__ __forceinline__ uint64_t make_key(float d, int j) {
    return (static_cast<uint64_t>(__float_as_uint(d)) << 32) | static_cast<uint32_t>(j);
}

// ---------------------------------------------------------------------------
// Lane-parallel selection by sorting networks.  A wave-wide broadcast insert
// costs every lane a full K-slot shift whenever ANY lane improves, which in a
// dense window of candidates is nearly every step.  Instead each lane sorts a
// chunk of C keys with a fixed network and merges it into its list: the cost
// is the same for every lane and does not depend on who accepts.
// ---------------------------------------------------------------------------
// Batcher's odd-even merge sort on N (a power of two) elements, as a list of
// compare-exchange pairs (a < b): 19 for N = 8, 63 for N = 16.
template <int N>
struct OddEvenNet {
    static constexpr int kMax = 8 * N;
    int n = 0;
    unsigned char a[kMax] = {}, b[kMax] = {};
    constexpr OddEvenNet() {
        for (int p = 1; p < N; p <<= 1)
            for (int k = p; k >= 1; k >>= 1)
                for (int j = k % p; j + k < N; j += 2 * k)
                    for (int i = 0; i < k && i < N - j - k; ++i)
                        if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
                            a[n] = static_cast<unsigned char>(i + j);
                            b[n] = static_cast<unsigned char>(i + j + k);
                            ++n;
                        }
    }
};

__device__ __forceinline__ void cswap(uint64_t& x, uint64_t& y) {
    const uint64_t lo = key_min(x, y);
    y = key_max(x, y);
    x = lo;
}

template <int N>
__device__ __forceinline__ void sort_keys(uint64_t (&x)[N]) {
    constexpr OddEvenNet<N> net{};
#pragma unroll
    for (int c = 0; c < net.n; ++c) cswap(x[net.a[c]], x[net.b[c]]);
}

// Merge an ascending chunk c[0, C) into the ascending list (C <= K, K a power
// of two): e[K-1-t] = min(e[K-1-t], c[t]) leaves the K smallest of both as a
// bitonic sequence, which half-cleaners then sort.
template <int K, int C>
__device__ __forceinline__ void merge_chunk(uint64_t (&e)[K], const uint64_t (&c)[C]) {
    static_assert(C <= K && (K & (K - 1)) == 0, "merge_chunk: C <= K, K a power of two");
#pragma unroll
    for (int t = 0; t < C; ++t) e[K - 1 - t] = c[t] < e[K - 1 - t] ? c[t] : e[K - 1 - t];
#pragma unroll
    for (int d = K / 2; d >= 1; d >>= 1)
#pragma unroll
        for (int j = 0; j < K; ++j)
            if ((j & d) == 0) cswap(e[j], e[j + d]);
}

// Flattened position f (< total) of a list of inclusive index ranges with
// inclusive size prefix `pre` -> point index.
__device__ __forceinline__ int range_point(const int2* rng, const int* pre, int nr, int f) {
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] > f) hi = mid; else lo = mid + 1;
    }
    return rng[lo].y - (pre[lo] - 1 - f);
}

// Inclusive wave sum whose lane-permute addresses are rebuilt at every call:
// the compiler would otherwise hoist them out of the kNN's loops and keep six
// registers live through the whole kernel.
__device__ __forceinline__ int wave_incl_sum_fresh(int v) {
    const int lane = opaque(static_cast<int>(lane_id()));
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __builtin_amdgcn_ds_bpermute((lane >= o ? lane - o : lane) << 2, v);
        if (lane >= o) v += t;
    }
    return v;
}

// In place: range i = (first, last) becomes (inclusive size prefix, last),
// so the prefix needs no array of its own; returns the total (uniform).
__device__ __forceinline__ int range_prefix_inplace(int2* rng, int nr) {
    const int lane = static_cast<int>(lane_id());
    int total = 0;
    for (int base = 0; base < nr; base += kWave) {
        const int idx = base + lane;
        const int2 rg = idx < nr ? rng[idx] : make_int2(0, -1);
        const int incl = total + wave_incl_sum_fresh(rg.y - rg.x + 1);
        if (idx < nr) rng[idx].x = incl;
        total = __builtin_amdgcn_readlane(incl, kWave - 1);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    return total;
}

// Flattened position f (< total) of ranges prepared by range_prefix_inplace
// -> point index.
__device__ __forceinline__ int range_point_inplace(const int2* rng, int nr, int f) {
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rng[mid].x > f) hi = mid; else lo = mid + 1;
    }
    const int2 r = rng[lo];
    return r.y - (r.x - 1 - f);
}

// Broadcast the n candidates staged in cand[0, n) (x, y, z, index bits; two
// per step, uniform LDS reads) to every lane; a lane inserts a candidate when
// on, it beats the lane's k-th key and (kTau) its squared distance is <= tau.
// Without kTau the distance test is implied: a lane's radius is its k-th
// distance when the stream starts, and the list only improves.
// cand[n] must hold a point at +inf when n is odd.
template <int K, bool kTau>
__device__ __forceinline__ void broadcast_insert(const float4* cand, int n, float4 q, bool on, uint32_t tau,
                                                 TopK<K>& top) {
    for (int i = 0; i < n; i += 2) {
        const float4 c0 = cand[i];
        const float4 c1 = cand[i + 1];
        const uint32_t d0 = __float_as_uint(dist2(q.x, q.y, q.z, c0.x, c0.y, c0.z));
        const uint32_t d1 = __float_as_uint(dist2(q.x, q.y, q.z, c1.x, c1.y, c1.z));
        const int j0 = __float_as_int(c0.w), j1 = __float_as_int(c1.w);
        const uint64_t k0 = (static_cast<uint64_t>(d0) << 32) | static_cast<uint32_t>(j0);
        const uint64_t k1 = (static_cast<uint64_t>(d1) << 32) | static_cast<uint32_t>(j1);
        // the distance alone decides whether to insert: a key that ties the
        // k-th distance with a larger index leaves the list as it is
        if (on && (!kTau || d0 <= tau) && d0 <= top.kth_bits()) top.insert_wave(k0);
        if (on && (!kTau || d1 <= tau) && d1 <= top.kth_bits()) top.insert_wave(k1);
    }
}

// Lists of a power-of-two length take the window through sorting networks.
template <int K>
constexpr bool kNetWindow = (K & (K - 1)) == 0;

// Merge the staged points cand[0, n) (n <= kWave; slots up to the next chunk
// boundary hold points at +inf) into every lane's list, C keys at a time.
template <int K>
__device__ __forceinline__ void window_merge(const float4* cand, int n, float4 q, TopK<K>& top) {
    constexpr int C = K <= 16 ? 8 : 16;  // K = 16: an 8-key chunk keeps the list and chunk within 80 VGPRs
    static_assert(kWave % C == 0, "window_merge: chunks tile the staged points");
#pragma unroll 1
    for (int c0 = 0; c0 < n; c0 += C) {
        uint64_t x[C];
#pragma unroll
        for (int t = 0; t < C; ++t) {
            const float4 p = cand[c0 + t];
            x[t] = make_key(dist2(q.x, q.y, q.z, p.x, p.y, p.z), __float_as_int(p.w));
        }
        sort_keys<C>(x);
        merge_chunk<K, C>(top.e, x);
    }
}

// ---------------------------------------------------------------------------
// Window selection on 32-bit coarse keys.  The window is at most 256 points
// staged in LDS; point t of it gets the key (distance bits with the low PB
// bits cleared) | t, PB = log2(window slots).  A coarse key orders by the
// distance rounded down to 2^PB ulps and then by window position, i.e. by
// index, so a compare-exchange is one min and one max (a 64-bit key needs a
// compare and four selects).  The K smallest coarse keys L then give the
// exact K smallest (distance, index) keys whenever no discarded point shares
// the coarse distance of L's last entry: every point of a smaller coarse
// distance has a smaller exact distance, and there are fewer than K of them.
// The exact keys of L are rebuilt from the staged points and sorted when two
// entries with one coarse distance came out of order; an ambiguous boundary
// sends the wave to the exact 64-bit merge.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cswap32(uint32_t& x, uint32_t& y) {
    const uint32_t a = x, b = y;
    x = a < b ? a : b;
    y = a < b ? b : a;
}

template <int N>
__device__ __forceinline__ void sort_keys32(uint32_t (&x)[N]) {
    constexpr OddEvenNet<N> net{};
#pragma unroll
    for (int c = 0; c < net.n; ++c) cswap32(x[net.a[c]], x[net.b[c]]);
}

// Coarse keys of the K nearest of the kSlots staged points win[0, kSlots)
// (points past the window at +inf) into e (ascending); *dmin receives the
// smallest key the selection discarded.
template <int K, int kSlots, int kPoints>
__device__ __forceinline__ void coarse_window(const float4* win, float4 q, uint32_t (&e)[K], uint32_t& dmin) {
    constexpr int C = K <= 8 ? 8 : 16;
    constexpr uint32_t kPosMask = kSlots - 1;
    static_assert((kSlots & (kSlots - 1)) == 0 && kPoints % C == 0 && kPoints <= kSlots,
                  "coarse_window: points in whole chunks within the slots");
#pragma unroll
    for (int s = 0; s < K; ++s) e[s] = 0xFFFFFFFFu;
    dmin = 0xFFFFFFFFu;
#pragma unroll 1
    for (int c0 = 0; c0 < kPoints; c0 += C) {
        uint32_t x[C];
#pragma unroll
        for (int t = 0; t < C; ++t) {
            const float4 p = win[c0 + t];
            x[t] = (__float_as_uint(dist2(q.x, q.y, q.z, p.x, p.y, p.z)) & ~kPosMask) | static_cast<uint32_t>(c0 + t);
        }
        sort_keys32<C>(x);
        // e[K-1-t] = min(e[K-1-t], x[t]) keeps the K smallest of both as a
        // bitonic sequence; the maxima are the discarded keys
#pragma unroll
        for (int t = 0; t < C; ++t) {
            const uint32_t a = e[K - 1 - t], b = x[t];
            e[K - 1 - t] = a < b ? a : b;
            const uint32_t hi = a < b ? b : a;
            dmin = hi < dmin ? hi : dmin;
        }
#pragma unroll
        for (int d = K / 2; d >= 1; d >>= 1)
#pragma unroll
            for (int j = 0; j < K; ++j)
                if ((j & d) == 0) cswap32(e[j], e[j + d]);
    }
}

// Exact (distance, index) keys of the coarse selection, in ascending order
// (a slot past the window keeps its +inf distance and index, as in the exact
// merge).
// Returns false when the selection may differ from the exact one (a
// discarded point shares the coarse distance of the last entry).
template <int K, int kSlots>
__device__ __forceinline__ bool exact_window(const float4* win, int wlo, float4 q, const uint32_t (&e)[K], uint32_t dmin,
                                             TopK<K>& top) {
    constexpr uint32_t kPosMask = kSlots - 1;
    bool unsorted = false;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int t = static_cast<int>(e[s] & kPosMask);  // every entry is a slot: kSlots > K
        const float4 p = win[t];
        const uint64_t key = make_key(dist2(q.x, q.y, q.z, p.x, p.y, p.z), wlo + t);
        top.e[s] = key;
        if (s > 0) unsorted = unsorted || top.e[s - 1] > key;
    }
    if (__ballot(unsorted)) sort_keys<K>(top.e);  // equal coarse distances out of exact order (rare)
    return e[K - 1] >= 0x7F800000u || (dmin & ~kPosMask) != (e[K - 1] & ~kPosMask);
}

// Stage the lanes' candidates (p, j) with ok set into cand[0, n), in lane
// order, followed by one point at +inf; returns n (uniform).
template <bool kFresh = false>
__device__ __forceinline__ int stage_candidates(float4* cand, float4 p, int j, bool ok) {
    const uint64_t m = __ballot(ok);
    const int n = __popcll(m);
    if (ok) cand[rank_below(m)] = make_float4(p.x, p.y, p.z, __int_as_float(j));
    if (static_cast<int>(lane_id()) == 0) {
        // kFresh: made here (not a loop-invariant vector the compiler would
        // keep live, and spill, across the caller's loops)
        const float inf = kFresh ? opaque(INFINITY) : INFINITY;
        cand[n] = make_float4(inf, inf, inf, 0.0f);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    return n;
}

// Stream the leaf ranges rng[0, nr): 64 candidates per load (one per lane,
// coalesced); those inside E are staged and broadcast.  Indices in
// [skip_lo, skip_hi] (the window) were seen by every lane already and are
// not staged.
// The ranges come prepared by range_prefix_inplace (total points: total).
template <int K, bool kTau>
__device__ __forceinline__ void stream_ranges(const float4* __restrict__ pts, const int2* rng, int nr, int total,
                                              int skip_lo, int skip_hi, float ex0, float ey0, float ez0, float ex1,
                                              float ey1, float ez1, float4 q, bool on, uint32_t tau, TopK<K>& top,
                                              int& streamed, float4* cand) {
    const int lane = static_cast<int>(lane_id());
    for (int f0 = 0; f0 < total; f0 += kWave) {
        const int f = f0 + lane;
        int j = f < total ? range_point_inplace(rng, nr, f) : -1;
        if (j >= skip_lo && j <= skip_hi) j = -1;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j >= 0) p = pts[j];
        const bool in_e = j >= 0 && p.x >= ex0 && p.x <= ex1 && p.y >= ey0 && p.y <= ey1 && p.z >= ez0 && p.z <= ez1;
        const int n = stage_candidates<K <= 16>(cand, p, j, in_e);
        streamed += n;
        broadcast_insert<K, kTau>(cand, n, q, on, tau, top);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// Waves per SIMD, measured on the 8M scene (K = 16): plain eight (64 VGPRs,
// 8 bytes per lane spilled; kNN 2.72 ms against 2.76 at seven with 72 VGPRs
// and 2.89 at six with 77), seeded eight too (64 VGPRs, no spill: 2.69
// against 2.73 ms at seven).  K = 24: seven (4M degenerate scene: kNN 2.41
// ms against 2.46 at six, 6.3 at eight); K = 32: five (16M mixture: 11.0 ms
// against 12.16 at its natural ~120 VGPRs, four waves, and 21.7 at six).
template <int K, bool kLearned>
__global__ void __launch_bounds__(kKnnBlock) __attribute__((amdgpu_waves_per_eu(K <= 16 ? 8 : (K <= 24 ? 7 : 5), 8)))
knn_tile_kernel(const float4* __restrict__ pts,
                                                             const TravNode* __restrict__ nodes,
                                                             const int32_t* __restrict__ d_counts, int64_t q_start,
                                                             int64_t n_q, const float* __restrict__ radius2,
                                                             int bucket, int32_t* __restrict__ out,
                                                             int2* __restrict__ queue, int32_t* queue_count,
                                                             int32_t* cand_total, CloudMap cm) {
    // per wave: the window's points, then (same memory) the range query's
    // stack, leaf ranges and their size prefix
    union TileScratch {
        struct {
            int stack[kTileStack];
            int2 range[kTileRanges];  // (first, last), then (size prefix, last)
        } q;
        float4 win[kWinSlots<K>];
    };
    __shared__ TileScratch s_scr[kTileWaves];
    __shared__ float4 s_cand[kTileWaves][kWave + 1];
    // Wave-uniform values, the query index and the query point are kept out
    // of the long-lived vector registers (scalar registers, or recomputed /
    // reloaded where used): every list length runs at an occupancy that
    // would otherwise spill them (K = 24 and 32: 2.61 -> 2.41 ms and 12.07 ->
    // 11.02 ms with the occupancies above).
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
    const int lane = static_cast<int>(lane_id());
    const int total = d_counts[0];
    const int64_t q_req = q_start + n_q;
    int64_t q_end = q_req;
    if (q_end > total) q_end = total;
    const int64_t tile0 = q_start + (static_cast<int64_t>(xcd_block()) * kTileWaves + w) * kWave;
    // positions asked for at or beyond the unique count get rows of -1 (the
    // range is checked against the input count until the counts are read)
    if (tile0 + lane >= total && tile0 + lane < q_req) write_missing_row<K>(out + (tile0 + lane - q_start) * K);
    if (tile0 >= q_end) return;  // whole wave idle (no block barriers below)
    // the tile's cloud: points and radix nodes [cbase, cbase + U)
    const int2 cloud = cm.at(tile0, total);
    const int cbase = __builtin_amdgcn_readfirstlane(cloud.x);
    const int U = __builtin_amdgcn_readfirstlane(cloud.y);
    const int cend = cbase + U;
    const int64_t qi = tile0 + lane;
    const bool valid = qi < q_end && qi < cend;
    const float4 q = valid ? pts[qi] : make_float4(0.f, 0.f, 0.f, 0.f);
    int* stk = s_scr[w].q.stack;
    int2* rng = s_scr[w].q.range;
    TopK<K> top;
    top.init();

    // ---- per-lane radius: the window
    int wlo, whi;
    {
        constexpr int hw = kWindow<K>;
        wlo = static_cast<int>(tile0 - hw > cbase ? tile0 - hw : cbase);
        whi = static_cast<int>(tile0 + kWave - 1 + hw < cend - 1 ? tile0 + kWave - 1 + hw : cend - 1);
        // wave-uniform: kept in scalar registers
        wlo = __builtin_amdgcn_readfirstlane(wlo);
        whi = __builtin_amdgcn_readfirstlane(whi);
        if constexpr (kNetWindow<K>) {
            // the whole window in LDS; slots past it hold points at +inf,
            // whose keys follow every finite one, so they reach the list only
            // when the window holds fewer than K points, and such a lane
            // never joins (its k-th distance is not finite)
            constexpr int kSlots = kWinSlots<K>;
            constexpr int kPoints = kWinPoints<K>;
            float4* win = s_scr[w].win;
#pragma unroll
            for (int t0 = 0; t0 < kSlots; t0 += kWave) {
                const int t = t0 + lane;
                win[t] = wlo + t <= whi ? pts[wlo + t] : make_float4(INFINITY, INFINITY, INFINITY, 0.f);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            bool exact_sel;
            {
                uint32_t ce[K];
                uint32_t dmin;
                coarse_window<K, kSlots, kPoints>(win, q, ce, dmin);
                exact_sel = exact_window<K, kSlots>(win, wlo, q, ce, dmin, top);
            }
            if (__ballot(!exact_sel)) {
                // a boundary the coarse keys cannot settle: the exact merge,
                // 64 staged points at a time
                top.init();
                for (int base = 0; base < kSlots; base += kWave) {
                    const float4 p = win[base + lane];
                    s_cand[w][lane] = make_float4(p.x, p.y, p.z, __int_as_float(wlo + base + lane));
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    window_merge<K>(s_cand[w], kWave, q, top);
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        } else {
            for (int base = wlo; base <= whi; base += kWave) {
                const int j = base + lane;
                const float4 p = j <= whi ? pts[j] : make_float4(0.f, 0.f, 0.f, 0.f);
                const int n = stage_candidates(s_cand[w], p, j, j <= whi);
                broadcast_insert<K, false>(s_cand[w], n, q, true, 0xFFFFFFFFu, top);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
    }
    uint32_t r2bits = top.kth_bits();  // >= 0x7F800000 when the window holds fewer than K points
    bool capped = false;
    if (kLearned && valid) {
        const float pr2 = radius2[qi - q_start];
        if (pr2 >= 0.0f && pr2 < INFINITY) {
            const uint32_t c = __float_as_uint(pr2 * kLearnedCap);
            if (c < r2bits) {
                r2bits = c;
                capped = true;
            }
        }
    }
    const bool finite = valid && r2bits < 0x7F800000u;
    const float r = finite ? sqrtf(__uint_as_float(r2bits)) : 0.0f;
    const float n_fin = wave_sum(finite ? 1.0f : 0.0f);
    const float mean_r = n_fin > 0.0f ? wave_sum(r) / n_fin : 0.0f;
    const bool join = finite && r <= 2.0f * mean_r;
    // padded radius: fl(rp * rp) > r2; each box bound is stepped one ulp
    // outward so it bounds the exact q -+ rp
    const float rp = r * 1.001f + 1e-19f;

    // ---- segments: runs of joining lanes without a spatial gap larger than
    // kSegGap mean radii between consecutive ones, so that a tile straddling
    // a jump of the Morton curve does not search the space in between
    uint64_t todo = __ballot(join);
    uint64_t breaks;
    {
        const uint64_t below = todo & ((1ull << lane) - 1ull);
        const int prev = below ? 63 - static_cast<int>(__builtin_clzll(below)) : lane;
        const float gx = __shfl(q.x, prev, kWave), gy = __shfl(q.y, prev, kWave), gz = __shfl(q.z, prev, kWave);
        const float gap = kSegGap * mean_r;
        breaks = __ballot(join && below != 0ull &&
                          (fabsf(q.x - gx) > gap || fabsf(q.y - gy) > gap || fabsf(q.z - gz) > gap));
    }
    bool lost = false;
    int streamed = 0;   // candidates broadcast to the wave (reported when cand_total is given)
    uint32_t lim = 0u;  // squared radius within which the lane's list is proven complete
    while (todo) {
        const int s0 = static_cast<int>(__builtin_ctzll(todo));
        const uint64_t after = s0 < 63 ? breaks & (~0ull << (s0 + 1)) : 0ull;
        const uint64_t upto = after ? (1ull << __builtin_ctzll(after)) - 1ull : ~0ull;
        const uint64_t seg = todo & upto;
        todo &= ~seg;
        const bool in_seg = (seg >> lane) & 1ull;
        const float rps = opaque(rp);  // recomputed per segment rather than held live
        const int ls = opaque(lane);
        // the query point loaded again (a cached read)
        const float4 qs = valid ? pts[tile0 + ls] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float ex0 = uniform_f(wave_min_at(in_seg ? nextafterf(qs.x - rps, -INFINITY) : INFINITY, ls));
        const float ey0 = uniform_f(wave_min_at(in_seg ? nextafterf(qs.y - rps, -INFINITY) : INFINITY, ls));
        const float ez0 = uniform_f(wave_min_at(in_seg ? nextafterf(qs.z - rps, -INFINITY) : INFINITY, ls));
        const float ex1 = uniform_f(wave_max_at(in_seg ? nextafterf(qs.x + rps, INFINITY) : -INFINITY, ls));
        const float ey1 = uniform_f(wave_max_at(in_seg ? nextafterf(qs.y + rps, INFINITY) : -INFINITY, ls));
        const float ez1 = uniform_f(wave_max_at(in_seg ? nextafterf(qs.z + rps, INFINITY) : -INFINITY, ls));
        uint32_t tau = 0u;
        if (in_seg) {
            tau = r2bits;
            if (kLearned && capped) {
                // Every point with dist2 <= rho2 lies inside E: the gap to the
                // nearest face is shrunk by 2^-20 twice, far more than the
                // rounding of the gap, the square and dist2 itself.  A lane may
                // so accept beyond its predicted radius (up to kRadiusStretch).
                float g = fminf(fminf(qs.x - ex0, ex1 - qs.x),
                                fminf(fminf(qs.y - ey0, ey1 - qs.y), fminf(qs.z - ez0, ez1 - qs.z)));
                g = fmaxf(g, 0.0f) * (1.0f - 0x1p-20f);
                const float rho2 = (g * g) * (1.0f - 0x1p-20f);
                const uint32_t t = __float_as_uint(fminf(rho2, __uint_as_float(r2bits) * kRadiusStretch));
                if (t > tau) tau = t;  // the capped ball itself lies inside E by construction
            }
            lim = tau;
        }
        // ---- depth-first range query for E, 64 subtrees per step; the
        // ranges are streamed whenever kRangeFlush of them are pending, each
        // batch only after its size is checked against the segment's budget
        int sp = 0, nr = 0, seen = 0;
        bool overflow = false;
        if (U <= bucket) {
            if (lane == 0) rng[0] = make_int2(cbase, cend - 1);
            nr = 1;
        } else {
            if (lane == 0) stk[0] = cbase;  // the cloud's root
            sp = 1;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        while (true) {
            while (sp > 0 && nr < kRangeFlush) {
                const int c = sp < kWave ? sp : kWave;
                sp -= c;
                const int node = lane < c ? stk[sp + lane] : -1;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                bool nodeL = false, nodeR = false, rangeL = false, rangeR = false;
                int a = 0, g = 0, b = 0;
                if (node >= 0) {
                    const TravNode n = nodes[node];
                    a = w_int(n.lo_l.w);
                    g = w_int(n.hi_l.w);
                    b = w_int(n.lo_r.w);
                    if (box_meets(n.lo_l, n.hi_l, ex0, ey0, ez0, ex1, ey1, ez1)) {
                        if (g - a < bucket) rangeL = true; else nodeL = true;
                    }
                    if (box_meets(n.lo_r, n.hi_r, ex0, ey0, ez0, ex1, ey1, ez1)) {
                        if (b - g - 1 < bucket) rangeR = true; else nodeR = true;
                    }
                }
                const uint64_t mnl = __ballot(nodeL), mnr = __ballot(nodeR);
                const uint64_t mrl = __ballot(rangeL), mrr = __ballot(rangeR);
                const int pushes = __popcll(mnl) + __popcll(mnr);
                if (sp + pushes > kTileStack) {
                    overflow = true;
                    sp = 0;
                    break;
                }
                const int pnl = sp + rank_below(mnl);
                const int pnr = sp + __popcll(mnl) + rank_below(mnr);
                const int prl = nr + rank_below(mrl);
                const int prr = nr + __popcll(mrl) + rank_below(mrr);
                if (nodeL) stk[pnl] = g;
                if (nodeR) stk[pnr] = g + 1;
                if (rangeL) rng[prl] = make_int2(a, g);
                if (rangeR) rng[prr] = make_int2(g + 1, b);
                sp += pushes;
                nr += __popcll(mrl) + __popcll(mrr);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
            if (overflow || nr == 0) break;
            const int total = range_prefix_inplace(rng, nr);
            if (seen + total > kSegBudget<K>) {
                overflow = true;
                break;
            }
            seen += total;
            stream_ranges<K, kLearned>(pts, rng, nr, total, wlo, whi, ex0, ey0, ez0, ex1, ey1, ez1, qs, in_seg, tau,
                                       top, streamed, s_cand[w]);
            nr = 0;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (sp == 0) break;
        }
        if (overflow) lost = lost || in_seg;  // the segment's lanes are deferred
    }
    // uncapped: every point within the window radius was a candidate;
    // capped: the k-th distance must lie within the radius the box proves
    const bool exact = join && !lost && (capped ? top.kth_bits() <= lim : true);
    const int64_t qo = tile0 + opaque(lane);  // the query index, recomputed
    if (valid && exact) write_row<K>(out + (qo - q_start) * K, top, cbase);
    const bool defer = valid && !exact;
    const uint64_t md = __ballot(defer);
    if (md) {
        int slot0 = 0;
        if (lane == __ffsll(static_cast<unsigned long long>(md)) - 1) slot0 = atomicAdd(queue_count, __popcll(md));
        slot0 = __shfl(slot0, __ffsll(static_cast<unsigned long long>(md)) - 1, kWave);
        // with the k-th key of its list, an upper bound on the k-th distance
        // when the list holds K real points (+inf bits otherwise)
        if (defer) queue[slot0 + rank_below(md)] = make_int2(static_cast<int32_t>(qo), static_cast<int>(top.kth_bits()));
    }
    if (cand_total && lane == 0 && streamed) atomicAdd(cand_total, streamed);
}

// ---------------------------------------------------------------------------
// Queued queries, one 16-lane group each (four per wave).  A query arrives
// with an upper bound on its k-th squared distance: the k-th key of the list
// its tile built from real points (+inf when that list held fewer than K).
// The list is seeded with the query's 64 Morton neighbours (group-sorted),
// and subtrees and candidates beyond the bound are skipped from the start.
// Subtrees are expanded 16 at a time from a per-group LDS stack, pruned by min(bound, current k-th distance) (strict, ties survive),
// and their leaf ranges streamed one point per lane.  The list is spread
// over the group - entry e*16 + l is register e of lane l - and a candidate
// that beats the k-th key is inserted by a ballot for its rank and a shift;
// the four groups of a wave insert their own candidates in the same steps.
// Most deferred queries have frontiers narrower than 16 subtrees, so a
// group does the work a whole wave did, at a quarter of the instructions.
// A query whose stack would overflow is marked in the queue (index
// complemented) and searched by knn_overflow_kernel, one lane per query: the
// per-lane search inlined here would raise the group kernel's register count
// for every query.
// ---------------------------------------------------------------------------
constexpr int kGroup = 16;
constexpr int kGroupsPerWave = kWave / kGroup;
constexpr int kGroupStack = 112;  // 20 KB of LDS per block: 8 blocks (8 waves/SIMD) per CU
constexpr int kGroupRanges = 2 * kGroup;  // one expansion adds at most two per lane

// Group-local collectives; gs = 16 * (group within the wave).
__device__ __forceinline__ uint32_t group_ballot(bool p, int gs) {
    return static_cast<uint32_t>(__ballot(p) >> gs) & 0xFFFFu;
}
__device__ __forceinline__ uint64_t group_shfl64(uint64_t v, int src) {
    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(v), src, kGroup));
    const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(v >> 32), src, kGroup));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t group_shfl_up64(uint64_t v) {
    const uint32_t lo = static_cast<uint32_t>(__shfl_up(static_cast<int>(v), 1, kGroup));
    const uint32_t hi = static_cast<uint32_t>(__shfl_up(static_cast<int>(v >> 32), 1, kGroup));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ int group_incl_sum(int v, int gl) {
#pragma unroll
    for (int o = 1; o < kGroup; o <<= 1) {
        const int t = __shfl_up(v, o, kGroup);
        if (gl >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t group_shfl_xor64(uint64_t v, int mask) {
    const uint32_t lo = static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), mask, kGroup));
    const uint32_t hi = static_cast<uint32_t>(__shfl_xor(static_cast<int>(v >> 32), mask, kGroup));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
// Ascending bitonic sort of 64 keys over a group: element r*16 + l is
// register r of lane l.  Partners 16 or 32 apart are in the same lane.
__device__ __forceinline__ void group_sort64(uint64_t (&x)[4], int gl) {
#pragma unroll
    for (int k = 2; k <= 4 * kGroup; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= kGroup) {
                const int rj = j / kGroup;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (r & rj) continue;
                    const bool asc = ((r * kGroup + gl) & k) == 0;
                    uint64_t& a = x[r];
                    uint64_t& b = x[r | rj];
                    const bool s = asc ? b < a : a < b;
                    const uint64_t lo = s ? b : a;
                    b = s ? a : b;
                    a = lo;
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint64_t o = group_shfl_xor64(x[r], j);
                    const bool asc = ((r * kGroup + gl) & k) == 0;
                    const bool keep_min = ((gl & j) == 0) == asc;
                    x[r] = keep_min ? (x[r] < o ? x[r] : o) : (x[r] < o ? o : x[r]);
                }
            }
        }
    }
}

// Sorted list of K keys spread over a group.
template <int K>
struct GroupList {
    static constexpr int E = (K + kGroup - 1) / kGroup;
    uint64_t key[E];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int e = 0; e < E; ++e) key[e] = ~0ull;
    }
    __device__ __forceinline__ uint64_t kth() const { return group_shfl64(key[(K - 1) / kGroup], (K - 1) % kGroup); }
    // x (uniform over the group) goes to its rank; entries behind it move up
    __device__ __forceinline__ void insert(uint64_t x, int gl, int gs) {
        int pos = 0;
#pragma unroll
        for (int e = 0; e < E; ++e) pos += __popc(group_ballot(key[e] < x, gs));
#pragma unroll
        for (int e = E - 1; e >= 0; --e) {
            const uint64_t up = group_shfl_up64(key[e]);
            const uint64_t carry = e > 0 ? group_shfl64(key[e > 0 ? e - 1 : 0], kGroup - 1) : ~0ull;
            const uint64_t prev = gl == 0 ? carry : up;
            const int idx = e * kGroup + gl;
            key[e] = idx >= K ? ~0ull : (idx == pos ? x : (idx > pos ? prev : key[e]));
        }
    }
};

template <int K>
// Latency-bound (one dependent record load per level): eight waves per SIMD.
__global__ void __launch_bounds__(kKnnBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) knn_group_kernel(const float4* __restrict__ pts,
                                                              const TravNode* __restrict__ nodes,
                                                              const int32_t* __restrict__ d_counts, int64_t q_start,
                                                              int2* __restrict__ queue,
                                                              const int32_t* __restrict__ queue_count, int bucket,
                                                              int32_t* __restrict__ out, int32_t* deferred_total,
                                                              CloudMap cm) {
    constexpr int kGroups = kTileWaves * kGroupsPerWave;
    __shared__ int2 s_stack[kGroups][kGroupStack];  // (node, box distance bits)
    __shared__ int2 s_rng[kGroups][kGroupRanges];
    __shared__ int s_pre[kGroups][kGroupRanges];
    const int lane = static_cast<int>(lane_id());
    const int gl = lane & (kGroup - 1);
    const int gs = lane & ~(kGroup - 1);
    const int gid = static_cast<int>(threadIdx.x) / kGroup;  // group within the block
    int2* stk = s_stack[gid];
    int2* rng = s_rng[gid];
    int* pre = s_pre[gid];
    const int total = d_counts[0];
    const int n = *queue_count;
    if (blockIdx.x == 0 && threadIdx.x == 0 && deferred_total) atomicAdd(deferred_total, n);
    const int n_groups = static_cast<int>(gridDim.x) * kGroups;
    for (int t = static_cast<int>(blockIdx.x) * kGroups + gid; t < n; t += n_groups) {
        const int2 entry = queue[t];
        const int qi = entry.x;
        const uint32_t bound = static_cast<uint32_t>(entry.y);
        const float4 q = pts[qi];
        // the query's cloud: points and radix nodes [cbase, cend)
        const int2 cloud = cm.at(qi, total);
        const int cbase = cloud.x, U = cloud.y, cend = cloud.x + cloud.y;
        // seed: the 64 Morton neighbours of the query, group-sorted; the K
        // nearest of them start the list
        int wbase = qi - 2 * kGroup;
        if (wbase > cend - 4 * kGroup) wbase = cend - 4 * kGroup;
        if (wbase < cbase) wbase = cbase;
        const int wend = wbase + 4 * kGroup - 1 < cend - 1 ? wbase + 4 * kGroup - 1 : cend - 1;
        GroupList<K> list;
        {
            uint64_t sk[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = wbase + r * kGroup + gl;
                sk[r] = ~0ull;
                if (j <= wend) {
                    const float4 p = pts[j];
                    sk[r] = make_key(dist2(q.x, q.y, q.z, p.x, p.y, p.z), j);
                }
            }
            group_sort64(sk, gl);
#pragma unroll
            for (int e = 0; e < GroupList<K>::E; ++e) list.key[e] = e * kGroup + gl < K ? sk[e] : ~0ull;
        }
        uint64_t kth = list.kth();
        int sp = 0, nr = 0;
        bool overflow = false;
        if (U <= 4 * kGroup) {
            sp = 0;  // the seed holds every point
        } else if (U <= bucket) {
            if (gl == 0) rng[0] = make_int2(cbase, cend - 1);
            nr = 1;
        } else {
            if (gl == 0) stk[0] = make_int2(cbase, 0);  // the cloud's root
            sp = 1;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        while (sp > 0 || nr > 0) {
            if (sp > 0) {
                const int c = sp < kGroup ? sp : kGroup;
                sp -= c;
                const int2 e = gl < c ? stk[sp + gl] : make_int2(-1, -1);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                const uint32_t kb = static_cast<uint32_t>(kth >> 32);
                const uint32_t kd = kb < bound ? kb : bound;
                bool nodeL = false, nodeR = false, rangeL = false, rangeR = false;
                int a = 0, g = 0, b = 0;
                uint32_t dlb = 0, drb = 0;
                if (gl < c && static_cast<uint32_t>(e.y) <= kd) {
                    const TravNode nd = nodes[e.x];
                    a = w_int(nd.lo_l.w);
                    g = w_int(nd.hi_l.w);
                    b = w_int(nd.lo_r.w);
                    dlb = __float_as_uint(box_dist2(q.x, q.y, q.z, nd.lo_l, nd.hi_l));
                    drb = __float_as_uint(box_dist2(q.x, q.y, q.z, nd.lo_r, nd.hi_r));
                    if (dlb <= kd) {
                        if (g - a < bucket) rangeL = true; else nodeL = true;
                    }
                    if (drb <= kd) {
                        if (b - g - 1 < bucket) rangeR = true; else nodeR = true;
                    }
                }
                const uint32_t mnl = group_ballot(nodeL, gs), mnr = group_ballot(nodeR, gs);
                const uint32_t mrl = group_ballot(rangeL, gs), mrr = group_ballot(rangeR, gs);
                const int pushes = __popc(mnl) + __popc(mnr);
                if (sp + pushes > kGroupStack) {
                    overflow = true;
                    break;
                }
                const uint32_t below = (1u << gl) - 1u;
                const int pnl = sp + __popc(mnl & below);
                const int pnr = sp + __popc(mnl) + __popc(mnr & below);
                const int prl = nr + __popc(mrl & below);
                const int prr = nr + __popc(mrl) + __popc(mrr & below);
                if (nodeL) stk[pnl] = make_int2(g, static_cast<int>(dlb));
                if (nodeR) stk[pnr] = make_int2(g + 1, static_cast<int>(drb));
                if (rangeL) rng[prl] = make_int2(a, g);
                if (rangeR) rng[prr] = make_int2(g + 1, b);
                sp += pushes;
                nr += __popc(mrl) + __popc(mrr);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
            if (nr > 0) {
                // inclusive size prefix of the ranges (nr <= kGroupRanges)
                int run = 0;
                for (int r0 = 0; r0 < nr; r0 += kGroup) {
                    const int idx = r0 + gl;
                    const int2 rg = idx < nr ? rng[idx] : make_int2(0, -1);
                    const int incl = run + group_incl_sum(rg.y - rg.x + 1, gl);
                    if (idx < nr) pre[idx] = incl;
                    run = __shfl(incl, kGroup - 1, kGroup);
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                for (int f0 = 0; f0 < run; f0 += kGroup) {
                    const int f = f0 + gl;
                    uint64_t ck = ~0ull;
                    if (f < run) {
                        const int j = range_point(rng, pre, nr, f);
                        if (j < wbase || j > wend) {  // seed points are in the list already
                            const float4 p = pts[j];
                            ck = make_key(dist2(q.x, q.y, q.z, p.x, p.y, p.z), j);
                        }
                    }
                    uint32_t m = group_ballot(ck < kth && static_cast<uint32_t>(ck >> 32) <= bound, gs);
                    while (m) {
                        const int src = static_cast<int>(__builtin_ctz(m));
                        m &= m - 1u;
                        const uint64_t x = group_shfl64(ck, src);
                        if (x < kth) {
                            list.insert(x, gl, gs);
                            kth = list.kth();
                        }
                    }
                }
                nr = 0;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
        if (overflow) {
            // marked for knn_overflow_kernel, which searches it with one lane
            if (gl == 0) queue[t].x = ~qi;
            continue;
        }
        int32_t* row = out + static_cast<int64_t>(qi - q_start) * K;
        const int first = __shfl(static_cast<int>(list.key[0]), 0, kGroup) - cbase;
#pragma unroll
        for (int e = 0; e < GroupList<K>::E; ++e) {
            const int idx = e * kGroup + gl;
            if (idx < K)
                row[idx] = list.key[e] == ~0ull ? first : static_cast<int32_t>(list.key[e] & 0xFFFFFFFFull) - cbase;
        }
    }
}

// Queue entries the group kernel marked (index complemented): one lane each,
// knn_search bounded by the entry's k-th distance bound.
template <int K>
__global__ void __launch_bounds__(kKnnBlock) knn_overflow_kernel(const float4* __restrict__ pts,
                                                                 const TravNode* __restrict__ nodes, ParentRef par,
                                                                 const int32_t* __restrict__ d_counts,
                                                                 int64_t q_start, const int2* __restrict__ queue,
                                                                 const int32_t* __restrict__ queue_count,
                                                                 int bucket, int32_t* __restrict__ out, CloudMap cm) {
    const int total = d_counts[0];
    const int n = *queue_count;
    const int stride = static_cast<int>(gridDim.x) * kKnnBlock;
    for (int t = static_cast<int>(blockIdx.x) * kKnnBlock + static_cast<int>(threadIdx.x); t < n; t += stride) {
        const int2 e = queue[t];
        if (e.x >= 0) continue;
        const int qi = ~e.x;
        const float4 q = pts[qi];
        const int2 cloud = cm.at(qi, total);
        TopK<K> top;
        top.init();
        knn_search<K>(pts, nodes, par, cloud.x, cloud.y, q.x, q.y, q.z, static_cast<uint32_t>(e.y), bucket, top);
        write_row<K>(out + static_cast<int64_t>(qi - q_start) * K, top, cloud.x);
    }
}

template <int K>
int knn_tiles_k(const float4* pts, const TravNode* nodes, ParentRef par, const int32_t* d_counts, int64_t q_start, int64_t n_q,
                const float* radius2, int bucket, int32_t* out, int2* queue, int32_t* queue_count,
                int32_t* deferred_total, hipStream_t st, CloudMap cm, int queue_blocks, int32_t* cand_total) {
    const int64_t tiles = (n_q + kWave - 1) / kWave;
    const unsigned g = static_cast<unsigned>((tiles + kTileWaves - 1) / kTileWaves);
    if (radius2)
        knn_tile_kernel<K, true><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, radius2, bucket, out,
                                                         queue, queue_count, cand_total, cm);
    else
        knn_tile_kernel<K, false><<<g, kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, n_q, nullptr, bucket, out,
                                                          queue, queue_count, cand_total, cm);
    SX_LAUNCH_CHECK("knn_tile_kernel");
    constexpr int kGroups = kTileWaves * kGroupsPerWave;
    int64_t gq = (n_q + kGroups - 1) / kGroups;  // one group per query at most
    if (gq > 8192) gq = 8192;
    if (queue_blocks > 0 && queue_blocks < gq) gq = queue_blocks;
    // The 16-lane groups stream a leaf range one point per lane, so they stop
    // descending at ranges of at most 16 points whatever the tiles use (8M
    // scene with tile ranges of 32: kNN 2.69 -> 2.60 ms; 8 or 12: 3.4 / 3.0 ms).
    const int group_bucket = bucket < kGroup ? bucket : kGroup;
    knn_group_kernel<K><<<static_cast<unsigned>(gq), kKnnBlock, 0, st>>>(pts, nodes, d_counts, q_start, queue,
                                                                       queue_count, group_bucket, out, deferred_total,
                                                                       cm);
    SX_LAUNCH_CHECK("knn_group_kernel");
    knn_overflow_kernel<K><<<64, kKnnBlock, 0, st>>>(pts, nodes, par, d_counts, q_start, queue, queue_count, bucket, out,
                                                     cm);
    SX_LAUNCH_CHECK("knn_overflow_kernel");
    return SX_OK;
}

// Index queries through the tiles.  queue: n_q int2; queue_count: one int32,
// cleared here; deferred_total accumulates the number of queued queries;
// cand_total (nullable) the candidates the tiles broadcast.
int knn_tiles_launch(const float4* pts, const TravNode* nodes, ParentRef par, const int32_t* d_counts, int64_t q_start,
                     int64_t n_q, int k, const float* radius2, int bucket, int32_t* out, int2* queue, int32_t* queue_count,
                     int32_t* deferred_total, hipStream_t st, CloudMap cm = CloudMap(), int queue_blocks = 0,
                     int32_t* cand_total = nullptr) {
    if (n_q <= 0) return SX_OK;
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0) return fail(SX_EINVAL, "knn: output must be 16-byte aligned");
    SX_TRY(fill_i32(queue_count, 0, 1, st));
    switch (k) {
        case 8: return knn_tiles_k<8>(pts, nodes, par, d_counts, q_start, n_q, radius2, bucket, out, queue, queue_count, deferred_total, st, cm, queue_blocks, cand_total);
        case 16: return knn_tiles_k<16>(pts, nodes, par, d_counts, q_start, n_q, radius2, bucket, out, queue, queue_count, deferred_total, st, cm, queue_blocks, cand_total);
        case 24: return knn_tiles_k<24>(pts, nodes, par, d_counts, q_start, n_q, radius2, bucket, out, queue, queue_count, deferred_total, st, cm, queue_blocks, cand_total);
        case 32: return knn_tiles_k<32>(pts, nodes, par, d_counts, q_start, n_q, radius2, bucket, out, queue, queue_count, deferred_total, st, cm, queue_blocks, cand_total);
        default: return fail(SX_EINVAL, "knn: k must be 8, 16, 24 or 32");
    }
}

}  // namespace sx
