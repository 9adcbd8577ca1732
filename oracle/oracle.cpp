// CPU ORACLE — test infrastructure only.
//
// A plain C++ restatement of Wehrdo/ShapeExtraction's normal-estimation path,
// used as the parity checker of the GPU library and as the CPU baseline of the
// benchmark.  Nothing in the product links or calls this file.  Each function
// cites the reference lines it restates; where the reference has no answer
// (exact kNN, FP64 PCA), the function states the definition it implements.
//
// Built with g++ -O3 -fopenmp -ffp-contract=off (oracle/Makefile): every float
// operation is a single IEEE round-to-nearest operation, which is what the GPU
// kernels do on their bit-exact paths.
//
// Pinning (tests/test_oracle.py): Morton codes by hand-derived known answers
// and round trips; sort/unique by numpy; the Karras restatement by an
// independent recursive trie split; the octree by brute-force enumeration of
// Morton cells; exact kNN by two independent methods (brute force, uniform
// grid); the FP64 Jacobi by numpy.linalg.eigh (LAPACK).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <utility>
#include <vector>

#include <omp.h>

namespace {

struct OTNode {  // OT::OTNode, src/Octree.hpp:19-35
    int32_t children[8];
    float corner[3];
    float cell_size;
    int32_t child_node_mask;
    int32_t child_leaf_mask;
};
static_assert(sizeof(OTNode) == 56, "OTNode layout");

// ---------------------------------------------------------------- Morton
inline uint32_t scale_of(int code_bits) { return 0xFFFFFFFFu >> (32 - code_bits / 3); }

// Bit-by-bit interleave: axis a's bit i goes to code bit 3i+a (x=0, y=1, z=2),
// the libmorton morton3D_64_encode convention (src/MortonUtils.cu:18).
inline uint64_t interleave_loop(uint32_t x, uint32_t y, uint32_t z, int bits) {
    uint64_t c = 0;
    for (int i = 0; i < bits; ++i) {
        c |= static_cast<uint64_t>((x >> i) & 1u) << (3 * i);
        c |= static_cast<uint64_t>((y >> i) & 1u) << (3 * i + 1);
        c |= static_cast<uint64_t>((z >> i) & 1u) << (3 * i + 2);
    }
    return c;
}
inline void deinterleave_loop(uint64_t c, int bits, uint32_t& x, uint32_t& y, uint32_t& z) {
    x = y = z = 0;
    for (int i = 0; i < bits; ++i) {
        x |= static_cast<uint32_t>((c >> (3 * i)) & 1u) << i;
        y |= static_cast<uint32_t>((c >> (3 * i + 1)) & 1u) << i;
        z |= static_cast<uint32_t>((c >> (3 * i + 2)) & 1u) << i;
    }
}

// pointToCode, src/MortonUtils.cu:7-19: (uint32)(bitscale * ((v - min) / range)).
// A zero range (all points equal) maps to 0; the reference divides by zero.
inline uint32_t quantize(float v, float mn, float range, int code_bits) {
    if (!(range > 0.0f)) return 0u;
    const float fs = static_cast<float>(scale_of(code_bits));
    float t = (v - mn) / range;
    float s = fs * t;
    if (s < 0.0f) s = 0.0f;
    if (s > fs) s = fs;
    return static_cast<uint32_t>(s);
}

// codeToPoint, src/MortonUtils.cu:21-28, uncontracted.
inline float dequantize(uint32_t raw, float mn, float range, int code_bits) {
    const float fs = static_cast<float>(scale_of(code_bits));
    float a = static_cast<float>(raw) / fs;
    float b = a * range;
    return b + mn;
}

// Point::distance2, src/PointCloud.hpp:30-35, order ((dx*dx + dy*dy) + dz*dz).
inline float dist2(const float* q, const float* p) {
    float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    float xx = dx * dx, yy = dy * dy, zz = dz * dz;
    float s = xx + yy;
    return s + zz;
}

// delta, src/RadixTree.cu:62-69, generalised to code_bits significant bits.
inline int delta(uint64_t a, uint64_t b, int code_bits) {
    uint64_t x = a ^ b;
    if (x == 0) return code_bits;
    return __builtin_clzll(x) - (64 - code_bits);
}

inline uint64_t prefix_of(uint64_t code, int level, int code_bits) {
    int sh = code_bits - 3 * level;
    return sh >= 64 ? 0ull : (code >> sh);
}

inline uint64_t dkey(float d, int32_t idx) {
    uint32_t b;
    std::memcpy(&b, &d, 4);
    return (static_cast<uint64_t>(b) << 32) | static_cast<uint32_t>(idx);
}

// ---------------------------------------------------------------- Jacobi
// Cyclic Jacobi (Golub & Van Loan, Matrix Computations, sec. 8.5) for a
// symmetric 3x3 in double; ascending eigenvalues, eigenvectors as columns.
void jacobi3(const double* Ain, double* evals, double* evecs) {
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = 0.5 * (Ain[3 * i + j] + Ain[3 * j + i]);
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        double dg = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (!(off > 1e-40 * dg)) break;
        const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
        for (int r = 0; r < 3; ++r) {
            int p = P[r], q = Q[r];
            double apq = a[p][q];
            if (apq == 0.0) continue;
            double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
            double t = std::fabs(theta) > 1e150 ? 0.5 / theta
                                                 : (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
            double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
            for (int k = 0; k < 3; ++k) {  // A <- A J
                double akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq;
                a[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {  // A <- J^T A
                double apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk;
                a[q][k] = s * apk + c * aqk;
            }
            a[p][q] = a[q][p] = 0.0;
            for (int k = 0; k < 3; ++k) {
                double vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int i, int j) { return a[i][i] < a[j][j] || (a[i][i] == a[j][j] && i < j); });
    for (int c = 0; c < 3; ++c) {
        evals[c] = a[idx[c]][idx[c]];
        for (int r = 0; r < 3; ++r) evecs[3 * r + c] = v[r][idx[c]];
    }
}

// FP32 Jacobi standing in for cusolverDnSsyevjBatched with tol 1e-7, 15
// sweeps, ascending order (src/NormalEstimation.hpp:250-299).  Column-major
// output like syevj: vectors[3*c + r].
void jacobi3_f32(const float* A, float* evals, float* vecs_colmajor) {
    float a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = A[3 * j + i];
    float fro = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) fro += a[i][j] * a[i][j];
    fro = std::sqrt(fro);
    for (int sweep = 0; sweep < 15; ++sweep) {
        float off = std::sqrt(2.0f * (a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2]));
        if (!(off > 1e-7f * fro)) break;
        const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
        for (int r = 0; r < 3; ++r) {
            int p = P[r], q = Q[r];
            float apq = a[p][q];
            if (apq == 0.0f) continue;
            float theta = (a[q][q] - a[p][p]) / (2.0f * apq);
            float t = (theta >= 0 ? 1.0f : -1.0f) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0f));
            float c = 1.0f / std::sqrt(t * t + 1.0f), s = t * c;
            for (int k = 0; k < 3; ++k) {
                float akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq;
                a[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {
                float apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk;
                a[q][k] = s * apk + c * aqk;
            }
            a[p][q] = a[q][p] = 0.0f;
            for (int k = 0; k < 3; ++k) {
                float vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int i, int j) { return a[i][i] < a[j][j] || (a[i][i] == a[j][j] && i < j); });
    for (int c = 0; c < 3; ++c) {
        evals[c] = a[idx[c]][idx[c]];
        for (int r = 0; r < 3; ++r) vecs_colmajor[3 * c + r] = v[r][idx[c]];
    }
}

// ---------------------------------------------------------------- reference kNN emulation
// PriorityQueue<T, MAX_Q> of src/dequeue.hpp restated: a fixed-capacity
// min-max heap (even levels min, odd levels max) whose insert, when full,
// replaces the maximum only if the maximum's priority is >= the new one.
template <typename T, int CAP>
struct MinMaxHeap {
    struct Item { T data; float pri; };
    Item h[CAP];
    int size = 0;
    static int lvl(int i) { int r = 0; for (int x = i + 1; x >>= 1;) ++r; return r; }
    void swp(int a, int b) { Item t = h[a]; h[a] = h[b]; h[b] = t; }
    void up_min(int i) { while (i >= 3) { int g = ((i - 1) / 2 - 1) / 2; if (h[i].pri < h[g].pri) { swp(i, g); i = g; } else break; } }
    void up_max(int i) { while (i >= 3) { int g = ((i - 1) / 2 - 1) / 2; if (h[i].pri > h[g].pri) { swp(i, g); i = g; } else break; } }
    void up(int i) {
        int p = (i - 1) / 2;
        if (lvl(i) % 2 == 0) {
            if (i > 0 && h[i].pri > h[p].pri) { swp(i, p); up_max(p); } else up_min(i);
        } else {
            if (i > 0 && h[i].pri < h[p].pri) { swp(i, p); up_min(p); } else up_max(i);
        }
    }
    // percolateDownMin / percolateDownMax (src/dequeue.hpp:213-289)
    void down(int i, bool is_min) {
        while (2 * i + 1 < size) {
            int c1 = 2 * i + 1;
            int cand[5] = {2 * i + 2, 2 * c1 + 1, 2 * c1 + 2, 2 * (2 * i + 2) + 1, 2 * (2 * i + 2) + 2};
            int m = c1;
            for (int d : cand)
                if (d < size && (is_min ? h[d].pri < h[m].pri : h[d].pri > h[m].pri)) m = d;
            bool better = is_min ? h[m].pri < h[i].pri : h[m].pri > h[i].pri;
            if (m > 2 * i + 2) {
                if (!better) break;
                swp(i, m);
                int pm = (m - 1) / 2;
                if (is_min ? h[m].pri > h[pm].pri : h[m].pri < h[pm].pri) swp(pm, m);
                i = m;
            } else {
                if (better) swp(i, m);
                break;
            }
        }
    }
    void down_at(int i) { down(i, lvl(i) % 2 == 0); }
    int max_idx() const { return size >= 3 ? (h[1].pri > h[2].pri ? 1 : 2) : (size == 2 ? 1 : 0); }
    void insert(T d, float pri) {
        if (size == CAP) {
            if (h[max_idx()].pri >= pri) remove_max(); else return;
        }
        h[size] = {d, pri};
        ++size;
        up(size - 1);
    }
    Item remove_min() { Item r = h[0]; --size; h[0] = h[size]; down_at(0); return r; }
    Item remove_max() { int m = max_idx(); Item r = h[m]; --size; h[m] = h[size]; down_at(m); return r; }
};

// nodeBoxDistance, src/Octree.cu:327-352: 0 strictly inside the cell, else the
// minimum distance to any of the six face planes (linear, not squared).
float node_box_distance(const float* q, const OTNode& n) {
    float x1 = n.corner[0], x2 = n.corner[0] + n.cell_size;
    float y1 = n.corner[1], y2 = n.corner[1] + n.cell_size;
    float z1 = n.corner[2], z2 = n.corner[2] + n.cell_size;
    bool inside = q[0] > x1 && q[0] < x2 && q[1] > y1 && q[1] < y2 && q[2] > z1 && q[2] < z2;
    if (inside) return 0.0f;
    float m = std::fabs(q[0] - x1);
    m = std::fmin(m, std::fabs(q[0] - x2));
    m = std::fmin(m, std::fabs(q[1] - y1));
    m = std::fmin(m, std::fabs(q[1] - y2));
    m = std::fmin(m, std::fabs(q[2] - z1));
    m = std::fmin(m, std::fabs(q[2] - z2));
    return m;
}

template <int K>
void knn_reference_one(const OTNode* nodes, const float* pts, const float* q, int32_t* out) {
    MinMaxHeap<int, 32> search;   // SEARCH_Q_SIZE, src/Octree.hpp:17
    MinMaxHeap<int, K> result;
    search.insert(0, 0.0f);
    while (search.size) {  // d stays INFINITY, so only the queue size ends the loop
        auto top = search.remove_min();
        const OTNode& n = nodes[top.data];
        if (n.child_leaf_mask)
            for (int i = 0; i < 8; ++i)
                if (n.child_leaf_mask & (1 << i)) result.insert(n.children[i], dist2(q, pts + 3 * n.children[i]));
        for (int i = 0; i < 8; ++i)
            if (n.child_node_mask & (1 << i)) search.insert(n.children[i], node_box_distance(q, nodes[n.children[i]]));
    }
    for (int i = 0; i < K; ++i) out[i] = result.size == 0 ? out[0] : result.remove_min().data;
}

// ---------------------------------------------------------------- exact kNN helpers
void top_k_push(std::vector<uint64_t>& heap, int k, uint64_t key) {
    if (static_cast<int>(heap.size()) < k) {
        heap.push_back(key);
        std::push_heap(heap.begin(), heap.end());
    } else if (key < heap.front()) {
        std::pop_heap(heap.begin(), heap.end());
        heap.back() = key;
        std::push_heap(heap.begin(), heap.end());
    }
}

void write_row(std::vector<uint64_t>& heap, int k, int32_t* out) {
    std::sort(heap.begin(), heap.end());
    for (int s = 0; s < k; ++s)
        out[s] = s < static_cast<int>(heap.size()) ? static_cast<int32_t>(heap[s] & 0xFFFFFFFFu)
                                                   : static_cast<int32_t>(heap[0] & 0xFFFFFFFFu);
}

}  // namespace

extern "C" {

int orc_abi_version(void) { return 1; }
int orc_threads(void) { return omp_get_max_threads(); }

// Global cube: min and max over all finite coordinates of all three axes
// (src/RadixTree.cu:208-241).  Returns the number of non-finite values.
int64_t orc_bbox(const float* x, const float* y, const float* z, int64_t n, float* out3) {
    float mn = std::numeric_limits<float>::infinity(), mx = -mn;
    int64_t bad = 0;
    const float* a[3] = {x, y, z};
    for (int ax = 0; ax < 3; ++ax)
        for (int64_t i = 0; i < n; ++i) {
            float v = a[ax][i];
            if (!std::isfinite(v)) { ++bad; continue; }
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
    out3[0] = mn;
    out3[1] = mx;
    out3[2] = mx - mn;
    return bad;
}

void orc_encode(const float* x, const float* y, const float* z, int64_t n, float mn, float range, int code_bits,
                uint64_t* out) {
    const int bits = code_bits / 3;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        out[i] = interleave_loop(quantize(x[i], mn, range, code_bits), quantize(y[i], mn, range, code_bits),
                                 quantize(z[i], mn, range, code_bits), bits);
}

void orc_decode(const uint64_t* codes, int64_t n, float mn, float range, int code_bits, float* pts) {
    const int bits = code_bits / 3;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        uint32_t rx, ry, rz;
        deinterleave_loop(codes[i], bits, rx, ry, rz);
        pts[3 * i] = dequantize(rx, mn, range, code_bits);
        pts[3 * i + 1] = dequantize(ry, mn, range, code_bits);
        pts[3 * i + 2] = dequantize(rz, mn, range, code_bits);
    }
}

uint64_t orc_interleave(uint32_t x, uint32_t y, uint32_t z, int code_bits) {
    return interleave_loop(x, y, z, code_bits / 3);
}

// Stable ascending sort of keys carrying their input index.
void orc_sort_pairs(const uint64_t* keys, int64_t n, uint64_t* keys_out, uint32_t* perm) {
    std::vector<std::pair<uint64_t, uint32_t>> v(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) v[i] = {keys[i], static_cast<uint32_t>(i)};
    std::sort(v.begin(), v.end());  // (key, index) order == stable order by key
    for (int64_t i = 0; i < n; ++i) { keys_out[i] = v[i].first; perm[i] = v[i].second; }
}

// removeDuplicates (src/RadixTree.cu:258-303): distinct codes in order, the
// first element of each run as representative.  Returns U.
int64_t orc_unique(const uint64_t* sorted, const uint32_t* perm, int64_t n, uint64_t* uniq, uint32_t* rep,
                   int32_t* unique_of_input) {
    int64_t u = -1;
    for (int64_t i = 0; i < n; ++i) {
        if (i == 0 || sorted[i] != sorted[i - 1]) {
            ++u;
            uniq[u] = sorted[i];
            if (rep) rep[u] = perm[i];
        }
        if (unique_of_input) unique_of_input[perm[i]] = static_cast<int32_t>(u);
    }
    return u + 1;
}

// constructTree (src/RadixTree.cu:99-194) for every internal node, restated.
// root_mode 1 = the reference's root length N-1 (src/RadixTree.cu:128-131),
// 0 = textbook N.  Arrays are sized U; parent[] of unclaimed nodes is -1.
void orc_karras(const uint64_t* codes, int64_t U, int code_bits, int root_mode, uint8_t* pn, int32_t* lc,
                uint8_t* hll, uint8_t* hlr, int32_t* parent) {
    const int64_t N = U - 1;
    for (int64_t i = 0; i < U; ++i) parent[i] = -1;
    for (int64_t i = 0; i < N; ++i) {
        const uint64_t ci = codes[i];
        int d = 1;
        if (i > 0) {
            int dr = delta(ci, codes[i + 1], code_bits), dl = delta(ci, codes[i - 1], code_bits);
            d = dr > dl ? 1 : (dr < dl ? -1 : 0);
        }
        int64_t l = 0;
        if (i == 0) {
            l = (root_mode == 1 && N >= 2) ? N - 1 : N;
        } else {
            int dmin = delta(ci, codes[i - d], code_bits);
            int64_t lmax = 2;
            while (i + lmax * d >= 0 && i + lmax * d <= N && delta(ci, codes[i + lmax * d], code_bits) > dmin) lmax *= 2;
            int64_t cut = d == -1 ? i : N - i;
            for (int64_t div = 2, t = lmax / 2; t >= 1; div *= 2, t = lmax / div)
                if (l + t <= cut && delta(ci, codes[i + (l + t) * d], code_bits) > dmin) l += t;
        }
        int64_t j = i + l * d;
        int dnode = delta(ci, codes[j], code_bits);
        pn[i] = static_cast<uint8_t>(dnode);
        int64_t s = 0;
        int log2c = 0;
        while ((int64_t(1) << log2c) < l) ++log2c;  // log2_ceil, src/RadixTree.cu:51-59
        int64_t maxdiv = int64_t(1) << log2c;
        int64_t scut = d == -1 ? i - 1 : N - i - 1;
        for (int64_t div = 2, t = (l + 1) / 2; div <= maxdiv; div <<= 1, t = (l + div - 1) / div)
            if (s + t <= scut && delta(ci, codes[i + (s + t) * d], code_bits) > dnode) s += t;
        int64_t gamma = i + s * d + std::min(d, 0);
        lc[i] = static_cast<int32_t>(gamma);
        hll[i] = std::min(i, j) == gamma;
        hlr[i] = std::max(i, j) == gamma + 1;
        if (!hll[i]) parent[gamma] = static_cast<int32_t>(i);
        if (!hlr[i]) parent[gamma + 1] = static_cast<int32_t>(i);
    }
}

// Independent binary radix trie (no Karras search): recursively split the
// range [a, b] after the last code whose bit below the common prefix is 0.
// The node covering [a, b] gets index a if it is a right child or the root,
// b if it is a left child (the Karras numbering).  Same outputs as orc_karras.
void orc_trie(const uint64_t* codes, int64_t U, int code_bits, uint8_t* pn, int32_t* lc, uint8_t* hll, uint8_t* hlr,
              int32_t* parent) {
    for (int64_t i = 0; i < U; ++i) parent[i] = -1;
    if (U < 2) return;
    struct Job { int64_t a, b, id, par; };
    std::vector<Job> st{{0, U - 1, 0, -1}};
    while (!st.empty()) {
        Job jb = st.back();
        st.pop_back();
        int p = delta(codes[jb.a], codes[jb.b], code_bits);  // common prefix of the whole range
        uint64_t bit = uint64_t(1) << (code_bits - 1 - p);
        int64_t g = jb.a;
        while (g + 1 <= jb.b && !(codes[g + 1] & bit)) ++g;  // last code with that bit clear
        pn[jb.id] = static_cast<uint8_t>(p);
        lc[jb.id] = static_cast<int32_t>(g);
        hll[jb.id] = (g == jb.a);
        hlr[jb.id] = (g + 1 == jb.b);
        parent[jb.id] = static_cast<int32_t>(jb.par);
        if (g > jb.a) st.push_back({jb.a, g, g, jb.id});
        if (g + 1 < jb.b) st.push_back({g + 1, jb.b, g + 1, jb.id});
    }
    parent[0] = -1;
}

// Octree from the radix tree (src/Octree.cu:213-319): edge counts, exclusive
// offsets, node chains (makeNodes, :114-161), leaf links (linkLeafNodes,
// :65-112), root at node 0 with cell_size = range (:257-261).  Returns the
// node count O (nodes written only when O <= cap).
int64_t orc_octree(const uint64_t* codes, int64_t U, int code_bits, float mn, float range, const uint8_t* pn,
                   const int32_t* lc, const uint8_t* hll, const uint8_t* hlr, const int32_t* parent, void* nodes_out,
                   int64_t cap) {
    if (U < 2) return 0;
    const int64_t M = U - 1;
    std::vector<int32_t> count(M), off(M + 1, 0);
    count[0] = 1;
    for (int64_t i = 1; i < M; ++i) count[i] = parent[i] < 0 ? 0 : pn[i] / 3 - pn[parent[i]] / 3;
    for (int64_t i = 0; i < M; ++i) off[i + 1] = off[i] + count[i];
    const int64_t O = off[M];
    if (O > cap || !nodes_out) return O;
    OTNode* nodes = static_cast<OTNode*>(nodes_out);
    std::memset(nodes, 0, sizeof(OTNode) * O);
    const int bits = code_bits / 3;
    auto corner = [&](uint64_t prefix, int level, float* c) {
        uint64_t code = level == 0 ? 0 : prefix << (code_bits - 3 * level);
        uint32_t rx, ry, rz;
        deinterleave_loop(code, bits, rx, ry, rz);
        c[0] = dequantize(rx, mn, range, code_bits);
        c[1] = dequantize(ry, mn, range, code_bits);
        c[2] = dequantize(rz, mn, range, code_bits);
    };
    const int root_level = pn[0] / 3;
    corner(prefix_of(codes[0], root_level, code_bits), root_level, nodes[0].corner);
    nodes[0].cell_size = range;
    auto set_child = [&](int node, int child, int slot) {
        nodes[node].children[slot] = child;
        nodes[node].child_node_mask |= 1 << slot;
    };
    for (int64_t i = 1; i < M; ++i) {
        int oct = off[i];
        const int n_new = count[i];
        for (int j = 0; j < n_new - 1; ++j) {
            int level = pn[i] / 3 - j;
            uint64_t pref = prefix_of(codes[i], level, code_bits);
            set_child(oct + 1, oct, static_cast<int>(pref & 7));
            corner(pref, level, nodes[oct].corner);
            nodes[oct].cell_size = range / static_cast<float>(1 << (level - root_level));
            ++oct;
        }
        if (n_new > 0) {
            int rp = parent[i];
            while (rp >= 0 && count[rp] == 0) rp = parent[rp];
            int top = pn[i] / 3 - n_new + 1;
            uint64_t pref = prefix_of(codes[i], top, code_bits);
            if (rp >= 0) set_child(off[rp], oct, static_cast<int>(pref & 7));
            corner(pref, top, nodes[oct].corner);
            nodes[oct].cell_size = range / static_cast<float>(1 << (top - root_level));
        }
    }
    for (int64_t i = 0; i < M; ++i) {
        for (int side = 0; side < 2; ++side) {
            if (!(side == 0 ? hll[i] : hlr[i])) continue;
            int leaf = lc[i] + side;
            int level = pn[i] / 3 + 1;
            int slot = static_cast<int>(prefix_of(codes[leaf], level, code_bits) & 7);
            int r = static_cast<int>(i);
            while (r >= 0 && count[r] == 0) r = parent[r];
            if (r < 0) continue;
            nodes[off[r]].children[slot] = leaf;
            nodes[off[r]].child_leaf_mask |= 1 << slot;
        }
    }
    return O;
}

// Brute-force Morton-cell census: for every level L in [root_level, bits),
// the cells (code prefixes of 3L bits) holding >= 2 of the sorted unique
// codes; writes (level, prefix, first, last) per cell in level-major order
// into cells[4*c..] when cells != NULL.  Returns the number of cells.
int64_t orc_octree_cells(const uint64_t* codes, int64_t U, int code_bits, int root_level, int64_t* cells) {
    int64_t c = 0;
    const int bits = code_bits / 3;
    for (int L = root_level; L < bits + 1; ++L) {
        int64_t a = 0;
        while (a < U) {
            uint64_t p = prefix_of(codes[a], L, code_bits);
            int64_t b = a;
            while (b + 1 < U && prefix_of(codes[b + 1], L, code_bits) == p) ++b;
            if (b > a) {
                if (cells) { cells[4 * c] = L; cells[4 * c + 1] = static_cast<int64_t>(p); cells[4 * c + 2] = a; cells[4 * c + 3] = b; }
                ++c;
            }
            a = b + 1;
        }
    }
    return c;
}

// Exact kNN by brute force: for each of the nq query points qpts (3 floats
// each), the k points of pts (3*U floats) with the smallest (distance2,
// index), distance2 in the uncontracted order of Point::distance2
// (src/PointCloud.hpp:30-35); rows padded with their first entry when U < k.
static void knn_brute_rows(const float* pts, int64_t U, const float* qpts, int64_t nq, int k, int32_t* out) {
    std::vector<float> X(U), Y(U), Z(U);
    for (int64_t i = 0; i < U; ++i) { X[i] = pts[3 * i]; Y[i] = pts[3 * i + 1]; Z[i] = pts[3 * i + 2]; }
#pragma omp parallel
    {
        std::vector<float> d(4096);
        std::vector<uint64_t> heap;
#pragma omp for schedule(dynamic, 16)
        for (int64_t t = 0; t < nq; ++t) {
            const float* q = qpts + 3 * t;
            const float qx = q[0], qy = q[1], qz = q[2];
            heap.clear();
            float thr = std::numeric_limits<float>::infinity();
            for (int64_t b0 = 0; b0 < U; b0 += 4096) {
                const int64_t b1 = std::min(U, b0 + 4096);
                for (int64_t j = b0; j < b1; ++j) {
                    float dx = qx - X[j], dy = qy - Y[j], dz = qz - Z[j];
                    float xx = dx * dx, yy = dy * dy, zz = dz * dz;
                    float s = xx + yy;
                    d[j - b0] = s + zz;
                }
                for (int64_t j = b0; j < b1; ++j) {
                    if (d[j - b0] <= thr) {
                        top_k_push(heap, k, dkey(d[j - b0], static_cast<int32_t>(j)));
                        if (static_cast<int>(heap.size()) == k) {
                            uint32_t hb = static_cast<uint32_t>(heap.front() >> 32);
                            std::memcpy(&thr, &hb, 4);
                        }
                    }
                }
            }
            write_row(heap, k, out + t * k);
        }
    }
}

// Index points [q0, q0+nq) as the queries: the query itself is included.
void orc_knn_brute(const float* pts, int64_t U, int64_t q0, int64_t nq, int k, int32_t* out) {
    knn_brute_rows(pts, U, pts + 3 * q0, nq, k, out);
}

// Arbitrary query points (the query API of Octree::knnSearch<k>,
// src/Octree.hpp:214-242).
void orc_knn_brute_points(const float* pts, int64_t U, const float* queries, int64_t nq, int k, int32_t* out) {
    knn_brute_rows(pts, U, queries, nq, k, out);
}

// Exact kNN over a uniform grid (independent of any tree): points bucketed by
// cell; a query visits Chebyshev rings r = 0, 1, ... of cells around its own
// and stops once the k-th best squared distance is below the squared lower
// bound (r*h)^2 of every unvisited ring, with a relative margin for float
// rounding.  Queries are given as indices into pts.
void orc_knn_grid(const float* pts, int64_t U, const int64_t* qidx, int64_t nq, int k, int32_t* out) {
    float lo[3] = {pts[0], pts[1], pts[2]}, hi[3] = {pts[0], pts[1], pts[2]};
    for (int64_t i = 0; i < U; ++i)
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], pts[3 * i + a]); hi[a] = std::max(hi[a], pts[3 * i + a]); }
    double ext = std::max({double(hi[0]) - lo[0], double(hi[1]) - lo[1], double(hi[2]) - lo[2], 1e-30});
    // about 8 points per occupied cell for surfaces: h ~ ext * sqrt(8 / U) (area-like scaling)
    double h = ext * std::sqrt(8.0 / std::max<int64_t>(U, 1));
    h = std::max(h, ext / 1.0e6);
    const int64_t G = std::min<int64_t>(int64_t(ext / h) + 1, 2000000);
    auto cell_of = [&](const float* p, int64_t* c) {
        for (int a = 0; a < 3; ++a) {
            int64_t v = static_cast<int64_t>((double(p[a]) - lo[a]) / h);
            c[a] = std::min<int64_t>(std::max<int64_t>(v, 0), G - 1);
        }
    };
    auto key_of = [&](int64_t cx, int64_t cy, int64_t cz) { return (uint64_t(cx) << 42) | (uint64_t(cy) << 21) | uint64_t(cz); };
    std::vector<std::pair<uint64_t, int32_t>> cells(U);
    for (int64_t i = 0; i < U; ++i) {
        int64_t c[3];
        cell_of(pts + 3 * i, c);
        cells[i] = {key_of(c[0], c[1], c[2]), static_cast<int32_t>(i)};
    }
    std::sort(cells.begin(), cells.end());
    std::vector<uint64_t> keys(U);
    for (int64_t i = 0; i < U; ++i) keys[i] = cells[i].first;
#pragma omp parallel
    {
        std::vector<uint64_t> heap;
#pragma omp for schedule(dynamic, 8)
        for (int64_t t = 0; t < nq; ++t) {
            const float* q = pts + 3 * qidx[t];
            int64_t c[3];
            cell_of(q, c);
            heap.clear();
            for (int64_t r = 0;; ++r) {
                for (int64_t dx = -r; dx <= r; ++dx)
                    for (int64_t dy = -r; dy <= r; ++dy)
                        for (int64_t dz = -r; dz <= r; ++dz) {
                            if (std::max({std::llabs(dx), std::llabs(dy), std::llabs(dz)}) != r) continue;
                            int64_t x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
                            if (x < 0 || y < 0 || z < 0 || x >= G || y >= G || z >= G) continue;
                            uint64_t kk = key_of(x, y, z);
                            auto it = std::lower_bound(keys.begin(), keys.end(), kk);
                            for (int64_t j = it - keys.begin(); j < U && keys[j] == kk; ++j) {
                                int32_t pi = cells[j].second;
                                top_k_push(heap, k, dkey(dist2(q, pts + 3 * pi), pi));
                            }
                        }
                if (static_cast<int>(heap.size()) == std::min<int64_t>(k, U)) {
                    uint32_t hb = static_cast<uint32_t>(heap.front() >> 32);
                    float kth;
                    std::memcpy(&kth, &hb, 4);
                    double bound = double(r) * h;  // every unvisited point is at least this far
                    if (double(kth) < bound * bound * (1.0 - 1e-6)) break;
                    if (static_cast<int64_t>(heap.size()) == U) break;
                }
                if (r > G) break;
            }
            write_row(heap, k, out + t * k);
        }
    }
}

// The reference kNN (knnSearchKernel<k>, src/Octree.hpp:136-193) emulated on
// the octree: diagnostic only (not exact, see the file header of knn.hip).
void orc_knn_reference(const void* nodes_in, const float* pts, int64_t U, int64_t q0, int64_t nq, int k,
                       int32_t* out) {
    const OTNode* nodes = static_cast<const OTNode*>(nodes_in);
    (void)U;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t t = 0; t < nq; ++t) {
        const float* q = pts + 3 * (q0 + t);
        int32_t* row = out + t * k;
        switch (k) {
            case 8: knn_reference_one<8>(nodes, pts, q, row); break;
            case 16: knn_reference_one<16>(nodes, pts, q, row); break;
            case 24: knn_reference_one<24>(nodes, pts, q, row); break;
            default: knn_reference_one<32>(nodes, pts, q, row); break;
        }
    }
}

// FP64 PCA: covariance of the neighbour rows in double about their mean
// (coordinates are the decoded float points), eigen-decomposed by jacobi3;
// writes the unit eigenvector of the smallest eigenvalue (sign unspecified)
// and the three ascending eigenvalues.
void orc_pca_f64(const float* pts, const int32_t* nn, int64_t nq, int k, double* normals, double* evals) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nq; ++t) {
        const int32_t* row = nn + t * k;
        double m[3] = {0, 0, 0};
        for (int s = 0; s < k; ++s)
            for (int a = 0; a < 3; ++a) m[a] += pts[3 * row[s] + a];
        for (int a = 0; a < 3; ++a) m[a] /= k;
        double A[9] = {0};
        for (int s = 0; s < k; ++s) {
            double e[3];
            for (int a = 0; a < 3; ++a) e[a] = pts[3 * row[s] + a] - m[a];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) A[3 * a + b] += e[a] * e[b];
        }
        double w[3], V[9];
        jacobi3(A, w, V);
        double nx = V[0], ny = V[3], nz = V[6];
        double l = std::sqrt(nx * nx + ny * ny + nz * nz);
        if (!(l > 0)) { nx = 1; ny = 0; nz = 0; l = 1; }
        normals[3 * t] = nx / l;
        normals[3 * t + 1] = ny / l;
        normals[3 * t + 2] = nz / l;
        if (evals) for (int a = 0; a < 3; ++a) evals[3 * t + a] = w[a];
    }
}

void orc_jacobi3(const double* A, double* evals, double* evecs) { jacobi3(A, evals, evecs); }

// Reference normals in FP32: calculateCovariance<k> (src/NormalEstimation.hpp:
// 96-120: centroid sum, /k, sums of products in neighbour order, /(k-1)), an
// FP32 Jacobi in place of syevj, column 0, then the viewpoint flip of
// normalsFromEigenVectors (:160-192).
void orc_normals_ref_f32(const float* pts, const int32_t* nn, int64_t q0, int64_t nq, int k, float vx, float vy,
                         float vz, float* normals) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nq; ++t) {
        const int32_t* row = nn + t * k;
        float c[3] = {0, 0, 0};
        for (int s = 0; s < k; ++s)
            for (int a = 0; a < 3; ++a) c[a] += pts[3 * row[s] + a];
        for (int a = 0; a < 3; ++a) c[a] /= static_cast<float>(k);
        float u[6] = {0, 0, 0, 0, 0, 0};
        for (int s = 0; s < k; ++s) {
            float m[3];
            for (int a = 0; a < 3; ++a) m[a] = pts[3 * row[s] + a] - c[a];
            u[0] += m[0] * m[0]; u[1] += m[0] * m[1]; u[2] += m[0] * m[2];
            u[3] += m[1] * m[1]; u[4] += m[1] * m[2]; u[5] += m[2] * m[2];
        }
        for (int a = 0; a < 6; ++a) u[a] /= static_cast<float>(k - 1);
        float A[9] = {u[0], u[1], u[2], u[1], u[3], u[4], u[2], u[4], u[5]};
        float w[3], V[9];
        jacobi3_f32(A, w, V);
        float n[3] = {V[0], V[1], V[2]};
        const float* p = pts + 3 * (q0 + t);
        float dot = (vx - p[0]) * n[0] + (vy - p[1]) * n[1] + (vz - p[2]) * n[2];
        if (dot < 0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
        for (int a = 0; a < 3; ++a) normals[3 * t + a] = n[a];
    }
}

// The reference pipeline on the host (CPU baseline): bbox, encode, sort,
// unique, Karras (reference root), octree, decode, reference kNN (32-slot
// queue, knn_mode 0) or the exact grid kNN (knn_mode 1),
// covariance + FP32 Jacobi + orientation.  times_ms[0..7] receive per-stage
// wall times: bbox+encode, sort, unique, radix tree, octree, decode, kNN,
// normals.  Returns U, or -1 on failure.
int64_t orc_pipeline(const float* x, const float* y, const float* z, int64_t n, int code_bits, int k, int knn_mode,
                     float* normals_out, double* times_ms) {
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    auto t0 = clk::now();
    float b[3];
    orc_bbox(x, y, z, n, b);
    std::vector<uint64_t> codes(n), sorted(n);
    orc_encode(x, y, z, n, b[0], b[2], code_bits, codes.data());
    auto t1 = clk::now();
    std::vector<uint32_t> perm(n);
    orc_sort_pairs(codes.data(), n, sorted.data(), perm.data());
    auto t2 = clk::now();
    std::vector<uint64_t> uq(n);
    std::vector<uint32_t> rep(n);
    int64_t U = orc_unique(sorted.data(), perm.data(), n, uq.data(), rep.data(), nullptr);
    auto t3 = clk::now();
    std::vector<uint8_t> pn(U), hll(U), hlr(U);
    std::vector<int32_t> lc(U), par(U);
    orc_karras(uq.data(), U, code_bits, 1, pn.data(), lc.data(), hll.data(), hlr.data(), par.data());
    auto t4 = clk::now();
    int64_t O = orc_octree(uq.data(), U, code_bits, b[0], b[2], pn.data(), lc.data(), hll.data(), hlr.data(),
                           par.data(), nullptr, 0);
    std::vector<OTNode> nodes(std::max<int64_t>(O, 1));
    orc_octree(uq.data(), U, code_bits, b[0], b[2], pn.data(), lc.data(), hll.data(), hlr.data(), par.data(),
               nodes.data(), O);
    auto t5 = clk::now();
    std::vector<float> pts(3 * U);
    orc_decode(uq.data(), U, b[0], b[2], code_bits, pts.data());
    auto t6 = clk::now();
    std::vector<int32_t> nn(U * k);
    if (knn_mode == 0) {
        orc_knn_reference(nodes.data(), pts.data(), U, 0, U, k, nn.data());
    } else {
        std::vector<int64_t> q(U);
        std::iota(q.begin(), q.end(), 0);
        orc_knn_grid(pts.data(), U, q.data(), U, k, nn.data());
    }
    auto t7 = clk::now();
    std::vector<float> nrm(3 * U);
    orc_normals_ref_f32(pts.data(), nn.data(), 0, U, k, 0, 0, 0, normals_out ? normals_out : nrm.data());
    auto t8 = clk::now();
    if (times_ms) {
        times_ms[0] = ms(t0, t1); times_ms[1] = ms(t1, t2); times_ms[2] = ms(t2, t3); times_ms[3] = ms(t3, t4);
        times_ms[4] = ms(t4, t5); times_ms[5] = ms(t5, t6); times_ms[6] = ms(t6, t7); times_ms[7] = ms(t7, t8);
    }
    return U;
}

}  // extern "C"

// ---------------------------------------------------------------- learned radius MLP
// y = w3 . relu(W2 relu(W1 x + b1) + b2) + b3 with x = (f - mu) / sigma,
// parameters flattened as W1[64][8], b1[64], W2[64][64] (out, in), b2[64],
// w3[64], b3.  FP64 accumulation.  With emulate = 1 the operands of every
// matrix product are rounded to bf16 at the points where the GPU kernels
// round them (inputs, weights, hidden activations, dz2, dz1); with emulate = 0
// everything stays FP64 (for finite-difference checks of the formulas).
namespace {
constexpr int kMI = 8, kMH = 64;
constexpr int kPW1 = 0, kPB1 = kPW1 + kMH * kMI, kPW2 = kPB1 + kMH, kPB2 = kPW2 + kMH * kMH, kPW3 = kPB2 + kMH,
              kPB3 = kPW3 + kMH, kPN = kPB3 + 1;

inline double to_bf16(double v) {
    float f = static_cast<float>(v);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

struct MlpSample {
    double x[kMI], z1[kMH], h1[kMH], z2[kMH], h2[kMH], y;
};

void mlp_fwd_one(const float* P, const float* f, const float* mu, const float* sigma, int emu, MlpSample& s) {
    auto R = [&](double v) { return emu ? to_bf16(v) : v; };
    for (int i = 0; i < kMI; ++i) {
        if (emu) {
            float inv = 1.0f / sigma[i];
            float xn = (f[i] - mu[i]) * inv;
            s.x[i] = to_bf16(xn);
        } else {
            s.x[i] = (static_cast<double>(f[i]) - mu[i]) / sigma[i];
        }
    }
    for (int o = 0; o < kMH; ++o) {
        double a = P[kPB1 + o];
        for (int i = 0; i < kMI; ++i) a += R(P[kPW1 + o * kMI + i]) * s.x[i];
        s.z1[o] = a;
        s.h1[o] = R(a > 0 ? a : 0.0);
    }
    for (int o = 0; o < kMH; ++o) {
        double a = P[kPB2 + o];
        for (int i = 0; i < kMH; ++i) a += R(P[kPW2 + o * kMH + i]) * s.h1[i];
        s.z2[o] = a;
        s.h2[o] = R(a > 0 ? a : 0.0);
    }
    double y = P[kPB3];
    for (int i = 0; i < kMH; ++i) y += R(P[kPW3 + i]) * s.h2[i];
    s.y = y;
}
}  // namespace

extern "C" {

int orc_mlp_num_params(void) { return kPN; }

void orc_mlp_forward(const float* P, const float* feats, int64_t n, const float* mu, const float* sigma, int emulate,
                     double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q) {
        MlpSample s;
        mlp_fwd_one(P, feats + kMI * q, mu, sigma, emulate, s);
        y[q] = s.y;
    }
}

// Gradient of mean((y - t)^2) over n samples; returns the loss.
double orc_mlp_grad(const float* P, const float* feats, const float* targets, int64_t n, const float* mu,
                    const float* sigma, int emulate, double* grad) {
    auto R = [&](double v) { return emulate ? to_bf16(v) : v; };
    std::vector<double> g(kPN, 0.0);
    double loss = 0;
    for (int64_t q = 0; q < n; ++q) {
        MlpSample s;
        mlp_fwd_one(P, feats + kMI * q, mu, sigma, emulate, s);
        double e = s.y - targets[q];
        loss += e * e;
        double dy = 2.0 * e / static_cast<double>(n);
        g[kPB3] += dy;
        double dz2[kMH], dz1[kMH];
        for (int o = 0; o < kMH; ++o) {
            g[kPW3 + o] += dy * s.h2[o];
            double d = s.z2[o] > 0 ? R(P[kPW3 + o]) * dy : 0.0;
            g[kPB2 + o] += d;
            dz2[o] = R(d);
        }
        for (int o = 0; o < kMH; ++o)
            for (int i = 0; i < kMH; ++i) g[kPW2 + o * kMH + i] += dz2[o] * s.h1[i];
        for (int i = 0; i < kMH; ++i) {
            double a = 0;
            for (int o = 0; o < kMH; ++o) a += R(P[kPW2 + o * kMH + i]) * dz2[o];
            double d = s.z1[i] > 0 ? a : 0.0;
            g[kPB1 + i] += d;
            dz1[i] = R(d);
        }
        for (int o = 0; o < kMH; ++o)
            for (int i = 0; i < kMI; ++i) g[kPW1 + o * kMI + i] += dz1[o] * s.x[i];
    }
    for (int j = 0; j < kPN; ++j) grad[j] = g[j];
    return loss / static_cast<double>(n);
}

}  // extern "C"
