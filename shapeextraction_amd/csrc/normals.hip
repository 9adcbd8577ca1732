// Per-point normals from k-neighbour rows: covariance, smallest eigenvector,
// orientation toward a viewpoint.  One query per lane; fuses the reference's
// findCovariancesKernel<k>, cusolverDnSsyevjBatched and
// normalsFromEigenVectors (src/NormalEstimation.hpp:96-192,235-315) so the
// 3x3 matrices never touch HBM.
//
// Covariance (FP32): neighbours are shifted by the query point (the
// differences of nearby float coordinates are exact or nearly so, also far
// from the origin), then two passes: the mean, and the six second moments
// about it accumulated by fma.  Each entry is then within about K roundings
// of its value, an error far below what the eigensolve below tolerates before
// it hands the query to FP64 (against FP64 PCA: about 3e-6 deg median, 2e-5
// deg p99; compensated sums, which this replaced, cost 40% of the kernel's
// time).  The reference's 1/(k-1) scale is omitted: it does not change
// eigenvectors.
//
// Eigensolve (FP32): closed form for the eigenvalues (trigonometric solution
// of the characteristic cubic, O. K. Smith, CACM 4(4), 1961) and the smallest
// eigenvector as the largest cross product of two rows of (A - l_min I)
// (J. Kopp, arXiv:physics/0610206).  Its error grows like eps*l_max/(l_mid -
// l_min); when that relative gap falls below kGapTol the query is redone in
// FP64: covariance in double and cyclic Jacobi rotations (Golub & Van Loan,
// Matrix Computations, 4th ed., sec. 8.5).
//
// Orientation: n is flipped when (viewpoint - p) . n < 0
// (src/NormalEstimation.hpp:185-189).  A zero covariance (a single point)
// gives (1,0,0) before orientation, as the identity eigenbasis of syevj would.
#include <hip/hip_runtime.h>

#include "sx_internal.hpp"

namespace sx {

constexpr float kGapTol = 1e-2f;
constexpr int kNormalsBlock = 256;

struct Vec3f { float x, y, z; };

__device__ __forceinline__ Vec3f cross3(Vec3f a, Vec3f b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float norm2(Vec3f a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

// Smallest eigenvector of the symmetric matrix [a00 a01 a02; . a11 a12; . . a22].
// Returns false when the closed form cannot be trusted (relative gap between
// the two smallest eigenvalues below kGapTol, or a degenerate matrix).
__device__ bool smallest_eigvec_f32(float a00, float a01, float a02, float a11, float a12, float a22, Vec3f& n) {
    float sc = fmaxf(fmaxf(fabsf(a00), fabsf(a11)), fabsf(a22));
    sc = fmaxf(sc, fmaxf(fmaxf(fabsf(a01), fabsf(a02)), fabsf(a12)));
    if (!(sc > 0.0f) || !isfinite(sc)) return false;
    const float inv = 1.0f / sc;
    a00 *= inv; a01 *= inv; a02 *= inv; a11 *= inv; a12 *= inv; a22 *= inv;
    const float q = (a00 + a11 + a22) * (1.0f / 3.0f);
    const float b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
    const float p1 = a01 * a01 + a02 * a02 + a12 * a12;
    const float p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0f * p1;
    if (!(p2 > 1e-30f)) return false;
    const float p = sqrtf(p2 * (1.0f / 6.0f));
    const float det = b00 * (b11 * b22 - a12 * a12) - a01 * (a01 * b22 - a12 * a02) + a02 * (a01 * a12 - b11 * a02);
    float r = det / (2.0f * p * p * p);
    r = fminf(fmaxf(r, -1.0f), 1.0f);
    const float phi = acosf(r) * (1.0f / 3.0f);
    const float l0 = q + 2.0f * p * cosf(phi);
    const float l2 = q + 2.0f * p * cosf(phi + 2.0943951023931957f);
    const float l1 = 3.0f * q - l0 - l2;
    if (!(l0 > 0.0f) || (l1 - l2) < kGapTol * l0) return false;
    const Vec3f r0{a00 - l2, a01, a02}, r1{a01, a11 - l2, a12}, r2{a02, a12, a22 - l2};
    const Vec3f c01 = cross3(r0, r1), c02 = cross3(r0, r2), c12 = cross3(r1, r2);
    const float n01 = norm2(c01), n02 = norm2(c02), n12 = norm2(c12);
    Vec3f c = c01;
    float nn = n01;
    if (n02 > nn) { c = c02; nn = n02; }
    if (n12 > nn) { c = c12; nn = n12; }
    if (!(nn > 1e-30f)) return false;
    const float s = 1.0f / sqrtf(nn);
    n = {c.x * s, c.y * s, c.z * s};
    return true;
}

// Cyclic Jacobi on a symmetric 3x3 in double; returns the eigenvector of the
// smallest eigenvalue.  A zero matrix returns (1,0,0).
__device__ void smallest_eigvec_f64(double a[3][3], double out[3]) {
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 16; ++sweep) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (!(off > 1e-36 * diag)) break;
#pragma unroll
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0;
            const int qq = pq == 0 ? 1 : 2;
            const double apq = a[p][qq];
            if (apq == 0.0) continue;
            const double theta = (a[qq][qq] - a[p][p]) / (2.0 * apq);
            double t;
            if (fabs(theta) > 1e150) t = 0.5 / theta;
            else t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
            a[p][p] -= t * apq;
            a[qq][qq] += t * apq;
            a[p][qq] = a[qq][p] = 0.0;
            const int r = 3 - p - qq;
            const double arp = a[r][p], arq = a[r][qq];
            a[r][p] = a[p][r] = c * arp - s * arq;
            a[r][qq] = a[qq][r] = s * arp + c * arq;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double vkp = v[k][p], vkq = v[k][qq];
                v[k][p] = c * vkp - s * vkq;
                v[k][qq] = s * vkp + c * vkq;
            }
        }
    }
    int m = 0;
    if (a[1][1] < a[m][m]) m = 1;
    if (a[2][2] < a[m][m]) m = 2;
    out[0] = v[0][m];
    out[1] = v[1][m];
    out[2] = v[2][m];
}

template <int K>
__global__ void __launch_bounds__(kNormalsBlock) normals_kernel(const float4* __restrict__ pts,
                                                                const int32_t* __restrict__ d_counts,
                                                                const int32_t* __restrict__ nn, int64_t q_start,
                                                                int64_t n_q, int precision, float vx, float vy,
                                                                float vz, float* __restrict__ normals,
                                                                uint8_t* __restrict__ flags, CloudMap cm) {
    // XCD-contiguous block order: neighbours of a block's queries are mostly
    // the neighbouring blocks' points, read from the same L2
    const int64_t t = static_cast<int64_t>(xcd_block()) * blockDim.x + threadIdx.x;
    if (t >= n_q) return;
    const int total = d_counts[0];
    const int64_t qi = q_start + t;
    if (qi >= total) {  // past the unique count (range checked against the input count): no normal
        normals[3 * t] = normals[3 * t + 1] = normals[3 * t + 2] = __builtin_nanf("");
        if (flags) flags[t] = 0;
        return;
    }
    // rows hold indices relative to the query's cloud
    const int2 cloud = cm.at(qi, total);
    if (qi - cloud.x >= cloud.y) return;  // padding slot of a batched cloud
    const float4 q = pts[qi];
    const float4* cpts = pts + cloud.x;
    const int32_t* row = nn + t * K;
    float dx[K], dy[K], dz[K];
#pragma unroll
    for (int s = 0; s < K; s += 4) {
        const int4 id = *reinterpret_cast<const int4*>(row + s);
        const float4 p0 = cpts[id.x], p1 = cpts[id.y], p2 = cpts[id.z], p3 = cpts[id.w];
        dx[s] = p0.x - q.x; dy[s] = p0.y - q.y; dz[s] = p0.z - q.z;
        dx[s + 1] = p1.x - q.x; dy[s + 1] = p1.y - q.y; dz[s + 1] = p1.z - q.z;
        dx[s + 2] = p2.x - q.x; dy[s + 2] = p2.y - q.y; dz[s + 2] = p2.z - q.z;
        dx[s + 3] = p3.x - q.x; dy[s + 3] = p3.y - q.y; dz[s + 3] = p3.z - q.z;
    }
    Vec3f n{1.0f, 0.0f, 0.0f};
    bool done = false;
    if (precision == SX_PREC_FP32) {
        float mx = 0.0f, my = 0.0f, mz = 0.0f;
#pragma unroll
        for (int s = 0; s < K; ++s) { mx += dx[s]; my += dy[s]; mz += dz[s]; }
        const float cx = mx / K, cy = my / K, cz = mz / K;
        float xx = 0.0f, xy = 0.0f, xz = 0.0f, yy = 0.0f, yz = 0.0f, zz = 0.0f;
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const float ex = dx[s] - cx, ey = dy[s] - cy, ez = dz[s] - cz;
            xx = __builtin_fmaf(ex, ex, xx); xy = __builtin_fmaf(ex, ey, xy); xz = __builtin_fmaf(ex, ez, xz);
            yy = __builtin_fmaf(ey, ey, yy); yz = __builtin_fmaf(ey, ez, yz); zz = __builtin_fmaf(ez, ez, zz);
        }
        done = smallest_eigvec_f32(xx, xy, xz, yy, yz, zz, n);
    }
    if (!done) {
        double sx = 0, sy = 0, sz = 0;
#pragma unroll
        for (int s = 0; s < K; ++s) { sx += dx[s]; sy += dy[s]; sz += dz[s]; }
        sx /= K; sy /= K; sz /= K;
        double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const double ex = dx[s] - sx, ey = dy[s] - sy, ez = dz[s] - sz;
            a[0][0] += ex * ex; a[0][1] += ex * ey; a[0][2] += ex * ez;
            a[1][1] += ey * ey; a[1][2] += ey * ez; a[2][2] += ez * ez;
        }
        a[1][0] = a[0][1]; a[2][0] = a[0][2]; a[2][1] = a[1][2];
        double v[3];
        smallest_eigvec_f64(a, v);
        const double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        n = {static_cast<float>(v[0] / l), static_cast<float>(v[1] / l), static_cast<float>(v[2] / l)};
    }
    const float ox = __fsub_rn(vx, q.x), oy = __fsub_rn(vy, q.y), oz = __fsub_rn(vz, q.z);
    const float dot = __fadd_rn(__fadd_rn(__fmul_rn(ox, n.x), __fmul_rn(oy, n.y)), __fmul_rn(oz, n.z));
    if (dot < 0.0f) { n.x = -n.x; n.y = -n.y; n.z = -n.z; }
    normals[3 * t] = n.x;
    normals[3 * t + 1] = n.y;
    normals[3 * t + 2] = n.z;
    if (flags) flags[t] = done ? 0 : 1;
}

int normals_launch(const float4* pts, const int32_t* d_counts, const int32_t* nn, int64_t q_start, int64_t n_q, int k,
                   int precision, float vx, float vy, float vz, float* normals, uint8_t* flags, hipStream_t st,
                   CloudMap cm = CloudMap()) {
    if (n_q <= 0) return SX_OK;
    if (reinterpret_cast<uintptr_t>(nn) % 16 != 0) return fail(SX_EINVAL, "normals: neighbour rows must be 16-byte aligned");
    if (precision != SX_PREC_FP32 && precision != SX_PREC_FP64) return fail(SX_EINVAL, "normals: bad precision mode");
    const unsigned g = grid_for(n_q, kNormalsBlock);
    switch (k) {
        case 8: normals_kernel<8><<<g, kNormalsBlock, 0, st>>>(pts, d_counts, nn, q_start, n_q, precision, vx, vy, vz, normals, flags, cm); break;
        case 16: normals_kernel<16><<<g, kNormalsBlock, 0, st>>>(pts, d_counts, nn, q_start, n_q, precision, vx, vy, vz, normals, flags, cm); break;
        case 24: normals_kernel<24><<<g, kNormalsBlock, 0, st>>>(pts, d_counts, nn, q_start, n_q, precision, vx, vy, vz, normals, flags, cm); break;
        case 32: normals_kernel<32><<<g, kNormalsBlock, 0, st>>>(pts, d_counts, nn, q_start, n_q, precision, vx, vy, vz, normals, flags, cm); break;
        default: return fail(SX_EINVAL, "normals: k must be 8, 16, 24 or 32");
    }
    SX_LAUNCH_CHECK("normals_kernel");
    return SX_OK;
}

}  // namespace sx
