// Learned kNN radius: octree features + a 8-64-64-1 MLP with BF16 weights,
// FP32 accumulation, on CDNA matrix cores (v_mfma_f32_16x16x16_bf16).
//
// The reference has no learned component; this module seeds the exact kNN
// descent with a per-query radius (SX_KNN_RADIUS) and never changes its
// result (the search re-runs when fewer than k points fall in the ball).
//
// Features of query u (the point's leaf in the radix tree / octree):
//   f0..f3  log2 of the point count of the octree cells at depths D0, D0-1,
//           D0-2, D0-3 containing u, D0 = depth of u's leaf-holding cell
//   f4      log2 of that cell's edge length (log2(range) - D0)
//   f5      D0 - root depth
//   f6, f7  occupied leaf slots and child-cell slots of that octree node
// Target: ln(squared distance to the k-th neighbour).
//
// MFMA layout (16x16x16, four bf16 per lane): lane l holds
//   A[i = l%16][k = 4*(l/16)+e], B[k = 4*(l/16)+e][j = l%16],
//   D[i = 4*(l/16)+e][j = l%16], e = 0..3.
// Weights are the A operand ([out][in] row-major, 8-byte loads) and the
// activations of 16 queries the B operand, so a layer's D tile (rows = units,
// columns = queries) is already in the B layout of the next layer.
// Weight gradients sum over queries; their operands are transposed through
// LDS.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "sx_internal.hpp"

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

namespace sx {

constexpr int kIn = 8, kInPad = 16, kH = 64;
// flattened FP32 parameter vector
constexpr int kOffW1 = 0;                     // [64][8]
constexpr int kOffB1 = kOffW1 + kH * kIn;     // [64]
constexpr int kOffW2 = kOffB1 + kH;           // [64][64] (out, in)
constexpr int kOffB2 = kOffW2 + kH * kH;      // [64]
constexpr int kOffW3 = kOffB2 + kH;           // [64]
constexpr int kOffB3 = kOffW3 + kH;           // [1]
constexpr int kNParams = kOffB3 + 1;          // 4801
constexpr int kMlpBlock = 256;                // 4 waves, 64 queries per wave
constexpr int kWavesPerBlock = kMlpBlock / 64;

// Round to nearest even: v_cvt_pk_bf16_f32 on gfx950, the same bits as
// adding 0x7FFF + the kept LSB and truncating, for the finite values here.
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f)); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ bf16x4_t pack4(float a, float b, float c, float d) {
    bf16x4_t v;
    v[0] = static_cast<short>(f2bf(a));
    v[1] = static_cast<short>(f2bf(b));
    v[2] = static_cast<short>(f2bf(c));
    v[3] = static_cast<short>(f2bf(d));
    return v;
}
__device__ __forceinline__ bf16x4_t relu_pack(f32x4_t z) {
    return pack4(fmaxf(z[0], 0.f), fmaxf(z[1], 0.f), fmaxf(z[2], 0.f), fmaxf(z[3], 0.f));
}
__device__ __forceinline__ bf16x4_t load4(const uint16_t* p) {
    const uint2 w = *reinterpret_cast<const uint2*>(p);
    bf16x4_t v;
    v[0] = static_cast<short>(w.x & 0xFFFF);
    v[1] = static_cast<short>(w.x >> 16);
    v[2] = static_cast<short>(w.y & 0xFFFF);
    v[3] = static_cast<short>(w.y >> 16);
    return v;
}
__device__ __forceinline__ f32x4_t mfma(bf16x4_t a, bf16x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct MlpLds;

// Device-side weights: BF16 copies laid out for the A operand.
struct MlpDev {
    const uint16_t* w1p;  // [64][16]  W1 padded with zero inputs 8..15
    const uint16_t* w2;   // [64][64]  W2[out][in]
    const uint16_t* w2t;  // [64][64]  W2T[in][out]
    const uint16_t* w3;   // [64]
    const float* p;       // FP32 master parameters (biases are read from here)
    const float* norm;    // mu[8], inv_sigma[8], margin
    const MlpLds* img;    // inference operands in their LDS layout (pack_image_kernel)
};

struct Fwd {
    f32x4_t z1[4], z2[4];
    bf16x4_t h1[4], h2[4];
    float y;  // valid in lanes 0..15: query l
};

// Forward of one tile of 16 queries; xb is the layer-1 B operand.
__device__ __forceinline__ void mlp_tile_forward(const MlpDev& m, bf16x4_t xb, Fwd& f) {
    const int l = static_cast<int>(lane_id());
    const int r = l % 16, g = l / 16;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4_t c;
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = m.p[kOffB1 + 16 * mt + 4 * g + e];
        const bf16x4_t a = load4(m.w1p + (16 * mt + r) * kInPad + 4 * g);
        f.z1[mt] = mfma(a, xb, c);
        f.h1[mt] = relu_pack(f.z1[mt]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4_t c;
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = m.p[kOffB2 + 16 * mt + 4 * g + e];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) c = mfma(load4(m.w2 + (16 * mt + r) * kH + 16 * kc + 4 * g), f.h1[kc], c);
        f.z2[mt] = c;
        f.h2[mt] = relu_pack(c);
    }
    f32x4_t c = {0.f, 0.f, 0.f, 0.f};
    if (g == 0) c[0] = m.p[kOffB3];
    const bf16x4_t zero4 = {0, 0, 0, 0};
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
        const bf16x4_t a = (r == 0) ? load4(m.w3 + 16 * kc + 4 * g) : zero4;
        c = mfma(a, f.h2[kc], c);
    }
    f.y = c[0];
}

// Layer-1 B operand for the training tile starting at query q0: lane l holds
// the normalised features 4*(l/16)+e of query q0 + l%16 (zero padding for
// l >= 32 and for queries >= n).
__device__ __forceinline__ bf16x4_t load_x(const MlpDev& m, const float* __restrict__ feats, int64_t q0, int64_t n) {
    const int l = static_cast<int>(lane_id());
    const int r = l % 16, g = l / 16;
    const int64_t q = q0 + r;
    if (g >= 2 || q >= n) return bf16x4_t{0, 0, 0, 0};
    const float4 v = *reinterpret_cast<const float4*>(feats + q * kIn + 4 * g);
    const float* mu = m.norm;
    const float* is = m.norm + kIn;
    return pack4((v.x - mu[4 * g]) * is[4 * g], (v.y - mu[4 * g + 1]) * is[4 * g + 1],
                 (v.z - mu[4 * g + 2]) * is[4 * g + 2], (v.w - mu[4 * g + 3]) * is[4 * g + 3]);
}

// ---------------------------------------------------------------- inference
// 64 queries (four 16-query tiles) per pass, so each weight operand is read
// from LDS once per 64 queries.  Layer 1 (8 -> 64) on v_mfma_f32_16x16x16_bf16
// (inputs padded to 16); layers 2 (64 -> 64) and 3 (64 -> 1) on gfx950's
// v_mfma_f32_16x16x32_bf16: lane l holds A[i = l%16][k = 8*(l/16)+e],
// B[k = 8*(l/16)+e][j = l%16], e = 0..7, and D as for 16x16x16.  The K index
// of a 32-unit chunk kc is permuted so that the D tiles 2kc and 2kc+1 of the
// previous layer, concatenated per lane, are already its B operand:
//   k = 8g + e  <->  hidden unit 32kc + 16*(e/4) + 4g + e%4,
// and the A operands (weights) are stored in LDS in that order, lane-major,
// so a wave's operand read is one contiguous block.  ReLU is taken after the
// BF16 rounding, on the packed halves as int16 (v_pk_max_i16): a negative
// BF16 value is a negative int16, so this equals rounding max(z, 0).
typedef short bf16x8s_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4_t mfma32(bf16x8s_t a, bf16x8s_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8s_t cat8(bf16x4_t a, bf16x4_t b) {
    return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
}
__device__ __forceinline__ bf16x4_t round_relu(f32x4_t z) {
    const bf16x4_t zero4 = {0, 0, 0, 0};
    return __builtin_elementwise_max(pack4(z[0], z[1], z[2], z[3]), zero4);
}
// Hidden unit of (chunk kc, lane group g, element e) in the permuted K order.
__host__ __device__ constexpr int perm_unit(int kc, int g, int e) { return 32 * kc + 16 * (e / 4) + 4 * g + e % 4; }

struct alignas(16) MlpLds {
    uint16_t w1x[4 * kWave * 4];  // layer 1, output tile mt: lane l's four BF16 at (mt*64 + l)*4
    uint16_t w2x[8 * kWave * 8];  // layer 2, (output tile mt, chunk kc): lane l's eight at ((mt*2+kc)*64 + l)*8
    uint16_t w3x[2 * kWave * 8];  // layer 3, chunk kc: row 0 only (zero for l%16 != 0)
    float b1[kH], b2[kH], b3;
    float mu[kIn], is[kIn], margin;
};

// The weights in the operand layouts (block-wide; sw may be global memory).
__device__ __forceinline__ void build_operands(const MlpDev& m, MlpLds& sw) {
    const int tid = static_cast<int>(threadIdx.x);
    for (int i = tid; i < 4 * kWave * 4; i += blockDim.x) {
        const int e = i % 4, l = (i / 4) % kWave, mt = i / (4 * kWave);
        sw.w1x[i] = m.w1p[(16 * mt + l % 16) * kInPad + 4 * (l / 16) + e];
    }
    for (int i = tid; i < 8 * kWave * 8; i += blockDim.x) {
        const int e = i % 8, l = (i / 8) % kWave, c = i / (8 * kWave), mt = c / 2, kc = c % 2;
        sw.w2x[i] = m.w2[(16 * mt + l % 16) * kH + perm_unit(kc, l / 16, e)];
    }
    for (int i = tid; i < 2 * kWave * 8; i += blockDim.x) {
        const int e = i % 8, l = (i / 8) % kWave, kc = i / (8 * kWave);
        sw.w3x[i] = (l % 16 == 0) ? m.w3[perm_unit(kc, l / 16, e)] : static_cast<uint16_t>(0);
    }
    for (int i = tid; i < kH; i += blockDim.x) {
        sw.b1[i] = m.p[kOffB1 + i];
        sw.b2[i] = m.p[kOffB2 + i];
    }
    for (int i = tid; i < kIn; i += blockDim.x) {
        sw.mu[i] = m.norm[i];
        sw.is[i] = m.norm[kIn + i];
    }
    if (tid == 0) {
        sw.b3 = m.p[kOffB3];
        sw.margin = m.norm[2 * kIn];
    }
}

// One block: the image every inference block copies into LDS, rebuilt
// whenever the parameters or the normaliser change.
__global__ void __launch_bounds__(256) pack_image_kernel(MlpDev m, MlpLds* img) { build_operands(m, *img); }

// Block-wide copy of the packed image into LDS (ends with a barrier).
__device__ __forceinline__ void stage_weights(const MlpDev& m, MlpLds& sw) {
    static_assert(sizeof(MlpLds) % 16 == 0, "stage_weights: 16-byte copies");
    const uint4* src = reinterpret_cast<const uint4*>(m.img);
    uint4* dst = reinterpret_cast<uint4*>(&sw);
    for (int i = static_cast<int>(threadIdx.x); i < static_cast<int>(sizeof(MlpLds) / 16); i += blockDim.x)
        dst[i] = src[i];
    __syncthreads();
}

__device__ __forceinline__ bf16x8s_t lds8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8s_t, *reinterpret_cast<const uint4*>(p));
}

// Layer-1 B operand of one tile: lane l holds the normalised features
// 4*(l/16)+e of its query (v = those four raw features; lanes l >= 32 and
// absent queries pass present = false and get zeros).
__device__ __forceinline__ bf16x4_t feature_operand(const MlpLds& sw, float4 v, bool present) {
    const int g = static_cast<int>(lane_id()) / 16;
    if (!present || g >= 2) return bf16x4_t{0, 0, 0, 0};
    const int o = 4 * g;
    return pack4((v.x - sw.mu[o]) * sw.is[o], (v.y - sw.mu[o + 1]) * sw.is[o + 1], (v.z - sw.mu[o + 2]) * sw.is[o + 2],
                 (v.w - sw.mu[o + 3]) * sw.is[o + 3]);
}

// y of the four tiles: y[t][0] in lanes 0..15 is the output of query 16t + l.
// The weights are re-read from LDS per call: the empty asm hides that the
// address is loop-invariant, so the compiler does not hoist every operand
// into registers (which would cost the kernel half its occupancy).
__device__ __forceinline__ void mlp_forward64(const MlpLds& s_w, const bf16x4_t (&xb)[4], f32x4_t (&y)[4]) {
    int zo = 0;
    asm volatile("" : "+v"(zo));
    const MlpLds& sw = *reinterpret_cast<const MlpLds*>(reinterpret_cast<const char*>(&s_w) + zo);
    const int l = static_cast<int>(lane_id());
    const int g = l / 16;
    bf16x4_t h1[4][4];  // [tile][output tile]
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const bf16x4_t a = load4(sw.w1x + (mt * kWave + l) * 4);
        const f32x4_t b = *reinterpret_cast<const f32x4_t*>(sw.b1 + 16 * mt + 4 * g);
#pragma unroll
        for (int t = 0; t < 4; ++t) h1[t][mt] = round_relu(mfma(a, xb[t], b));
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) y[t] = f32x4_t{g == 0 ? sw.b3 : 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {  // layer-2 output tiles 2kc, 2kc+1 = layer-3 chunk kc
        bf16x4_t h2[4][2];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int mt = 2 * kc + half;
            const f32x4_t b = *reinterpret_cast<const f32x4_t*>(sw.b2 + 16 * mt + 4 * g);
            f32x4_t c[4] = {b, b, b, b};
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                const bf16x8s_t a = lds8(sw.w2x + ((mt * 2 + kk) * kWave + l) * 8);
#pragma unroll
                for (int t = 0; t < 4; ++t) c[t] = mfma32(a, cat8(h1[t][2 * kk], h1[t][2 * kk + 1]), c[t]);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) h2[t][half] = round_relu(c[t]);
        }
        const bf16x8s_t a3 = lds8(sw.w3x + (kc * kWave + l) * 8);
#pragma unroll
        for (int t = 0; t < 4; ++t) y[t] = mfma32(a3, cat8(h2[t][0], h2[t][1]), y[t]);
    }
}

// mode 0: y (raw prediction); mode 1: exp(y + margin) (squared radius).
// Persistent: each wave walks groups of 64 queries.
__global__ void __launch_bounds__(kMlpBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) mlp_forward_kernel(MlpDev m, const float* __restrict__ feats, int64_t n,
                                                                int mode, float* __restrict__ out) {
    __shared__ MlpLds s_w;
    stage_weights(m, s_w);
    const int w = static_cast<int>(threadIdx.x) / kWave;
    const int l = static_cast<int>(lane_id());
    const int r = l % 16, g = l / 16;
    const int64_t groups = (n + kWave - 1) / kWave;
    for (int64_t grp = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w; grp < groups;
         grp += static_cast<int64_t>(gridDim.x) * kWavesPerBlock) {
        const int64_t q0 = grp * kWave;
        bf16x4_t xb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t q = q0 + 16 * t + r;
            const bool present = g < 2 && q < n;
            const float4 v = present ? *reinterpret_cast<const float4*>(feats + q * kIn + 4 * g)
                                     : make_float4(0.f, 0.f, 0.f, 0.f);
            xb[t] = feature_operand(s_w, v, present);
        }
        f32x4_t y[4];
        mlp_forward64(s_w, xb, y);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t q = q0 + 16 * t + l;
            if (l < 16 && q < n) out[q] = mode == 0 ? y[t][0] : __expf(y[t][0] + s_w.margin);
        }
    }
}

// Gradient of mean((y - t)^2) over n samples.  Each wave handles 64 samples
// (4 tiles) and writes its partial gradient (kNParams floats) to
// partial[wave]; loss_partial[wave] receives its sum of squared errors.
__global__ void __launch_bounds__(kMlpBlock) mlp_grad_kernel(MlpDev m, const float* __restrict__ feats,
                                                             const float* __restrict__ target, int64_t n,
                                                             float inv_n, float* __restrict__ partial,
                                                             float* __restrict__ loss_partial) {
    __shared__ uint16_t s_a[kMlpBlock / kWave][kH][16];  // dz2 or dz1 tile, [unit][query]
    __shared__ uint16_t s_b[kMlpBlock / kWave][kH][16];  // h1 tile, [unit][query]
    __shared__ uint16_t s_x[kMlpBlock / kWave][16][16];  // x tile, [feature][query]
    const int wl = static_cast<int>(threadIdx.x) / kWave;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int l = static_cast<int>(lane_id());
    const int r = l % 16, g = l / 16;
    f32x4_t gW2[4][4], gW1[4];
    float gb1[4][4], gb2[4][4], gw3[4][4];
    float gb3 = 0.f, sse = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gW1[a] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 4; ++b) { gW2[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f}; gb1[a][b] = gb2[a][b] = gw3[a][b] = 0.f; }
    }
#pragma unroll 1
    for (int t = 0; t < 4; ++t) {
        const int64_t q0 = wave * 64 + 16 * t;
        const bf16x4_t xb = load_x(m, feats, q0, n);
        Fwd f;
        mlp_tile_forward(m, xb, f);
        // dL/dy for query q0 + r, broadcast to every lane of column r
        float dy = 0.f;
        if (l < 16 && q0 + l < n) {
            const float e = f.y - target[q0 + l];
            dy = 2.0f * e * inv_n;
            sse += e * e;
        }
        const float dyq = __shfl(dy, r, kWave);
        gb3 += dy;
        // layer 3 and dz2
        bf16x4_t dz2b[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            float dz[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int o = 16 * mt + 4 * g + e;
                gw3[mt][e] += dyq * bf2f(static_cast<uint16_t>(f.h2[mt][e]));
                dz[e] = f.z2[mt][e] > 0.f ? bf2f(m.w3[o]) * dyq : 0.f;
                gb2[mt][e] += dz[e];
            }
            dz2b[mt] = pack4(dz[0], dz[1], dz[2], dz[3]);
        }
        // dh1 = W2^T dz2, then dz1
        bf16x4_t dz1b[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) c = mfma(load4(m.w2t + (16 * mt + r) * kH + 16 * kc + 4 * g), dz2b[kc], c);
            float dz[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                dz[e] = f.z1[mt][e] > 0.f ? c[e] : 0.f;
                gb1[mt][e] += dz[e];
            }
            dz1b[mt] = pack4(dz[0], dz[1], dz[2], dz[3]);
        }
        // weight gradients: transpose operands through LDS ([unit][query])
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                s_a[wl][16 * mt + 4 * g + e][r] = static_cast<uint16_t>(dz2b[mt][e]);
                s_b[wl][16 * mt + 4 * g + e][r] = static_cast<uint16_t>(f.h1[mt][e]);
            }
#pragma unroll
        for (int e = 0; e < 4; ++e) s_x[wl][4 * g + e][r] = static_cast<uint16_t>(xb[e]);
        __syncthreads();
        // dW2[o][i] += sum_q dz2[o][q] h1[i][q]: A = dz2 (o x q), B = h1^T (q x i)
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) {
            const bf16x4_t a = load4(&s_a[wl][16 * mo + r][4 * g]);
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) gW2[mo][mi] = mfma(a, load4(&s_b[wl][16 * mi + r][4 * g]), gW2[mo][mi]);
        }
        __syncthreads();
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int e = 0; e < 4; ++e) s_a[wl][16 * mt + 4 * g + e][r] = static_cast<uint16_t>(dz1b[mt][e]);
        __syncthreads();
        // dW1[o][f] += sum_q dz1[o][q] x[f][q]: A = dz1 (o x q), B = x^T (q x f)
        const bf16x4_t bx = load4(&s_x[wl][r][4 * g]);
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) gW1[mo] = mfma(load4(&s_a[wl][16 * mo + r][4 * g]), bx, gW1[mo]);
        __syncthreads();
    }
    // reduce the per-lane bias and w3 partial sums over the 16 query columns
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                gb1[a][b] += __shfl_xor(gb1[a][b], o, kWave);
                gb2[a][b] += __shfl_xor(gb2[a][b], o, kWave);
                gw3[a][b] += __shfl_xor(gw3[a][b], o, kWave);
            }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        gb3 += __shfl_xor(gb3, o, kWave);
        sse += __shfl_xor(sse, o, kWave);
    }
    float* P = partial + wave * kNParams;
#pragma unroll
    for (int mo = 0; mo < 4; ++mo)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int o = 16 * mo + 4 * g + e;
#pragma unroll
            for (int mi = 0; mi < 4; ++mi) P[kOffW2 + o * kH + 16 * mi + r] = gW2[mo][mi][e];
            if (r < kIn) P[kOffW1 + o * kIn + r] = gW1[mo][e];
            if (r == 0) {
                P[kOffB1 + o] = gb1[mo][e];
                P[kOffB2 + o] = gb2[mo][e];
                P[kOffW3 + o] = gw3[mo][e];
            }
        }
    if (l == 0) {
        P[kOffB3] = gb3;
        loss_partial[wave] = sse;
    }
}

__global__ void mlp_reduce_kernel(const float* __restrict__ partial, int64_t waves, float* __restrict__ grad) {
    const int j = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= kNParams) return;
    float s = 0.f;
    for (int64_t w = 0; w < waves; ++w) s += partial[w * kNParams + j];
    grad[j] = s;
}

__global__ void mlp_sum_kernel(const float* __restrict__ v, int64_t n, float* __restrict__ out) {
    float s = 0.f;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    __shared__ float ws[4];
    if (lane_id() == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = ws[0] + ws[1] + ws[2] + ws[3];
}

// Writes the BF16 operand copies of parameter j (padding is zero from init).
__device__ __forceinline__ void write_bf16(int j, float v, uint16_t* w1p, uint16_t* w2, uint16_t* w2t, uint16_t* w3) {
    if (j < kOffB1) {
        const int o = j / kIn, f = j % kIn;
        w1p[o * kInPad + f] = f2bf(v);
    } else if (j >= kOffW2 && j < kOffB2) {
        const int o = (j - kOffW2) / kH, i = (j - kOffW2) % kH;
        w2[o * kH + i] = f2bf(v);
        w2t[i * kH + o] = f2bf(v);
    } else if (j >= kOffW3 && j < kOffB3) {
        w3[j - kOffW3] = f2bf(v);
    }
}

__global__ void mlp_sync_bf16_kernel(const float* __restrict__ p, uint16_t* w1p, uint16_t* w2, uint16_t* w2t,
                                     uint16_t* w3) {
    const int j = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < kH * kInPad) w1p[j] = 0;
    __syncthreads();
    if (j < kNParams) write_bf16(j, p[j], w1p, w2, w2t, w3);
}

__global__ void mlp_adam_kernel(float* __restrict__ p, float* __restrict__ mo, float* __restrict__ ve,
                                const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1,
                                float c2, float wd, uint16_t* w1p, uint16_t* w2, uint16_t* w2t, uint16_t* w3) {
    const int j = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= kNParams) return;
    const float gj = grad[j] + wd * p[j];
    const float m1 = b1 * mo[j] + (1.f - b1) * gj;
    const float v1 = b2 * ve[j] + (1.f - b2) * gj * gj;
    mo[j] = m1;
    ve[j] = v1;
    const float pj = p[j] - lr * (m1 * c1) / (sqrtf(v1 * c2) + eps);
    p[j] = pj;
    write_bf16(j, pj, w1p, w2, w2t, w3);
}

// ---------------------------------------------------------------- features
// Textbook root (the octree is built on the traversal tree): cell counts come
// from each octree node's (parent, point count), one dependent load per level,
// and depths from the prefix-length array.  Reference root: the walk climbs
// the traversal tree, one int2 (parent, prefix length) per radix node, about
// three steps per octree level.
struct FeatTree {
    const TravNode* trav;
    const int2* up;      // reference root
    const int2* aux;     // textbook root
    const uint8_t* pn;   // textbook root
    int64_t oct_cap;     // octree links are followed only inside the octree's capacity
    __device__ __forceinline__ int prefix(int node) const { return aux ? pn[node] : up[node].y; }
};

// Features of index point u (U >= 2 points in the index).
__device__ __forceinline__ void query_features(const FeatTree& tr, const int32_t* __restrict__ leaf_parent,
                                               const int32_t* __restrict__ leaf_oct,
                                               const sx_otnode* __restrict__ oct, float log2_range, int root_level,
                                               int64_t u, float (&f)[kIn]) {
    int c = leaf_parent[u] >> 1;
    const int d0 = tr.prefix(c) / 3;
    const int on = leaf_oct[u];
    if (tr.aux) {
        // the octree node holding u is the cell at depth d0; its ancestors are
        // the cells at d0-1, d0-2, ... up to the root, which repeats
        int2 a = (on >= 0 && on < tr.oct_cap) ? tr.aux[on] : make_int2(-1, 1);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            f[s] = __log2f(static_cast<float>(a.y));
            if (s < 3 && a.x >= 0 && a.x < tr.oct_cap) a = tr.aux[a.x];
        }
    } else {
        int pc = tr.up[c].x;
        int2 upc = pc >= 0 ? tr.up[pc] : make_int2(-1, -1);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int L = max(d0 - s, root_level);
            // highest ancestor whose prefix still holds 3L bits
            while (pc >= 0 && upc.y >= 3 * L) {
                c = pc;
                pc = upc.x;
                upc = pc >= 0 ? tr.up[pc] : make_int2(-1, -1);
            }
            const int cnt = w_int(tr.trav[c].lo_r.w) - w_int(tr.trav[c].lo_l.w) + 1;
            f[s] = __log2f(static_cast<float>(cnt));
        }
    }
    f[4] = log2_range - static_cast<float>(d0);
    f[5] = static_cast<float>(d0 - root_level);
    if (on >= 0 && on < tr.oct_cap) {
        f[6] = static_cast<float>(__popc(oct[on].child_leaf_mask));
        f[7] = static_cast<float>(__popc(oct[on].child_node_mask));
    }
}

// The cloud of point u: its first slot s0, unique points U, log2 of its cube
// range.  One cloud (kBatch false): slot 0, the point count, bounds[2]; a
// batch: cloud c = the cloud of u's 64-slot tile, cube at bounds[4c].
template <bool kBatch>
__device__ __forceinline__ void feature_cloud(const CloudMap& cm, const float* __restrict__ bounds, int total,
                                              int64_t u, int64_t& s0, int& U, float& log2_range) {
    if (kBatch) {
        const int cid = cm.tile_cloud[u / kCloudAlign];
        s0 = cm.base[cid];
        U = cm.size[cid];
        log2_range = __log2f(bounds[4 * cid + 2]);
    } else {
        s0 = 0;
        U = total;
        log2_range = __log2f(bounds[2]);
    }
}

template <bool kBatch>
__global__ void __launch_bounds__(256) features_kernel(FeatTree tr, const int32_t* __restrict__ leaf_parent,
                                                       const int32_t* __restrict__ leaf_oct,
                                                       const sx_otnode* __restrict__ oct,
                                                       const float* __restrict__ bounds,
                                                       const int32_t* __restrict__ d_counts, int64_t q_start,
                                                       int64_t n_q, float* __restrict__ feats, CloudMap cm) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_q) return;
    const int total = d_counts[0];
    const int64_t u = q_start + t;
    float f[kIn] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (u < total) {
        int64_t s0;
        int U;
        float log2_range;
        feature_cloud<kBatch>(cm, bounds, total, u, s0, U, log2_range);
        if (u - s0 < U && U >= 2) query_features(tr, leaf_parent, leaf_oct, oct, log2_range, tr.prefix(s0) / 3, u, f);
    }
    float4* out = reinterpret_cast<float4*>(feats + t * kIn);
    out[0] = make_float4(f[0], f[1], f[2], f[3]);
    out[1] = make_float4(f[4], f[5], f[6], f[7]);
}

// Features and the MLP in one pass: each wave computes the features of its
// 64 queries (one per lane), stages them in LDS and runs mlp_forward64 on
// them, r2 = exp(y + margin): the same arithmetic as features_kernel followed
// by mlp_forward_kernel (mode 1), without the 32-byte feature rows in HBM.
template <bool kBatch>
__global__ void __launch_bounds__(kMlpBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) radius_kernel(FeatTree tr, const int32_t* __restrict__ leaf_parent,
                                                           const int32_t* __restrict__ leaf_oct,
                                                           const sx_otnode* __restrict__ oct,
                                                           const float* __restrict__ bounds,
                                                           const int32_t* __restrict__ d_counts, int64_t q_start,
                                                           int64_t n_q, MlpDev m, float* __restrict__ r2,
                                                           CloudMap cm) {
    __shared__ MlpLds s_w;
    __shared__ float4 s_f[kWavesPerBlock][2][kWave];  // features 0-3 and 4-7 of the wave's queries
    stage_weights(m, s_w);
    const int w = static_cast<int>(threadIdx.x) / kWave;
    const int l = static_cast<int>(lane_id());
    const int r = l % 16, g = l / 16;
    const int U = d_counts[0];  // a batch: the slot count
    const float log2_range = kBatch ? 0.0f : __log2f(bounds[2]);
    const int root_level = !kBatch && U >= 2 ? tr.prefix(0) / 3 : 0;
    // persistent: each wave walks groups of 64 queries (no block barriers below)
    const int64_t groups = (n_q + kWave - 1) / kWave;
    for (int64_t grp = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + w; grp < groups;
         grp += static_cast<int64_t>(gridDim.x) * kWavesPerBlock) {
        const int64_t q0 = grp * kWave;
        {
            const int64_t u = q_start + q0 + l;
            float f[kIn] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (kBatch) {
                if (q0 + l < n_q && u < U) {
                    int64_t s0;
                    int Uc;
                    float lr;
                    feature_cloud<true>(cm, bounds, U, u, s0, Uc, lr);
                    if (u - s0 < Uc && Uc >= 2)
                        query_features(tr, leaf_parent, leaf_oct, oct, lr, tr.prefix(s0) / 3, u, f);
                }
            } else if (q0 + l < n_q && u < U && U >= 2) {
                query_features(tr, leaf_parent, leaf_oct, oct, log2_range, root_level, u, f);
            }
            s_f[w][0][l] = make_float4(f[0], f[1], f[2], f[3]);
            s_f[w][1][l] = make_float4(f[4], f[5], f[6], f[7]);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        bf16x4_t xb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bool present = g < 2 && q0 + 16 * t + r < n_q;
            xb[t] = feature_operand(s_w, present ? s_f[w][g & 1][16 * t + r] : make_float4(0.f, 0.f, 0.f, 0.f),
                                    present);
        }
        f32x4_t y[4];
        mlp_forward64(s_w, xb, y);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t q = q0 + 16 * t + l;
            if (l < 16 && q < n_q) r2[q] = __expf(y[t][0] + s_w.margin);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // s_f is refilled by the next group
    }
}

}  // namespace sx

// ---------------------------------------------------------------- C ABI
struct sx_mlp {
    void* mem = nullptr;
    float* p = nullptr;      // master params
    float* mo = nullptr;     // Adam first moment
    float* ve = nullptr;     // Adam second moment
    float* grad = nullptr;
    float* norm = nullptr;   // mu[8], inv_sigma[8], margin
    float* loss = nullptr;   // [1]
    uint16_t* w1p = nullptr;
    uint16_t* w2 = nullptr;
    uint16_t* w2t = nullptr;
    uint16_t* w3 = nullptr;
    sx::MlpLds* img = nullptr;  // inference operands, packed
    float* partial = nullptr;
    float* loss_partial = nullptr;
    int64_t max_batch = 0;
    int64_t step = 0;
};

namespace {

size_t mlp_layout(sx_mlp* m, void* base) {
    sx::Carver c(base);
    m->p = c.take<float>(sx::kNParams);
    m->mo = c.take<float>(sx::kNParams);
    m->ve = c.take<float>(sx::kNParams);
    m->grad = c.take<float>(sx::kNParams);
    m->norm = c.take<float>(2 * sx::kIn + 1);
    m->loss = c.take<float>(1);
    m->w1p = c.take<uint16_t>(sx::kH * sx::kInPad);
    m->w2 = c.take<uint16_t>(sx::kH * sx::kH);
    m->w2t = c.take<uint16_t>(sx::kH * sx::kH);
    m->w3 = c.take<uint16_t>(sx::kH);
    m->img = c.take<sx::MlpLds>(1);
    // whole blocks of waves: every wave of the last block writes its (zero) partial
    const int64_t waves = ((m->max_batch + 63) / 64 + sx::kWavesPerBlock - 1) / sx::kWavesPerBlock * sx::kWavesPerBlock;
    m->partial = c.take<float>(static_cast<size_t>(waves) * sx::kNParams);
    m->loss_partial = c.take<float>(static_cast<size_t>(waves));
    return c.off;
}

sx::MlpDev dev_view(const sx_mlp* m) { return sx::MlpDev{m->w1p, m->w2, m->w2t, m->w3, m->p, m->norm, m->img}; }

// Rebuild the packed inference operands (after any change of the BF16
// weights, biases or normaliser; stream-ordered after that change).
int pack_image(sx_mlp* m, hipStream_t st) {
    sx::pack_image_kernel<<<1, 256, 0, st>>>(dev_view(m), m->img);
    SX_LAUNCH_CHECK("pack_image_kernel");
    return SX_OK;
}

// Grid of a persistent inference kernel (64 queries per wave and step): the
// blocks the queries need, at most the blocks the device holds at once (a
// block that starts after the first ones finish would double the tail).  The
// occupancy is queried once per device and kernel.
std::atomic<int64_t> g_resident_forward[64], g_resident_radius[64], g_resident_radius_batch[64];

template <typename Kern>
int persistent_grid(Kern kernel, std::atomic<int64_t>* cache, int64_t n, unsigned* grid) {
    const int64_t groups = (n + sx::kWave - 1) / sx::kWave;
    int64_t blocks = (groups + sx::kWavesPerBlock - 1) / sx::kWavesPerBlock;
    int dev = 0;
    SX_HIP(hipGetDevice(&dev));
    int64_t resident = dev >= 0 && dev < 64 ? cache[dev].load(std::memory_order_relaxed) : 0;
    if (resident <= 0) {
        int cus = 0, per_cu = 0;
        SX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        SX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, sx::kMlpBlock, 0));
        resident = static_cast<int64_t>(cus > 0 ? cus : 1) * (per_cu > 0 ? per_cu : 1);
        if (dev >= 0 && dev < 64) cache[dev].store(resident, std::memory_order_relaxed);
    }
    if (blocks > resident) blocks = resident;
    *grid = static_cast<unsigned>(blocks < 1 ? 1 : blocks);
    return SX_OK;
}

int check_mlp(const sx_mlp* m) { return m ? SX_OK : sx::fail(SX_EINVAL, "null mlp"); }

}  // namespace

// sx_index / sx_batch internals needed by the feature kernels
struct sx_index;
struct sx_batch;
namespace sx {
int index_feature_arrays(sx_index* ix, const TravNode** trav, const int2** up, const int2** oct_aux,
                         const uint8_t** pn, const int32_t** leaf_parent, const int32_t** leaf_oct,
                         const sx_otnode** oct, int64_t* oct_cap, const float** bounds, const int32_t** counts,
                         int64_t* limit, void* stream);
// the same for a batch built with per-cloud octrees (SX_ESTATE without them)
int batch_feature_arrays(sx_batch* b, BatchFeatures* out, void* stream);
}

using sx::fail;

extern "C" {

int sx_mlp_create(sx_mlp** out, int64_t max_batch) {
    if (!out) return fail(SX_EINVAL, "sx_mlp_create: null out");
    *out = nullptr;
    if (max_batch <= 0 || max_batch > (1ll << 26)) return fail(SX_EINVAL, "sx_mlp_create: max_batch out of range");
    sx_mlp* m = new sx_mlp();
    m->max_batch = max_batch;
    const size_t bytes = mlp_layout(m, nullptr);
    if (hipMalloc(&m->mem, bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete m;
        return fail(SX_ENOMEM, "sx_mlp_create: allocation failed");
    }
    mlp_layout(m, m->mem);
    if (sx::fill_u8(m->mem, 0, bytes - bytes % 16, nullptr) != SX_OK || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(m->mem);
        delete m;
        return fail(SX_EHIP, "sx_mlp_create: initialisation failed");
    }
    *out = m;
    return SX_OK;
}

int sx_mlp_destroy(sx_mlp* m) {
    if (!m) return SX_OK;
    int rc = SX_OK;
    if (m->mem && hipFree(m->mem) != hipSuccess) rc = fail(SX_EHIP, "sx_mlp_destroy: hipFree failed");
    delete m;
    return rc;
}

int sx_mlp_num_params(void) { return sx::kNParams; }

// Host FP32 parameters in the flattened order W1[64][8], b1, W2[64][64], b2,
// w3[64], b3; resets the Adam state.
int sx_mlp_set_params(sx_mlp* m, const float* h_params, void* stream) {
    SX_TRY(check_mlp(m));
    if (!h_params) return fail(SX_EINVAL, "sx_mlp_set_params: null params");
    hipStream_t st = sx::as_stream(stream);
    SX_HIP(hipMemcpyAsync(m->p, h_params, sizeof(float) * sx::kNParams, hipMemcpyHostToDevice, st));
    SX_TRY(sx::fill_u8(m->mo, 0, sizeof(float) * sx::kNParams, st));
    SX_TRY(sx::fill_u8(m->ve, 0, sizeof(float) * sx::kNParams, st));
    sx::mlp_sync_bf16_kernel<<<(sx::kNParams + 1023) / 1024 * 1, 1024, 0, st>>>(m->p, m->w1p, m->w2, m->w2t, m->w3);
    SX_LAUNCH_CHECK("mlp_sync_bf16_kernel");
    SX_TRY(pack_image(m, st));
    SX_HIP(hipStreamSynchronize(st));
    m->step = 0;
    return SX_OK;
}

int sx_mlp_get_params(sx_mlp* m, float* h_params, void* stream) {
    SX_TRY(check_mlp(m));
    if (!h_params) return fail(SX_EINVAL, "sx_mlp_get_params: null params");
    hipStream_t st = sx::as_stream(stream);
    SX_HIP(hipMemcpyAsync(h_params, m->p, sizeof(float) * sx::kNParams, hipMemcpyDeviceToHost, st));
    SX_HIP(hipStreamSynchronize(st));
    return SX_OK;
}

// Glorot-uniform weights, zero biases, from a host RNG seeded with `seed`.
int sx_mlp_init(sx_mlp* m, uint64_t seed, void* stream) {
    SX_TRY(check_mlp(m));
    std::vector<float> p(sx::kNParams, 0.f);
    std::mt19937_64 gen(seed);
    auto fill = [&](int off, int fan_in, int fan_out, int count) {
        const float a = std::sqrt(6.0f / static_cast<float>(fan_in + fan_out));
        std::uniform_real_distribution<float> d(-a, a);
        for (int i = 0; i < count; ++i) p[off + i] = d(gen);
    };
    fill(sx::kOffW1, sx::kIn, sx::kH, sx::kH * sx::kIn);
    fill(sx::kOffW2, sx::kH, sx::kH, sx::kH * sx::kH);
    fill(sx::kOffW3, sx::kH, 1, sx::kH);
    return sx_mlp_set_params(m, p.data(), stream);
}

// mu[8], sigma[8] (features are used as (f - mu) / sigma) and the log-margin
// added to the prediction when producing squared radii.
int sx_mlp_set_normalizer(sx_mlp* m, const float* h_mu, const float* h_sigma, float margin, void* stream) {
    SX_TRY(check_mlp(m));
    if (!h_mu || !h_sigma) return fail(SX_EINVAL, "sx_mlp_set_normalizer: null pointer");
    float h[2 * sx::kIn + 1];
    for (int i = 0; i < sx::kIn; ++i) {
        if (!(h_sigma[i] > 0.f)) return fail(SX_EINVAL, "sx_mlp_set_normalizer: sigma must be positive");
        h[i] = h_mu[i];
        h[sx::kIn + i] = 1.0f / h_sigma[i];
    }
    h[2 * sx::kIn] = margin;
    hipStream_t st = sx::as_stream(stream);
    SX_HIP(hipMemcpyAsync(m->norm, h, sizeof(h), hipMemcpyHostToDevice, st));
    SX_TRY(pack_image(m, st));
    SX_HIP(hipStreamSynchronize(st));
    return SX_OK;
}

int sx_mlp_forward(sx_mlp* m, const float* d_feats, int64_t n, float* d_out, void* stream) {
    SX_TRY(check_mlp(m));
    if (n < 0 || (n > 0 && (!d_feats || !d_out))) return fail(SX_EINVAL, "sx_mlp_forward: bad arguments");
    if (n == 0) return SX_OK;
    unsigned grid = 1;
    SX_TRY(persistent_grid(sx::mlp_forward_kernel, g_resident_forward, n, &grid));
    sx::mlp_forward_kernel<<<grid, sx::kMlpBlock, 0, sx::as_stream(stream)>>>(dev_view(m), d_feats, n, 0, d_out);
    SX_LAUNCH_CHECK("mlp_forward_kernel");
    return SX_OK;
}

int sx_mlp_predict_radius2(sx_mlp* m, const float* d_feats, int64_t n, float* d_radius2, void* stream) {
    SX_TRY(check_mlp(m));
    if (n < 0 || (n > 0 && (!d_feats || !d_radius2))) return fail(SX_EINVAL, "sx_mlp_predict_radius2: bad arguments");
    if (n == 0) return SX_OK;
    unsigned grid = 1;
    SX_TRY(persistent_grid(sx::mlp_forward_kernel, g_resident_forward, n, &grid));
    sx::mlp_forward_kernel<<<grid, sx::kMlpBlock, 0, sx::as_stream(stream)>>>(dev_view(m), d_feats, n, 1, d_radius2);
    SX_LAUNCH_CHECK("mlp_forward_kernel");
    return SX_OK;
}

// Gradient of the mean squared error over n samples into d_grad (kNParams
// floats, device) and the summed squared error into *d_loss (device, nullable).
int sx_mlp_backward(sx_mlp* m, const float* d_feats, const float* d_targets, int64_t n, float* d_grad, float* d_loss,
                    void* stream) {
    SX_TRY(check_mlp(m));
    if (n <= 0 || n > m->max_batch || !d_feats || !d_targets)
        return fail(SX_EINVAL, "sx_mlp_backward: n must be in [1, max_batch] with non-null inputs");
    hipStream_t st = sx::as_stream(stream);
    const int64_t waves = (n + 63) / 64;
    sx::mlp_grad_kernel<<<sx::grid_for(waves * 64, sx::kMlpBlock), sx::kMlpBlock, 0, st>>>(
        dev_view(m), d_feats, d_targets, n, 1.0f / static_cast<float>(n), m->partial, m->loss_partial);
    SX_LAUNCH_CHECK("mlp_grad_kernel");
    // waves beyond n inside the last block wrote nothing: reduce only real waves
    sx::mlp_reduce_kernel<<<(sx::kNParams + 255) / 256, 256, 0, st>>>(m->partial, waves, d_grad ? d_grad : m->grad);
    SX_LAUNCH_CHECK("mlp_reduce_kernel");
    if (d_loss) {
        sx::mlp_sum_kernel<<<1, 256, 0, st>>>(m->loss_partial, waves, d_loss);
        SX_LAUNCH_CHECK("mlp_sum_kernel");
    }
    return SX_OK;
}

// One Adam step on a minibatch (lr, betas 0.9/0.999, eps 1e-8, decoupled
// weight decay wd); the summed squared error before the step goes to *d_loss (nullable).
int sx_mlp_adam_step(sx_mlp* m, const float* d_feats, const float* d_targets, int64_t n, float lr, float wd,
                     float* d_loss, void* stream) {
    SX_TRY(sx_mlp_backward(m, d_feats, d_targets, n, m->grad, d_loss, stream));
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f / (1.0f - std::pow(b1, static_cast<float>(m->step)));
    const float c2 = 1.0f / (1.0f - std::pow(b2, static_cast<float>(m->step)));
    sx::mlp_adam_kernel<<<(sx::kNParams + 255) / 256, 256, 0, sx::as_stream(stream)>>>(
        m->p, m->mo, m->ve, m->grad, lr, b1, b2, 1e-8f, c1, c2, wd, m->w1p, m->w2, m->w2t, m->w3);
    SX_LAUNCH_CHECK("mlp_adam_kernel");
    SX_TRY(pack_image(m, sx::as_stream(stream)));
    return SX_OK;
}

// Features of index points [q_start, q_start + n_q) into d_feats (n_q x 8 floats).
int sx_radius_features(sx_index* index, int64_t q_start, int64_t n_q, float* d_feats, void* stream) {
    const sx::TravNode* trav;
    const int32_t *lp, *lo, *counts;
    const int2 *up, *aux;
    const uint8_t* pn;
    const sx_otnode* oct;
    const float* bounds;
    int64_t limit;
    int64_t oct_cap;
    SX_TRY(sx::index_feature_arrays(index, &trav, &up, &aux, &pn, &lp, &lo, &oct, &oct_cap, &bounds, &counts,
                                    &limit, stream));
    if (q_start < 0 || n_q < 0 || q_start + n_q > limit) return fail(SX_EINVAL, "sx_radius_features: bad query range");
    if (n_q == 0) return SX_OK;
    if (!d_feats || reinterpret_cast<uintptr_t>(d_feats) % 16 != 0)
        return fail(SX_EINVAL, "sx_radius_features: output must be 16-byte aligned");
    sx::features_kernel<false><<<sx::grid_for(n_q, 256), 256, 0, sx::as_stream(stream)>>>(
        sx::FeatTree{trav, up, aux, pn, oct_cap}, lp, lo, oct, bounds, counts, q_start, n_q, d_feats, sx::CloudMap());
    SX_LAUNCH_CHECK("features_kernel");
    return SX_OK;
}

// Squared search radius of index points [q_start, q_start + n_q): features and
// MLP in one kernel, equal to sx_radius_features + sx_mlp_predict_radius2.
int sx_radius_predict(sx_index* index, sx_mlp* mlp, int64_t q_start, int64_t n_q, float* d_radius2, void* stream) {
    SX_TRY(check_mlp(mlp));
    const sx::TravNode* trav;
    const int32_t *lp, *lo, *counts;
    const int2 *up, *aux;
    const uint8_t* pn;
    const sx_otnode* oct;
    const float* bounds;
    int64_t limit;
    int64_t oct_cap;
    SX_TRY(sx::index_feature_arrays(index, &trav, &up, &aux, &pn, &lp, &lo, &oct, &oct_cap, &bounds, &counts,
                                    &limit, stream));
    if (q_start < 0 || n_q < 0 || q_start + n_q > limit) return fail(SX_EINVAL, "sx_radius_predict: bad query range");
    if (n_q == 0) return SX_OK;
    if (!d_radius2) return fail(SX_EINVAL, "sx_radius_predict: null output");
    unsigned grid = 1;
    SX_TRY(persistent_grid(sx::radius_kernel<false>, g_resident_radius, n_q, &grid));
    sx::radius_kernel<false><<<grid, sx::kMlpBlock, 0, sx::as_stream(stream)>>>(
        sx::FeatTree{trav, up, aux, pn, oct_cap}, lp, lo, oct, bounds, counts, q_start, n_q, dev_view(mlp), d_radius2,
        sx::CloudMap());
    SX_LAUNCH_CHECK("radius_kernel");
    return SX_OK;
}

// Features of a batch's slots [0, n_rows) (n_rows x 8 floats; rows of padding
// slots are zero); each cloud's rows equal sx_radius_features of a standalone
// index of that cloud.
int sx_batch_radius_features(sx_batch* batch, float* d_feats, int64_t n_rows, void* stream) {
    sx::BatchFeatures bf;
    SX_TRY(sx::batch_feature_arrays(batch, &bf, stream));
    if (n_rows < 0 || n_rows > bf.limit) return fail(SX_EINVAL, "sx_batch_radius_features: n_rows outside [0, slot capacity]");
    if (n_rows == 0) return SX_OK;
    if (!d_feats || reinterpret_cast<uintptr_t>(d_feats) % 16 != 0)
        return fail(SX_EINVAL, "sx_batch_radius_features: output must be 16-byte aligned");
    sx::features_kernel<true><<<sx::grid_for(n_rows, 256), 256, 0, sx::as_stream(stream)>>>(
        sx::FeatTree{bf.trav, nullptr, bf.aux, bf.pn, bf.oct_cap}, bf.leaf_parent, bf.leaf_oct, bf.oct, bf.bounds,
        bf.slots, 0, n_rows, d_feats, bf.cm);
    SX_LAUNCH_CHECK("features_kernel");
    return SX_OK;
}

// Squared search radius of a batch's slots [0, n_rows): features and MLP in one
// kernel, equal per cloud to sx_radius_predict of a standalone index.
int sx_batch_radius_predict(sx_batch* batch, sx_mlp* mlp, float* d_radius2, int64_t n_rows, void* stream) {
    SX_TRY(check_mlp(mlp));
    sx::BatchFeatures bf;
    SX_TRY(sx::batch_feature_arrays(batch, &bf, stream));
    if (n_rows < 0 || n_rows > bf.limit) return fail(SX_EINVAL, "sx_batch_radius_predict: n_rows outside [0, slot capacity]");
    if (n_rows == 0) return SX_OK;
    if (!d_radius2) return fail(SX_EINVAL, "sx_batch_radius_predict: null output");
    unsigned grid = 1;
    SX_TRY(persistent_grid(sx::radius_kernel<true>, g_resident_radius_batch, n_rows, &grid));
    sx::radius_kernel<true><<<grid, sx::kMlpBlock, 0, sx::as_stream(stream)>>>(
        sx::FeatTree{bf.trav, nullptr, bf.aux, bf.pn, bf.oct_cap}, bf.leaf_parent, bf.leaf_oct, bf.oct, bf.bounds,
        bf.slots, 0, n_rows, dev_view(mlp), d_radius2, bf.cm);
    SX_LAUNCH_CHECK("radius_kernel");
    return SX_OK;
}

}  // extern "C"
