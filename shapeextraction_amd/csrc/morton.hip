// Cube bounds, Morton encode and decode.
//
// bbox:   one streaming pass over x, y, z (12 B/point) into per-block partials
//         and a one-block reduction; replaces the six cub reductions and the
//         host min/max of RadixTree::encodePoints (src/RadixTree.cu:208-241).
// encode: 4 points per thread, float4 loads, 20 B/point of HBM traffic;
//         replaces makeCodes / pointToCode (src/RadixTree.cu:21-39,
//         src/MortonUtils.cu:7-19).
// decode: codeToPoint per unique code (src/MortonUtils.cu:21-28,
//         src/Octree.cu:14-24), written both as the API's 12-byte AoS points
//         and as the 16-byte float4 copy the traversal kernels load.
#include <hip/hip_runtime.h>

#include <cmath>

#include "sx_internal.hpp"

namespace sx {

constexpr int kBBoxBlock = 256;
constexpr int kBBoxMaxBlocks = 4096;

__device__ __forceinline__ void bbox_acc(float v, float& mn, float& mx, int& bad) {
    const bool ok = isfinite(v);
    mn = fminf(mn, ok ? v : INFINITY);
    mx = fmaxf(mx, ok ? v : -INFINITY);
    bad += ok ? 0 : 1;
}

__device__ __forceinline__ void block_minmax(float& mn, float& mx, int& bad) {
    __shared__ float s_mn[kBBoxBlock / kWave], s_mx[kBBoxBlock / kWave];
    __shared__ int s_bad[kBBoxBlock / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, o, kWave));
        mx = fmaxf(mx, __shfl_xor(mx, o, kWave));
        bad += __shfl_xor(bad, o, kWave);
    }
    const int w = static_cast<int>(threadIdx.x) / kWave;
    if (lane_id() == 0) { s_mn[w] = mn; s_mx[w] = mx; s_bad[w] = bad; }
    __syncthreads();
    mn = s_mn[0]; mx = s_mx[0]; bad = s_bad[0];
#pragma unroll
    for (int i = 1; i < kBBoxBlock / kWave; ++i) {
        mn = fminf(mn, s_mn[i]);
        mx = fmaxf(mx, s_mx[i]);
        bad += s_bad[i];
    }
    __syncthreads();
}

// Per-block (min, max, non-finite count) into partial / bad_partial, then
// one block reduces them.  A last-block-done reduction in one kernel would
// need an agent-scope release per block, i.e. an L2 write-back on a GPU whose
// XCDs have separate L2s; the kernel boundary orders the partials instead.
// kVec: x, y, z are 16-byte aligned and read as float4.
// clear (nullable): n_clear int32 zeroed by block 0 for the caller, so that
// its counts need no launch of their own (stream order makes the zeros
// visible to every later kernel).
template <bool kVec>
__global__ void __launch_bounds__(kBBoxBlock) bbox_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                          const float* __restrict__ z, int64_t n, float2* partial,
                                                          int* bad_partial, int32_t* clear, int n_clear) {
    if (clear && blockIdx.x == 0 && static_cast<int>(threadIdx.x) < n_clear) clear[threadIdx.x] = 0;
    float mn = INFINITY, mx = -INFINITY;
    int bad = 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t rest = 0;
    if (kVec) {
        const int64_t n4 = n / 4;
        const float4* x4 = reinterpret_cast<const float4*>(x);
        const float4* y4 = reinterpret_cast<const float4*>(y);
        const float4* z4 = reinterpret_cast<const float4*>(z);
#pragma unroll 4
        for (int64_t i = tid; i < n4; i += stride) {
            const float4 a = x4[i], b = y4[i], c = z4[i];
            bbox_acc(a.x, mn, mx, bad); bbox_acc(a.y, mn, mx, bad); bbox_acc(a.z, mn, mx, bad); bbox_acc(a.w, mn, mx, bad);
            bbox_acc(b.x, mn, mx, bad); bbox_acc(b.y, mn, mx, bad); bbox_acc(b.z, mn, mx, bad); bbox_acc(b.w, mn, mx, bad);
            bbox_acc(c.x, mn, mx, bad); bbox_acc(c.y, mn, mx, bad); bbox_acc(c.z, mn, mx, bad); bbox_acc(c.w, mn, mx, bad);
        }
        rest = n4 * 4;
    }
    for (int64_t i = rest + tid; i < n; i += stride) {
        bbox_acc(x[i], mn, mx, bad);
        bbox_acc(y[i], mn, mx, bad);
        bbox_acc(z[i], mn, mx, bad);
    }
    block_minmax(mn, mx, bad);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = make_float2(mn, mx);
        bad_partial[blockIdx.x] = bad;
    }
}

__global__ void __launch_bounds__(kBBoxBlock) bbox_final_kernel(const float2* __restrict__ partial,
                                                                const int* __restrict__ bad_partial, int n_blocks,
                                                                float* bounds, int32_t* d_bad) {
    float mn = INFINITY, mx = -INFINITY;
    int bad = 0;
    for (int b = threadIdx.x; b < n_blocks; b += blockDim.x) {
        const float2 p = partial[b];
        mn = fminf(mn, p.x);
        mx = fmaxf(mx, p.y);
        bad += bad_partial[b];
    }
    block_minmax(mn, mx, bad);
    if (threadIdx.x == 0) {
        bounds[0] = mn;
        bounds[1] = mx;
        bounds[2] = __fsub_rn(mx, mn);
        if (d_bad) *d_bad = bad;
    }
}

inline int bbox_blocks(int64_t n) {
    int64_t b = (n + kBBoxBlock * 16 - 1) / (kBBoxBlock * 16);
    if (b < 1) b = 1;
    if (b > kBBoxMaxBlocks) b = kBBoxMaxBlocks;
    return static_cast<int>(b);
}

size_t bbox_ws_bytes() {
    Carver c(nullptr);
    c.take<float2>(kBBoxMaxBlocks);
    c.take<int>(kBBoxMaxBlocks);
    return c.off;
}

int bbox_launch(const float* x, const float* y, const float* z, int64_t n, float* bounds, int32_t* d_bad, void* ws,
                hipStream_t st, int32_t* clear = nullptr, int n_clear = 0) {
    if (n_clear > kBBoxBlock) return fail(SX_EINVAL, "bbox_launch: too many words to clear");
    Carver c(ws);
    float2* partial = c.take<float2>(kBBoxMaxBlocks);
    int* badp = c.take<int>(kBBoxMaxBlocks);
    const int nb = bbox_blocks(n);
    const bool vec = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y) |
                       reinterpret_cast<uintptr_t>(z)) & 15u) == 0;
    if (vec)
        bbox_kernel<true><<<nb, kBBoxBlock, 0, st>>>(x, y, z, n, partial, badp, clear, n_clear);
    else
        bbox_kernel<false><<<nb, kBBoxBlock, 0, st>>>(x, y, z, n, partial, badp, clear, n_clear);
    SX_LAUNCH_CHECK("bbox_kernel");
    bbox_final_kernel<<<1, kBBoxBlock, 0, st>>>(partial, badp, nb, bounds, d_bad);
    SX_LAUNCH_CHECK("bbox_final_kernel");
    return SX_OK;
}

__device__ __forceinline__ uint64_t encode_one(float x, float y, float z, float mn, float range, float fscale) {
    return interleave3(quantize(x, mn, range, fscale), quantize(y, mn, range, fscale), quantize(z, mn, range, fscale));
}

constexpr int kEncodeBlock = 256;

__global__ void __launch_bounds__(kEncodeBlock) encode_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                              const float* __restrict__ z, int64_t n,
                                                              const float* __restrict__ bounds, int code_bits,
                                                              uint64_t* __restrict__ codes) {
    const float mn = bounds[0], range = bounds[2];
    const float fscale = static_cast<float>(axis_scale(code_bits));
    const int64_t i4 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
    if (i4 + 3 < n) {
        const float4 vx = *reinterpret_cast<const float4*>(x + i4);
        const float4 vy = *reinterpret_cast<const float4*>(y + i4);
        const float4 vz = *reinterpret_cast<const float4*>(z + i4);
        ulonglong2 c01, c23;
        c01.x = encode_one(vx.x, vy.x, vz.x, mn, range, fscale);
        c01.y = encode_one(vx.y, vy.y, vz.y, mn, range, fscale);
        c23.x = encode_one(vx.z, vy.z, vz.z, mn, range, fscale);
        c23.y = encode_one(vx.w, vy.w, vz.w, mn, range, fscale);
        reinterpret_cast<ulonglong2*>(codes + i4)[0] = c01;
        reinterpret_cast<ulonglong2*>(codes + i4)[1] = c23;
    } else {
        for (int64_t i = i4; i < n; ++i) codes[i] = encode_one(x[i], y[i], z[i], mn, range, fscale);
    }
}

// Scalar variant for unaligned inputs.
__global__ void __launch_bounds__(kEncodeBlock) encode_kernel_scalar(const float* __restrict__ x,
                                                                     const float* __restrict__ y,
                                                                     const float* __restrict__ z, int64_t n,
                                                                     const float* __restrict__ bounds, int code_bits,
                                                                     uint64_t* __restrict__ codes) {
    const float mn = bounds[0], range = bounds[2];
    const float fscale = static_cast<float>(axis_scale(code_bits));
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) codes[i] = encode_one(x[i], y[i], z[i], mn, range, fscale);
}

int encode_launch(const float* x, const float* y, const float* z, int64_t n, const float* bounds, int code_bits,
                  uint64_t* codes, hipStream_t st) {
    if (n <= 0) return SX_OK;
    const bool aligned = (reinterpret_cast<uintptr_t>(x) % 16 == 0) && (reinterpret_cast<uintptr_t>(y) % 16 == 0) &&
                         (reinterpret_cast<uintptr_t>(z) % 16 == 0) && (reinterpret_cast<uintptr_t>(codes) % 16 == 0);
    if (aligned) {
        encode_kernel<<<grid_for((n + 3) / 4, kEncodeBlock), kEncodeBlock, 0, st>>>(x, y, z, n, bounds, code_bits,
                                                                                     codes);
        SX_LAUNCH_CHECK("encode_kernel");
    } else {
        encode_kernel_scalar<<<grid_for(n, kEncodeBlock), kEncodeBlock, 0, st>>>(x, y, z, n, bounds, code_bits,
                                                                                  codes);
        SX_LAUNCH_CHECK("encode_kernel_scalar");
    }
    return SX_OK;
}

__global__ void __launch_bounds__(256) decode_kernel(const uint64_t* __restrict__ codes, const int32_t* d_count,
                                                     int64_t n_max, const float* __restrict__ bounds, int code_bits,
                                                     float* __restrict__ pts12, float4* __restrict__ pts16) {
    const int64_t n = scan_count(d_count, n_max);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float mn = bounds[0], range = bounds[2];
    const float fscale = static_cast<float>(axis_scale(code_bits));
    const uint64_t c = codes[i];
    const float px = dequantize(compact3(c), fscale, range, mn);
    const float py = dequantize(compact3(c >> 1), fscale, range, mn);
    const float pz = dequantize(compact3(c >> 2), fscale, range, mn);
    if (pts12) {
        pts12[3 * i] = px;
        pts12[3 * i + 1] = py;
        pts12[3 * i + 2] = pz;
    }
    if (pts16) pts16[i] = make_float4(px, py, pz, 0.0f);
}

// The 12-byte (API layout) copy of the float4 points, [0, *d_count).
__global__ void __launch_bounds__(256) pack_points12_kernel(const float4* __restrict__ pts16, const int32_t* d_count,
                                                            int64_t n_max, float* __restrict__ pts12) {
    const int64_t n = scan_count(d_count, n_max);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pts16[i];
    pts12[3 * i] = p.x;
    pts12[3 * i + 1] = p.y;
    pts12[3 * i + 2] = p.z;
}

int pack_points12_launch(const float4* pts16, const int32_t* d_count, int64_t n_max, float* pts12, hipStream_t st) {
    if (n_max <= 0) return SX_OK;
    pack_points12_kernel<<<grid_for(n_max, 256), 256, 0, st>>>(pts16, d_count, n_max, pts12);
    SX_LAUNCH_CHECK("pack_points12_kernel");
    return SX_OK;
}

int decode_launch(const uint64_t* codes, const int32_t* d_count, int64_t n_max, const float* bounds, int code_bits,
                  float* pts12, float4* pts16, hipStream_t st) {
    if (n_max <= 0) return SX_OK;
    decode_kernel<<<grid_for(n_max, 256), 256, 0, st>>>(codes, d_count, n_max, bounds, code_bits, pts12, pts16);
    SX_LAUNCH_CHECK("decode_kernel");
    return SX_OK;
}

// ---------------------------------------------------------------------------
// Batched clouds: cloud c holds input points [off[c], off[c+1]).
// ---------------------------------------------------------------------------
// kSegBBoxSplit blocks per cloud reduce strided shares of its points into
// partials; one thread per cloud then gives bounds[4c] = (min, max, range)
// over its three axes (the same min/max as bbox_kernel, so the same bits);
// non-finite coordinates add to *d_bad.
constexpr int kSegBBoxSplit = 16;

__global__ void __launch_bounds__(kBBoxBlock) seg_bbox_kernel(const float* __restrict__ x, const float* __restrict__ y,
                                                              const float* __restrict__ z,
                                                              const int64_t* __restrict__ off, float* partial,
                                                              int32_t* d_bad) {
    const int c = blockIdx.x / kSegBBoxSplit, part = blockIdx.x % kSegBBoxSplit;
    float mn = INFINITY, mx = -INFINITY;
    int bad = 0;
    const int64_t stride = static_cast<int64_t>(kSegBBoxSplit) * blockDim.x;
    for (int64_t i = off[c] + static_cast<int64_t>(part) * blockDim.x + threadIdx.x; i < off[c + 1]; i += stride) {
        bbox_acc(x[i], mn, mx, bad);
        bbox_acc(y[i], mn, mx, bad);
        bbox_acc(z[i], mn, mx, bad);
    }
    block_minmax(mn, mx, bad);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = mn;
        partial[2 * blockIdx.x + 1] = mx;
        if (bad) atomicAdd(d_bad, bad);
    }
}

__global__ void __launch_bounds__(256) seg_bbox_final_kernel(const float* __restrict__ partial, int n_clouds,
                                                             float* bounds) {
    const int c = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_clouds) return;
    float mn = INFINITY, mx = -INFINITY;
    for (int p = 0; p < kSegBBoxSplit; ++p) {
        mn = fminf(mn, partial[2 * (c * kSegBBoxSplit + p)]);
        mx = fmaxf(mx, partial[2 * (c * kSegBBoxSplit + p) + 1]);
    }
    bounds[4 * c] = mn;
    bounds[4 * c + 1] = mx;
    bounds[4 * c + 2] = __fsub_rn(mx, mn);
    bounds[4 * c + 3] = 0.0f;
}

size_t seg_bbox_ws_bytes(int max_clouds) { return sizeof(float) * 2 * kSegBBoxSplit * max_clouds; }

int seg_bbox_launch(const float* x, const float* y, const float* z, const int64_t* off, int n_clouds, float* bounds,
                    int32_t* d_bad, float* partial, hipStream_t st) {
    if (n_clouds <= 0) return SX_OK;
    seg_bbox_kernel<<<n_clouds * kSegBBoxSplit, kBBoxBlock, 0, st>>>(x, y, z, off, partial, d_bad);
    SX_LAUNCH_CHECK("seg_bbox_kernel");
    seg_bbox_final_kernel<<<grid_for(n_clouds, 256), 256, 0, st>>>(partial, n_clouds, bounds);
    SX_LAUNCH_CHECK("seg_bbox_final_kernel");
    return SX_OK;
}

constexpr int kSegEncodeItems = 8;
__global__ void __launch_bounds__(kEncodeBlock) seg_encode_kernel(const float* __restrict__ x,
                                                                  const float* __restrict__ y,
                                                                  const float* __restrict__ z, int64_t n,
                                                                  const int64_t* __restrict__ off, int n_clouds,
                                                                  const float* __restrict__ bounds, int code_bits,
                                                                  uint64_t* __restrict__ codes) {
    // kSegEncodeItems points per thread; the clouds of the block's first and
    // last point are found once per block (uniform: scalar loads), and a
    // point searches only when the block straddles a cloud boundary
    const int64_t span = static_cast<int64_t>(blockDim.x) * kSegEncodeItems;
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * span;
    const int64_t b1 = (b0 + span < n ? b0 + span : n) - 1;
    const int c0 = cloud_of(off, n_clouds, b0), c1 = cloud_of(off, n_clouds, b1);
    const float fscale = static_cast<float>(axis_scale(code_bits));
#pragma unroll
    for (int r = 0; r < kSegEncodeItems; ++r) {
        const int64_t i = b0 + static_cast<int64_t>(r) * blockDim.x + threadIdx.x;
        if (i >= n) break;
        const int c = c0 == c1 ? c0 : c0 + cloud_of(off + c0, c1 - c0 + 1, i);
        codes[i] = encode_one(x[i], y[i], z[i], bounds[4 * c], bounds[4 * c + 2], fscale);
    }
}

int seg_encode_launch(const float* x, const float* y, const float* z, int64_t n, const int64_t* off, int n_clouds,
                      const float* bounds, int code_bits, uint64_t* codes, hipStream_t st) {
    if (n <= 0) return SX_OK;
    seg_encode_kernel<<<grid_for(n, kEncodeBlock * kSegEncodeItems), kEncodeBlock, 0, st>>>(x, y, z, n, off, n_clouds, bounds,
                                                                            code_bits, codes);
    SX_LAUNCH_CHECK("seg_encode_kernel");
    return SX_OK;
}

// Points of every slot [0, *d_slots): a cloud's unique codes with its own
// cube; the padding slots after a cloud's last point are zero.
__global__ void __launch_bounds__(256) seg_decode_kernel(const uint64_t* __restrict__ codes, const int32_t* d_slots,
                                                         int64_t n_max, CloudMap cm, const int32_t* __restrict__ cloud_of_tile,
                                                         const float* __restrict__ bounds, int code_bits,
                                                         float4* __restrict__ pts16) {
    const int64_t n = scan_count(d_slots, n_max);
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 cloud = cm.at(i, static_cast<int>(n));
    if (i - cloud.x >= cloud.y) {
        pts16[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const int c = cloud_of_tile[i / kCloudAlign];
    const float mn = bounds[4 * c], range = bounds[4 * c + 2];
    const float fscale = static_cast<float>(axis_scale(code_bits));
    const uint64_t v = codes[i];
    pts16[i] = make_float4(dequantize(compact3(v), fscale, range, mn), dequantize(compact3(v >> 1), fscale, range, mn),
                           dequantize(compact3(v >> 2), fscale, range, mn), 0.0f);
}

int seg_decode_launch(const uint64_t* codes, const int32_t* d_slots, int64_t n_max, CloudMap cm, const float* bounds,
                      int code_bits, float4* pts16, hipStream_t st) {
    if (n_max <= 0) return SX_OK;
    seg_decode_kernel<<<grid_for(n_max, 256), 256, 0, st>>>(codes, d_slots, n_max, cm, cm.tile_cloud, bounds,
                                                            code_bits, pts16);
    SX_LAUNCH_CHECK("seg_decode_kernel");
    return SX_OK;
}

inline bool valid_code_bits(int b) { return b == 63 || b == 30; }

}  // namespace sx

using sx::fail;

extern "C" {

int sx_bbox(const float* d_x, const float* d_y, const float* d_z, int64_t n, float* d_bounds, int32_t* d_bad, void* ws,
            size_t* ws_bytes, void* stream) {
    if (!ws_bytes) return fail(SX_EINVAL, "sx_bbox: ws_bytes is null");
    if (!ws) {
        *ws_bytes = sx::bbox_ws_bytes();
        return SX_OK;
    }
    if (*ws_bytes < sx::bbox_ws_bytes()) return fail(SX_ENOMEM, "sx_bbox: workspace too small");
    if (n <= 0 || !d_x || !d_y || !d_z || !d_bounds) return fail(SX_EINVAL, "sx_bbox: bad arguments");
    return sx::bbox_launch(d_x, d_y, d_z, n, d_bounds, d_bad, ws, sx::as_stream(stream));
}

int sx_morton_encode(const float* d_x, const float* d_y, const float* d_z, int64_t n, const float* d_bounds,
                     int code_bits, uint64_t* d_codes, void* stream) {
    if (!sx::valid_code_bits(code_bits)) return fail(SX_EINVAL, "sx_morton_encode: code_bits must be 30 or 63");
    if (n < 0 || (n > 0 && (!d_x || !d_y || !d_z || !d_bounds || !d_codes)))
        return fail(SX_EINVAL, "sx_morton_encode: bad arguments");
    return sx::encode_launch(d_x, d_y, d_z, n, d_bounds, code_bits, d_codes, sx::as_stream(stream));
}

int sx_morton_decode(const uint64_t* d_codes, const int32_t* d_count, int64_t n_max, const float* d_bounds,
                     int code_bits, float* d_points, void* stream) {
    if (!sx::valid_code_bits(code_bits)) return fail(SX_EINVAL, "sx_morton_decode: code_bits must be 30 or 63");
    if (n_max < 0 || (n_max > 0 && (!d_codes || !d_bounds || !d_points)))
        return fail(SX_EINVAL, "sx_morton_decode: bad arguments");
    return sx::decode_launch(d_codes, d_count, n_max, d_bounds, code_bits, d_points, nullptr, sx::as_stream(stream));
}

}  // extern "C"
