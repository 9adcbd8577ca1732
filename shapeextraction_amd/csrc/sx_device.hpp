// Device-side building blocks shared by the kernels: Morton quantisation and
// bit interleave, the Karras prefix length, and the two distance functions
// whose rounding must match the host (CPU) restatement bit for bit.
//
// Every float operation on a bit-exact path is a single IEEE
// round-to-nearest-even operation written in an explicit order (__fadd_rn /
// __fsub_rn / __fmul_rn / __fdiv_rn, plain operators in this HIP version), and
// no pair may be contracted into an FMA: the build passes -ffp-contract=off
// and the pragma below repeats it for every file including this header.  The
// host restatement uses the same operation order.
#pragma once

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sx {

constexpr int kWave = 64;

// 2^(code_bits/3) - 1: 2097151 for 63-bit codes, 1023 for 30-bit codes
// (bitscale of src/MortonUtils.cu:13).
__host__ __device__ __forceinline__ uint32_t axis_scale(int code_bits) {
    return 0xFFFFFFFFu >> (32 - code_bits / 3);
}

// Spread the low 21 bits of v so that bit i lands on bit 3i.
__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Inverse of spread3: gather bits 0, 3, 6, ... into the low 21 bits.
__device__ __forceinline__ uint32_t compact3(uint64_t x) {
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10C30C30C30C30C3ull;
    x = (x ^ (x >> 4)) & 0x100F00F00F00F00Full;
    x = (x ^ (x >> 8)) & 0x001F0000FF0000FFull;
    x = (x ^ (x >> 16)) & 0x001F00000000FFFFull;
    x = (x ^ (x >> 32)) & 0x00000000001FFFFFull;
    return static_cast<uint32_t>(x);
}

// x -> bits 3i, y -> 3i+1, z -> 3i+2 (libmorton's morton3D_64_encode order).
__device__ __forceinline__ uint64_t interleave3(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

// (uint32)(bitscale * ((v - min) / range)) with the reference's operation
// order (src/MortonUtils.cu:13-16).  For points inside the cube the product
// lies in [0, bitscale]; the clamp only guards points outside caller-given
// bounds.  A non-positive range maps every point to 0.
__device__ __forceinline__ uint32_t quantize(float v, float min_coord, float range, float fscale) {
    if (!(range > 0.0f)) return 0u;
    float t = __fdiv_rn(__fsub_rn(v, min_coord), range);
    float s = __fmul_rn(fscale, t);
    s = fminf(fmaxf(s, 0.0f), fscale);
    return static_cast<uint32_t>(s);
}

// (float)raw / bitscale, correctly rounded.  For the two Morton scales
// (2^21 - 1 and 2^10 - 1) one reciprocal product plus one Newton correction,
// q = fma(fma(-q0, d, x), 1/d, q0) with q0 = x * (1/d), equals the IEEE
// quotient for every raw in [0, 2^21); that equality is checked exhaustively
// on the host by tests/test_arith.py.  Any other scale takes the full division.
__device__ __forceinline__ float raw_over_scale(uint32_t raw, float fscale) {
    const float x = static_cast<float>(raw);
    float r;
    if (fscale == 2097151.0f) r = 0x1.000008p-21f;
    else if (fscale == 1023.0f) r = 0x1.00401p-10f;
    else return __fdiv_rn(x, fscale);
    const float q0 = __fmul_rn(x, r);
    return __fmaf_rn(__fmaf_rn(-q0, fscale, x), r, q0);
}

// ((float)raw / bitscale) * range + min, uncontracted (src/MortonUtils.cu:25-27).
__device__ __forceinline__ float dequantize(uint32_t raw, float fscale, float range, float min_coord) {
    return __fadd_rn(__fmul_rn(raw_over_scale(raw, fscale), range), min_coord);
}

// Length of the common prefix of two codes of `code_bits` significant bits
// (delta of src/RadixTree.cu:62-69, generalised from clz - 1 to code_bits).
__device__ __forceinline__ int prefix_len(uint64_t a, uint64_t b, int code_bits) {
    uint64_t x = a ^ b;
    return x == 0 ? code_bits : __clzll(static_cast<long long>(x)) - (64 - code_bits);
}

// Point::distance2 (src/PointCloud.hpp:30-35) with a fixed, uncontracted
// order: ((dx*dx + dy*dy) + dz*dz), d = q - p.
__device__ __forceinline__ float dist2(float qx, float qy, float qz, float px, float py, float pz) {
    float dx = __fsub_rn(qx, px), dy = __fsub_rn(qy, py), dz = __fsub_rn(qz, pz);
    return __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
}

// Squared distance from q to the box [lo, hi] in the same operation order as
// dist2.  For every point p inside the box, box_dist2(q) <= dist2(q, p)
// exactly: each rounded operation is monotone, |fl(q-p)| >= fl(gap) per axis.
__device__ __forceinline__ float box_gap(float q, float lo, float hi) {
    return q < lo ? __fsub_rn(lo, q) : (q > hi ? __fsub_rn(q, hi) : 0.0f);
}
__device__ __forceinline__ float box_dist2(float qx, float qy, float qz, float4 lo, float4 hi) {
    float gx = box_gap(qx, lo.x, hi.x), gy = box_gap(qy, lo.y, hi.y), gz = box_gap(qz, lo.z, hi.z);
    return __fadd_rn(__fadd_rn(__fmul_rn(gx, gx), __fmul_rn(gy, gy)), __fmul_rn(gz, gz));
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

}  // namespace sx
