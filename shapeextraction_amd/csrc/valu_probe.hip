// VALU issue-rate probe for the roofline of the VALU-bound kernels.
//
// Standalone program (not part of libshapes_hip.so), built beside the
// library.  Each probe kernel runs one instruction pattern over eight
// independent register chains per lane, kIters times, on 8,192 blocks of 256
// threads (eight waves per SIMD on 256 CUs); the rate is wave-instructions
// per second of the timed launches (HIP events, five launches after one
// warm-up).  The patterns are written in inline asm so that the compiler
// issues exactly the instruction measured; every result is stored, none is a
// memory access.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kIters = 2048;
constexpr int kBlocks = 8192;
constexpr int kThreads = 256;

#define CHECK(x)                                                                       \
    do {                                                                               \
        const hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));        \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

enum Op {
    kFmaF32,
    kMulF32,
    kAddU32,
    kXorB32,
    kMaxF32,
    kMinU32,
    kLshlB32,
    kCmpU32,
    kCmpU64,
    kSelE64,
    kCmpSel,
    kFmaF64,
};

// Instructions per lane per iteration of each pattern.
__host__ __device__ constexpr int per_iter(int op) { return op == kCmpSel ? 12 : 8; }

template <int OP>
__global__ void __launch_bounds__(kThreads) probe(float* out, const uint32_t* in) {
    uint32_t u0 = threadIdx.x, u1 = u0 + 1, u2 = u0 + 2, u3 = u0 + 3, u4 = u0 + 4, u5 = u0 + 5, u6 = u0 + 6,
             u7 = u0 + 7;
    float f0 = u0 * 1e-3f, f1 = f0 + 1, f2 = f0 + 2, f3 = f0 + 3, f4 = f0 + 4, f5 = f0 + 5, f6 = f0 + 6, f7 = f0 + 7;
    uint64_t w0 = u0, w1 = u1, w2 = u2, w3 = u3;
    double d0 = f0, d1 = f1, d2 = f2, d3 = f3, d4 = f4, d5 = f5, d6 = f6, d7 = f7;
    const uint32_t t = in[threadIdx.x], y = in[threadIdx.x + kThreads];
    const float b = 1.0001f, c = 0.5f;
    const double db = 1.0001, dc = 0.5;
    uint64_t m;
    asm volatile("v_cmp_lt_u32_e64 %0, %1, %2" : "=s"(m) : "v"(u0), "v"(t));
    uint64_t acc = 0;
    for (int it = 0; it < kIters; ++it) {
        if constexpr (OP == kFmaF32) {
#define P(x) asm volatile("v_fma_f32 %0, %1, %2, %0" : "+v"(x) : "v"(b), "v"(c));
            P(f0) P(f1) P(f2) P(f3) P(f4) P(f5) P(f6) P(f7)
#undef P
        } else if constexpr (OP == kMulF32) {
#define P(x) asm volatile("v_mul_f32 %0, %1, %0" : "+v"(x) : "v"(b));
            P(f0) P(f1) P(f2) P(f3) P(f4) P(f5) P(f6) P(f7)
#undef P
        } else if constexpr (OP == kAddU32) {
#define P(x) asm volatile("v_add_u32 %0, %1, %0" : "+v"(x) : "v"(t));
            P(u0) P(u1) P(u2) P(u3) P(u4) P(u5) P(u6) P(u7)
#undef P
        } else if constexpr (OP == kXorB32) {
#define P(x) asm volatile("v_xor_b32 %0, %1, %0" : "+v"(x) : "v"(t));
            P(u0) P(u1) P(u2) P(u3) P(u4) P(u5) P(u6) P(u7)
#undef P
        } else if constexpr (OP == kMaxF32) {
#define P(x) asm volatile("v_max_f32 %0, %1, %0" : "+v"(x) : "v"(b));
            P(f0) P(f1) P(f2) P(f3) P(f4) P(f5) P(f6) P(f7)
#undef P
        } else if constexpr (OP == kMinU32) {
#define P(x) asm volatile("v_min_u32 %0, %1, %0" : "+v"(x) : "v"(t));
            P(u0) P(u1) P(u2) P(u3) P(u4) P(u5) P(u6) P(u7)
#undef P
        } else if constexpr (OP == kLshlB32) {
#define P(x) asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(x));
            P(u0) P(u1) P(u2) P(u3) P(u4) P(u5) P(u6) P(u7)
#undef P
        } else if constexpr (OP == kCmpU32) {
            uint64_t m0, m1, m2, m3, m4, m5, m6, m7;
#define P(r, x) asm volatile("v_cmp_lt_u32_e64 %0, %1, %2" : "=s"(r) : "v"(x), "v"(t));
            P(m0, u0) P(m1, u1) P(m2, u2) P(m3, u3) P(m4, u4) P(m5, u5) P(m6, u6) P(m7, u7)
#undef P
            asm volatile("" : : "s"(m0), "s"(m1), "s"(m2), "s"(m3), "s"(m4), "s"(m5), "s"(m6), "s"(m7));
        } else if constexpr (OP == kCmpU64) {
            uint64_t m0, m1, m2, m3, m4, m5, m6, m7;
#define P(r, x, z) asm volatile("v_cmp_lt_u64_e64 %0, %1, %2" : "=s"(r) : "v"(x), "v"(z));
            P(m0, w0, w1) P(m1, w1, w2) P(m2, w2, w3) P(m3, w3, w0) P(m4, w0, w2) P(m5, w1, w3) P(m6, w2, w0)
            P(m7, w3, w1)
#undef P
            asm volatile("" : : "s"(m0), "s"(m1), "s"(m2), "s"(m3), "s"(m4), "s"(m5), "s"(m6), "s"(m7));
        } else if constexpr (OP == kSelE64) {
#define P(x) asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(x) : "v"(y), "s"(m));
            P(u0) P(u1) P(u2) P(u3) P(u4) P(u5) P(u6) P(u7)
#undef P
        } else if constexpr (OP == kCmpSel) {
            // a compare feeding two selects, as in a sorted-list insert
            uint64_t c0, c1, c2, c3;
#define P(x, z, r)                                                                                  \
    asm volatile("v_cmp_lt_u32_e64 %2, %3, %4\n\tv_cndmask_b32_e64 %0, %0, %5, %2\n\t"               \
                 "v_cndmask_b32_e64 %1, %1, %5, %2"                                                  \
                 : "+v"(x), "+v"(z), "=s"(r)                                                         \
                 : "v"(x), "v"(t), "v"(y));
            P(u0, u4, c0) P(u1, u5, c1) P(u2, u6, c2) P(u3, u7, c3)
#undef P
        } else if constexpr (OP == kFmaF64) {
#define P(x) asm volatile("v_fma_f64 %0, %1, %2, %0" : "+v"(x) : "v"(db), "v"(dc));
            P(d0) P(d1) P(d2) P(d3) P(d4) P(d5) P(d6) P(d7)
#undef P
        }
    }
    acc = m + w0 + w1 + w2 + w3;
    const float r = f0 + f1 + f2 + f3 + f4 + f5 + f6 + f7 + static_cast<float>(u0 + u1 + u2 + u3 + u4 + u5 + u6 + u7) +
                    static_cast<float>(d0 + d1 + d2 + d3 + d4 + d5 + d6 + d7) + static_cast<float>(acc & 1u);
    out[static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x] = r;
}

template <int OP>
void run(const char* name, float* out, const uint32_t* in) {
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    probe<OP><<<kBlocks, kThreads>>>(out, in);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < 5; ++r) probe<OP><<<kBlocks, kThreads>>>(out, in);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    ms /= 5.0f;
    const double waves = static_cast<double>(kBlocks) * (kThreads / 64);
    const double instr = waves * kIters * per_iter(OP);
    std::printf("%-34s %8.3f ms  %7.1f G wave-instructions/s\n", name, ms, instr / (ms * 1e-3) / 1e9);
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
}

}  // namespace

int main() {
    float* out = nullptr;
    uint32_t* in = nullptr;
    CHECK(hipMalloc(&out, sizeof(float) * kBlocks * kThreads));
    CHECK(hipMalloc(&in, sizeof(uint32_t) * 2 * kThreads));
    uint32_t h[2 * kThreads];
    for (int i = 0; i < 2 * kThreads; ++i) h[i] = (static_cast<uint32_t>(i) * 2654435761u) >> 8;
    CHECK(hipMemcpy(in, h, sizeof(h), hipMemcpyHostToDevice));
    run<kFmaF32>("v_fma_f32", out, in);
    run<kMulF32>("v_mul_f32", out, in);
    run<kAddU32>("v_add_u32", out, in);
    run<kXorB32>("v_xor_b32", out, in);
    run<kMaxF32>("v_max_f32", out, in);
    run<kMinU32>("v_min_u32", out, in);
    run<kLshlB32>("v_lshlrev_b32", out, in);
    run<kCmpU32>("v_cmp_lt_u32_e64", out, in);
    run<kCmpU64>("v_cmp_lt_u64_e64", out, in);
    run<kSelE64>("v_cndmask_b32_e64", out, in);
    run<kCmpSel>("v_cmp_lt_u32_e64 + 2 v_cndmask (12)", out, in);
    run<kFmaF64>("v_fma_f64", out, in);
    CHECK(hipFree(out));
    CHECK(hipFree(in));
    return 0;
}
