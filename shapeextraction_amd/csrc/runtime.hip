// Device runtime plumbing exported through the C ABI so that harnesses in any
// language (Python ctypes, C, C++) can allocate HBM, move data and time work
// without linking HIP themselves.  Fills are kernels rather than
// hipMemsetAsync so that they are safe inside captured graphs.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "sx_internal.hpp"

namespace sx {

__global__ void fill_u8_kernel(uint8_t* dst, uint8_t value, size_t bytes) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t words = bytes / 16;
    uint4 v;
    const uint32_t b = value;
    v.x = v.y = v.z = v.w = b | (b << 8) | (b << 16) | (b << 24);
    // dst is 16-byte aligned by contract of the callers (hipMalloc / Carver).
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) d4[i] = v;
    for (size_t i = words * 16 + static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < bytes; i += stride)
        dst[i] = value;
}

__global__ void fill_i32_kernel(int32_t* dst, int32_t value, int64_t n) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = value;
}

__global__ void scatter_i32_kernel(const int32_t* __restrict__ src, const uint32_t* __restrict__ idx, int64_t n,
                                   int32_t* __restrict__ dst) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = src[i];
}

// dst[idx[i]] = src[i] for i < n.
int scatter_i32_launch(const int32_t* src, const uint32_t* idx, int64_t n, int32_t* dst, hipStream_t st) {
    if (n <= 0) return SX_OK;
    scatter_i32_kernel<<<grid_for(n, 256), 256, 0, st>>>(src, idx, n, dst);
    SX_LAUNCH_CHECK("scatter_i32_kernel");
    return SX_OK;
}

int fill_u8(void* dst, uint8_t value, size_t bytes, hipStream_t st) {
    if (bytes == 0) return SX_OK;
    if (reinterpret_cast<uintptr_t>(dst) % 16 != 0) return fail(SX_EINVAL, "fill_u8: destination not 16-byte aligned");
    unsigned g = grid_for(static_cast<int64_t>(bytes / 16 + 1), 256);
    if (g > 4096) g = 4096;
    fill_u8_kernel<<<g, 256, 0, st>>>(static_cast<uint8_t*>(dst), value, bytes);
    SX_LAUNCH_CHECK("fill_u8_kernel");
    return SX_OK;
}

int fill_i32(int32_t* dst, int32_t value, int64_t n, hipStream_t st) {
    if (n <= 0) return SX_OK;
    unsigned g = grid_for(n, 256);
    if (g > 4096) g = 4096;
    fill_i32_kernel<<<g, 256, 0, st>>>(dst, value, n);
    SX_LAUNCH_CHECK("fill_i32_kernel");
    return SX_OK;
}

}  // namespace sx

using sx::fail;

extern "C" {

const char* sx_last_error(void) { return sx::g_last_error.c_str(); }

int sx_abi_version(void) { return SX_ABI_VERSION; }

int sx_device_count(int* n) {
    if (!n) return fail(SX_EINVAL, "sx_device_count: null");
    SX_HIP(hipGetDeviceCount(n));
    return SX_OK;
}

int sx_set_device(int device) {
    SX_HIP(hipSetDevice(device));
    return SX_OK;
}

int sx_device_name(int device, char* buf, size_t buf_len) {
    if (!buf || buf_len == 0) return fail(SX_EINVAL, "sx_device_name: null buffer");
    hipDeviceProp_t prop;
    SX_HIP(hipGetDeviceProperties(&prop, device));
    std::snprintf(buf, buf_len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return SX_OK;
}

int sx_malloc(void** d_ptr, size_t bytes) {
    if (!d_ptr) return fail(SX_EINVAL, "sx_malloc: null");
    *d_ptr = nullptr;
    if (bytes == 0) return SX_OK;
    hipError_t e = hipMalloc(d_ptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *d_ptr = nullptr;
        return fail(SX_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return SX_OK;
}

int sx_free(void* d_ptr) {
    if (d_ptr) SX_HIP(hipFree(d_ptr));
    return SX_OK;
}

int sx_host_malloc(void** h_ptr, size_t bytes) {
    if (!h_ptr) return fail(SX_EINVAL, "sx_host_malloc: null");
    *h_ptr = nullptr;
    if (bytes == 0) return SX_OK;
    hipError_t e = hipHostMalloc(h_ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *h_ptr = nullptr;
        return fail(SX_ENOMEM, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    }
    return SX_OK;
}

int sx_host_free(void* h_ptr) {
    if (h_ptr) SX_HIP(hipHostFree(h_ptr));
    return SX_OK;
}

int sx_memcpy_htod(void* d_dst, const void* h_src, size_t bytes, void* stream) {
    if (bytes == 0) return SX_OK;
    if (!d_dst || !h_src) return fail(SX_EINVAL, "sx_memcpy_htod: null pointer");
    SX_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, sx::as_stream(stream)));
    return SX_OK;
}

int sx_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes, void* stream) {
    if (bytes == 0) return SX_OK;
    if (!h_dst || !d_src) return fail(SX_EINVAL, "sx_memcpy_dtoh: null pointer");
    SX_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, sx::as_stream(stream)));
    SX_HIP(hipStreamSynchronize(sx::as_stream(stream)));
    return SX_OK;
}

int sx_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes, void* stream) {
    if (bytes == 0) return SX_OK;
    if (!d_dst || !d_src) return fail(SX_EINVAL, "sx_memcpy_dtod: null pointer");
    SX_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, sx::as_stream(stream)));
    return SX_OK;
}

int sx_fill_u8(void* d_dst, uint8_t value, size_t bytes, void* stream) {
    if (bytes && !d_dst) return fail(SX_EINVAL, "sx_fill_u8: null pointer");
    return sx::fill_u8(d_dst, value, bytes, sx::as_stream(stream));
}

int sx_stream_create(void** stream) {
    if (!stream) return fail(SX_EINVAL, "sx_stream_create: null");
    hipStream_t s;
    SX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return SX_OK;
}

int sx_stream_destroy(void* stream) {
    if (stream) SX_HIP(hipStreamDestroy(sx::as_stream(stream)));
    return SX_OK;
}

int sx_stream_sync(void* stream) {
    SX_HIP(hipStreamSynchronize(sx::as_stream(stream)));
    return SX_OK;
}

int sx_device_sync(void) {
    SX_HIP(hipDeviceSynchronize());
    return SX_OK;
}

int sx_event_create(void** event) {
    if (!event) return fail(SX_EINVAL, "sx_event_create: null");
    hipEvent_t e;
    SX_HIP(hipEventCreate(&e));
    *event = e;
    return SX_OK;
}

int sx_event_destroy(void* event) {
    if (event) SX_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return SX_OK;
}

int sx_event_record(void* event, void* stream) {
    if (!event) return fail(SX_EINVAL, "sx_event_record: null");
    SX_HIP(hipEventRecord(static_cast<hipEvent_t>(event), sx::as_stream(stream)));
    return SX_OK;
}

int sx_event_elapsed_ms(void* start, void* stop, float* ms) {
    if (!start || !stop || !ms) return fail(SX_EINVAL, "sx_event_elapsed_ms: null");
    SX_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    SX_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return SX_OK;
}

}  // extern "C"
