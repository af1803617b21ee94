// mans_runtime.hip -- device, memory, copy, stream and event entry points of the C ABI.
// Replaces the CUDA runtime calls of CUHDGPUMemoryBuffer<T> and CUHDGPUDecoderMemory
// (include/cuhd_gpu_memory_buffer.hxx:18-68, src/multians_gpu_decoder_memory.cc:25-92
// of the reference) and the CUERR message (include/cuhd_cuda_definitions.h:21-28).
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "mans_internal.h"

namespace {
thread_local char g_msg[512] = "no error";

int hip_fail(hipError_t e, const char* what) {
    return set_error(MANS_EHIP, "%s: %s", what, hipGetErrorString(e));
}
}  // namespace

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" {

const char* mans_error_string(int code) {
    switch (code) {
        case MANS_OK: return "success";
        case MANS_EINVAL: return "invalid argument";
        case MANS_EHIP: return "HIP runtime error";
        case MANS_ENOMEM: return "out of memory or workspace too small";
        case MANS_ETABLE: return "decoder table rejected";
        case MANS_ECORRUPT: return "stream inconsistent with table or sizes";
        case MANS_EUNSUPPORTED: return "unsupported";
        default: return "unknown error";
    }
}

const char* mans_last_error(void) { return g_msg; }

int mans_set_device(int device) {
    hipError_t e = hipSetDevice(device);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipSetDevice");
}

int mans_get_device_count(int* count) {
    if (!count) return set_error(MANS_EINVAL, "count is NULL");
    hipError_t e = hipGetDeviceCount(count);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipGetDeviceCount");
}

int mans_malloc(void** ptr, size_t bytes) {
    if (!ptr) return set_error(MANS_EINVAL, "ptr is NULL");
    *ptr = nullptr;
    if (!bytes) return MANS_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) return set_error(MANS_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    e = hipMemset(*ptr, 0, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMemset");
    return MANS_OK;
}

int mans_free(void* ptr) {
    if (!ptr) return MANS_OK;
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipFree");
}

int mans_memset(void* dst, int value, size_t bytes, void* stream) {
    hipError_t e = hipMemsetAsync(dst, value, bytes, (hipStream_t)stream);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipMemsetAsync");
}

int mans_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
    if (e == hipSuccess && !stream) e = hipStreamSynchronize(nullptr);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipMemcpy H2D");
}

int mans_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess && !stream) e = hipStreamSynchronize(nullptr);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipMemcpy D2H");
}

int mans_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipMemcpy D2D");
}

int mans_host_alloc_pinned(void** ptr, size_t bytes) {
    if (!ptr) return set_error(MANS_EINVAL, "ptr is NULL");
    hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    return e == hipSuccess ? MANS_OK : set_error(MANS_ENOMEM, "hipHostMalloc: %s", hipGetErrorString(e));
}

int mans_host_free_pinned(void* ptr) {
    if (!ptr) return MANS_OK;
    hipError_t e = hipHostFree(ptr);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipHostFree");
}

int mans_stream_create(void** stream) {
    if (!stream) return set_error(MANS_EINVAL, "stream is NULL");
    hipStream_t s;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    *stream = (void*)s;
    return MANS_OK;
}

int mans_stream_destroy(void* stream) {
    hipError_t e = hipStreamDestroy((hipStream_t)stream);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipStreamDestroy");
}

int mans_sync(void* stream) {
    hipError_t e = stream ? hipStreamSynchronize((hipStream_t)stream) : hipDeviceSynchronize();
    return e == hipSuccess ? MANS_OK : hip_fail(e, "synchronize");
}

int mans_event_create(void** ev) {
    if (!ev) return set_error(MANS_EINVAL, "ev is NULL");
    hipEvent_t e0;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    *ev = (void*)e0;
    return MANS_OK;
}

int mans_event_destroy(void* ev) {
    hipError_t e = hipEventDestroy((hipEvent_t)ev);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipEventDestroy");
}

int mans_event_record(void* ev, void* stream) {
    hipError_t e = hipEventRecord((hipEvent_t)ev, (hipStream_t)stream);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipEventRecord");
}

int mans_event_elapsed_ms(void* start, void* stop, float* ms) {
    if (!ms) return set_error(MANS_EINVAL, "ms is NULL");
    hipError_t e = hipEventSynchronize((hipEvent_t)stop);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
    return e == hipSuccess ? MANS_OK : hip_fail(e, "hipEventElapsedTime");
}

}  // extern "C"
