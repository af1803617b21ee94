// cuhd_gpu_memory_buffer.h -- one device allocation mirroring a host array.
// Drop-in for the reference's CUHDGPUMemoryBuffer<T> (include/cuhd_gpu_memory_buffer.h:20-38,
// include/cuhd_gpu_memory_buffer.hxx:18-68): zero-filled allocate, free in the
// destructor, blocking host<->device copies, lazy allocate on the first upload.
// Errors print a message and exit(1), as the reference's CUERR does.
#ifndef CUHD_GPU_MEMORY_BUFFER_
#define CUHD_GPU_MEMORY_BUFFER_

#include <cstdio>
#include <cstdlib>
#include <memory>

#include "cuhd_constants.h"
#include "mans_c_api.h"

namespace cuhd {

// print + exit(1) on a failed C-ABI call (the reference's CUERR behaviour)
inline void mans_check(int rc, const char* file, int line) {
    if (rc != MANS_OK) {
        std::printf("MANS error: %s (%s) : %s, line %d\n", mans_error_string(rc), mans_last_error(),
                    file, line);
        std::exit(1);
    }
}
#define MANS_CHECK(call) ::cuhd::mans_check((call), __FILE__, __LINE__)

template <typename T>
class CUHDGPUMemoryBuffer {
public:
    CUHDGPUMemoryBuffer(T* buffer, size_t size) : buffer_(buffer), size_(size) {}
    ~CUHDGPUMemoryBuffer() { free(); }
    CUHDGPUMemoryBuffer(const CUHDGPUMemoryBuffer&) = delete;
    CUHDGPUMemoryBuffer& operator=(const CUHDGPUMemoryBuffer&) = delete;

    T* get() { return buffer_device_; }
    size_t size() const { return size_; }

    void allocate() {
        if (is_allocated_) return;
        void* p = nullptr;
        MANS_CHECK(mans_malloc(&p, size_ * sizeof(T)));
        buffer_device_ = static_cast<T*>(p);
        is_allocated_ = true;
    }
    void free() {
        if (!is_allocated_) return;
        MANS_CHECK(mans_free(buffer_device_));
        buffer_device_ = nullptr;
        is_allocated_ = false;
    }
    void cpy_host_to_device() {
        if (!is_allocated_) allocate();
        MANS_CHECK(mans_memcpy_h2d(buffer_device_, buffer_, size_ * sizeof(T), nullptr));
    }
    void cpy_device_to_host() {
        MANS_CHECK(mans_memcpy_d2h(buffer_, buffer_device_, size_ * sizeof(T), nullptr));
    }

private:
    T* buffer_;
    T* buffer_device_ = nullptr;
    bool is_allocated_ = false;
    size_t size_;
};

}  // namespace cuhd

#endif
