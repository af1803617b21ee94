// cuhd_gpu_input_buffer.h -- device copy of a CUHDInputBuffer, compressed size + 4 units
// (reference include/cuhd_gpu_input_buffer.h, src/multians_gpu_input_buffer.cc:13-19).
#ifndef CUHD_GPU_INPUT_BUFFER_
#define CUHD_GPU_INPUT_BUFFER_

#include "cuhd_gpu_memory_buffer.h"
#include "cuhd_input_buffer.h"

namespace cuhd {
class CUHDGPUInputBuffer : public CUHDGPUMemoryBuffer<UNIT_TYPE> {
public:
    explicit CUHDGPUInputBuffer(std::shared_ptr<CUHDInputBuffer> input_buffer);

private:
    std::shared_ptr<CUHDInputBuffer> input_buffer_;
};
}  // namespace cuhd

#endif
