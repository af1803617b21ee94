// cuhd_gpu_codetable.h -- device copy of a CUHDCodetable
// (reference include/cuhd_gpu_codetable.h:18-24).
#ifndef CUHD_GPU_CODETABLE_
#define CUHD_GPU_CODETABLE_

#include "cuhd_codetable.h"
#include "cuhd_gpu_memory_buffer.h"

namespace cuhd {
class CUHDGPUCodetable : public CUHDGPUMemoryBuffer<CUHDCodetableItem> {
public:
    explicit CUHDGPUCodetable(std::shared_ptr<CUHDCodetable> codetable);

private:
    std::shared_ptr<CUHDCodetable> table_;
};
}  // namespace cuhd

#endif
