// cuhd_subsequence_sync_point.h -- 16-byte sync record of the reference
// (include/cuhd_subsequence_sync_point.h:17-22), kept for source compatibility.
// This decoder's workspace does not hold an array of these (see cuhd_gpu_decoder_memory.h).
#ifndef CUHD_SUBSEQUENCE_SYNC_POINT_
#define CUHD_SUBSEQUENCE_SYNC_POINT_

#include "cuhd_constants.h"

namespace cuhd {
struct CUHDSubsequenceSyncPoint {
    std::uint32_t unit;
    std::uint32_t bit;
    std::uint32_t output_size;
    std::uint32_t state;
};
}  // namespace cuhd

#endif
