// cuhd_subsequence_sync_point.h -- 16-byte sync record of the reference
// (include/cuhd_subsequence_sync_point.h:17-22): per subsequence, the unit (relative to the
// subsequence) and bit where its last step starts, the symbols before its first step, and
// the state before its last step.  CUHDGPUDecoderMemory::get_sync_info() returns an array
// of these for the last decode (filled from this decoder's proven checkpoints).
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
