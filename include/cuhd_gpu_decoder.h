// cuhd_gpu_decoder.h -- GPU decoder entry point.
// Drop-in for the reference's cuhd::CUHDGPUDecoder::decode (include/cuhd_gpu_decoder.h:22-36).
// Same parameters and meaning:
//   input_size                 compressed units, without padding
//   output_size                symbols to produce
//   initial_state/initial_bit  the stream's first_state / first_bit
//   number_of_states           L (table rows)
//   max_codeword_length        ignored (the reference ignores it too)
//   preferred_subsequence_size, threads_per_block
//                              must equal the ones `aux` was built with (as the reference
//                              needs); they shape only the reference sync state that
//                              aux's accessors return.  This decoder picks its own
//                              sync-chunk and checkpoint lengths (results do not depend
//                              on them).
// Like the reference, decode() may return before the last kernel completes; the next
// blocking copy (cpy_device_to_host) or CUHDUtil::time synchronises.  Integrity faults
// found on the device (rejected table) are reported at the next
// CUHDGPUDecoderMemory::retrieve_sync_data() / decode_status() call.
#ifndef CUHD_GPU_DECODER_
#define CUHD_GPU_DECODER_

#include <memory>

#include "cuhd_constants.h"
#include "cuhd_gpu_codetable.h"
#include "cuhd_gpu_decoder_memory.h"
#include "cuhd_gpu_input_buffer.h"
#include "cuhd_gpu_output_buffer.h"
#include "cuhd_subsequence_sync_point.h"

namespace cuhd {
class CUHDGPUDecoder {
public:
    static void decode(std::shared_ptr<cuhd::CUHDGPUInputBuffer> input, size_t input_size,
                       std::shared_ptr<cuhd::CUHDGPUOutputBuffer> output, size_t output_size,
                       std::shared_ptr<cuhd::CUHDGPUCodetable> table,
                       std::shared_ptr<cuhd::CUHDGPUDecoderMemory> aux, STATE_TYPE initial_state,
                       std::uint32_t initial_bit, std::uint32_t number_of_states,
                       size_t max_codeword_length, size_t preferred_subsequence_size,
                       size_t threads_per_block);
    // 0 when the last decode that used `aux` found no fault, else a MANS_E* code
    static int decode_status(std::shared_ptr<cuhd::CUHDGPUDecoderMemory> aux);
};
}  // namespace cuhd

#endif
