// cuhd_gpu_decoder_memory.h -- decoder workspace.
// Drop-in for the reference's CUHDGPUDecoderMemory (include/cuhd_gpu_decoder_memory.h:20-45,
// src/multians_gpu_decoder_memory.cc:13-92).  The constructor arguments keep their
// meaning; the allocation is this decoder's workspace (sync-chunk records, checkpoints,
// offsets and status words) sized by mans_workspace_bytes for num_units, which is
// valid for any subsequence size and block size passed to CUHDGPUDecoder::decode.
// The accessors keep their names; their arrays are laid out by this decoder:
//   get_sync_info()          workspace base (device)
//   get_output_sizes()       per-chunk output offsets region (device)
//   get_sequence_synced_*()  status words (device) and their pinned host mirror,
//                            filled by retrieve_sync_data()
#ifndef CUHD_GPU_DECODER_MEMORY_
#define CUHD_GPU_DECODER_MEMORY_

#include <cstddef>
#include <cstdint>

#include "cuhd_constants.h"
#include "cuhd_subsequence_sync_point.h"
#include "cuhd_util.h"

namespace cuhd {
class CUHDGPUDecoderMemory {
public:
    CUHDGPUDecoderMemory(size_t num_units, size_t subsequence_size, size_t num_threads);
    ~CUHDGPUDecoderMemory();
    void allocate();
    void free();
    void retrieve_sync_data();
    cuhd::CUHDSubsequenceSyncPoint* get_sync_info();
    std::uint32_t* get_output_sizes();
    std::uint8_t* get_sequence_synced_device();
    std::uint8_t* get_sequence_synced_host();

    void* workspace() { return workspace_; }
    std::uint64_t workspace_bytes() const { return bytes_; }
    size_t num_units() const { return num_units_; }

private:
    const size_t num_units_;
    const size_t num_subseq_;
    void* workspace_ = nullptr;
    std::uint8_t* status_host_ = nullptr;
    std::uint64_t bytes_ = 0;
    bool is_allocated_ = false;
};
}  // namespace cuhd

#endif
