// cuhd_gpu_decoder_memory.h -- decoder workspace.
// Drop-in for the reference's CUHDGPUDecoderMemory (include/cuhd_gpu_decoder_memory.h:20-45,
// src/multians_gpu_decoder_memory.cc:13-92).  The constructor arguments keep their
// meaning (units, subsequence size S, threads per block).  allocate() takes this decoder's
// workspace (sync-chunk records, checkpoints, offsets and status words, sized by
// mans_workspace_bytes for num_units) and the reference's own arrays, zero-filled:
//   get_sync_info()           CUHDSubsequenceSyncPoint per subsequence of S units (device)
//   get_output_sizes()        the exclusive scan of symbols per subsequence (device)
//   get_sequence_synced_*()   one flag per block of `threads` subsequences (device, and its
//                             pinned host mirror filled by retrieve_sync_data())
// After a CUHDGPUDecoder::decode with this memory, the first call of an accessor (or of
// retrieve_sync_data) fills them for that decode from the workspace (mans_ref_sync_fill:
// the contents the reference leaves there, see mans_c_api.h mans_ref_sync); decodes that
// nobody inspects pay nothing for them.
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
    size_t subsequence_size() const { return subsequence_size_; }
    size_t num_threads() const { return num_threads_; }
    // called by CUHDGPUDecoder::decode: the decode whose sync state the accessors describe
    void note_decode(const void* decode_args);

private:
    void fill();
    const size_t num_units_;
    const size_t subsequence_size_;
    const size_t num_threads_;
    const size_t num_subseq_;
    const size_t num_blocks_;
    void* workspace_ = nullptr;
    std::uint64_t bytes_ = 0;
    cuhd::CUHDSubsequenceSyncPoint* sync_info_ = nullptr;
    std::uint32_t* output_sizes_ = nullptr;
    std::uint8_t* synced_device_ = nullptr;
    std::uint8_t* synced_host_ = nullptr;
    alignas(8) unsigned char last_[256];     // mans_decode_args of the last decode
    bool have_last_ = false;
    bool filled_ = false;
    bool is_allocated_ = false;
};
}  // namespace cuhd

#endif
