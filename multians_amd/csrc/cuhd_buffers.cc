// cuhd_buffers.cc -- host data model, GPU buffers, decoder memory and utilities of the
// drop-in C++ layer (include/cuhd_*.h).  Device work goes through the C ABI only.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>

#include "cuhd_codetable.h"
#include "cuhd_gpu_codetable.h"
#include "cuhd_gpu_decoder_memory.h"
#include "cuhd_gpu_input_buffer.h"
#include "cuhd_gpu_output_buffer.h"
#include "cuhd_input_buffer.h"
#include "cuhd_output_buffer.h"
#include "cuhd_util.h"
#include "mans_c_api.h"

// ------------------------------------------------------------------ host model

CUHDCodetable::CUHDCodetable(size_t num_entries)
    : size_(num_entries), num_entries_(num_entries),
      table_(new CUHDCodetableItem[num_entries]()) {}
size_t CUHDCodetable::get_size() { return size_; }
size_t CUHDCodetable::get_num_entries() { return num_entries_; }
size_t CUHDCodetable::get_max_codeword_length() { return MAX_CODEWORD_LENGTH; }
CUHDCodetableItem* CUHDCodetable::get() { return table_.get(); }

CUHDInputBuffer::CUHDInputBuffer(UNIT_TYPE* buffer, size_t size, size_t first_bit,
                                 size_t first_state)
    : first_bit_(first_bit), first_state_(first_state), compressed_size_(size),
      buffer_(new UNIT_TYPE[size + 4]()) {
    // decoder order = encoder emission reversed; 4 zero units of padding
    std::reverse_copy(buffer, buffer + size, buffer_.get());
}
UNIT_TYPE* CUHDInputBuffer::get_compressed_data() { return buffer_.get(); }
size_t CUHDInputBuffer::get_first_bit() { return first_bit_; }
size_t CUHDInputBuffer::get_first_state() { return first_state_; }
size_t CUHDInputBuffer::get_compressed_size() { return compressed_size_; }
size_t CUHDInputBuffer::get_unit_size() { return unit_size_; }

CUHDOutputBuffer::CUHDOutputBuffer(size_t size)
    : uncompressed_size_(size), buffer_(new SYMBOL_TYPE[size]()) {}
std::shared_ptr<SYMBOL_TYPE[]>& CUHDOutputBuffer::get_decompressed_data() { return buffer_; }
void CUHDOutputBuffer::reverse() { std::reverse(buffer_.get(), buffer_.get() + uncompressed_size_); }
size_t CUHDOutputBuffer::get_uncompressed_size() { return uncompressed_size_; }
size_t CUHDOutputBuffer::get_symbol_size() { return symbol_size_; }

// ------------------------------------------------------------------ GPU buffers

namespace cuhd {

CUHDGPUCodetable::CUHDGPUCodetable(std::shared_ptr<CUHDCodetable> codetable)
    : CUHDGPUMemoryBuffer<CUHDCodetableItem>(codetable->get(), codetable->get_size()),
      table_(codetable) {}

CUHDGPUInputBuffer::CUHDGPUInputBuffer(std::shared_ptr<CUHDInputBuffer> input_buffer)
    : CUHDGPUMemoryBuffer<UNIT_TYPE>(input_buffer->get_compressed_data(),
                                     input_buffer->get_compressed_size() + 4),
      input_buffer_(input_buffer) {}

CUHDGPUOutputBuffer::CUHDGPUOutputBuffer(std::shared_ptr<CUHDOutputBuffer> output_buffer)
    : CUHDGPUMemoryBuffer<SYMBOL_TYPE>(output_buffer->get_decompressed_data().get(),
                                       output_buffer->get_uncompressed_size()),
      output_buffer_(output_buffer) {}

// ------------------------------------------------------------------ decoder memory

static_assert(sizeof(mans_decode_args) <= 256, "decode args fit the memory's copy");

CUHDGPUDecoderMemory::CUHDGPUDecoderMemory(size_t num_units, size_t subsequence_size,
                                           size_t num_threads)
    : num_units_(num_units),
      subsequence_size_(subsequence_size),
      num_threads_(num_threads),
      num_subseq_(subsequence_size ? SDIV(num_units, subsequence_size) : 0),
      num_blocks_(num_threads ? SDIV(num_subseq_, num_threads) : 0) {}

CUHDGPUDecoderMemory::~CUHDGPUDecoderMemory() { free(); }

void CUHDGPUDecoderMemory::allocate() {
    if (is_allocated_) return;
    if (!subsequence_size_ || !num_threads_) {
        std::printf("MANS error: decoder memory needs a subsequence size and a thread count\n");
        std::exit(1);
    }
    MANS_CHECK(mans_workspace_bytes(num_units_, 0, &bytes_));
    MANS_CHECK(mans_malloc(&workspace_, bytes_));
    // the reference's arrays (mans_malloc zero-fills, like its cudaMemset)
    void* p = nullptr;
    MANS_CHECK(mans_malloc(&p, std::max<size_t>(num_subseq_, 1) * sizeof(CUHDSubsequenceSyncPoint)));
    sync_info_ = static_cast<CUHDSubsequenceSyncPoint*>(p);
    MANS_CHECK(mans_malloc(&p, std::max<size_t>(num_subseq_, 1) * sizeof(std::uint32_t)));
    output_sizes_ = static_cast<std::uint32_t*>(p);
    MANS_CHECK(mans_malloc(&p, std::max<size_t>(num_blocks_, 1)));
    synced_device_ = static_cast<std::uint8_t*>(p);
    MANS_CHECK(mans_host_alloc_pinned(&p, std::max<size_t>(num_blocks_, 1)));
    synced_host_ = static_cast<std::uint8_t*>(p);
    std::memset(synced_host_, 0, std::max<size_t>(num_blocks_, 1));
    is_allocated_ = true;
}

void CUHDGPUDecoderMemory::note_decode(const void* decode_args) {
    std::memcpy(last_, decode_args, sizeof(mans_decode_args));
    have_last_ = true;
    filled_ = false;
}

void CUHDGPUDecoderMemory::fill() {
    if (!is_allocated_ || !have_last_ || filled_) return;
    mans_ref_sync r;
    std::memset(&r, 0, sizeof(r));
    r.sync_points = reinterpret_cast<std::uint32_t*>(sync_info_);
    r.output_sizes = output_sizes_;
    r.synced_flags = synced_device_;
    r.subsequence_units = (std::uint32_t)subsequence_size_;
    r.threads = (std::uint32_t)num_threads_;
    mans_decode_args a;
    std::memcpy(&a, last_, sizeof(a));
    a.ref_sync = &r;
    MANS_CHECK(mans_ref_sync_fill(&a));
    filled_ = true;
}

void CUHDGPUDecoderMemory::retrieve_sync_data() {
    fill();
    MANS_CHECK(mans_memcpy_d2h(synced_host_, synced_device_, std::max<size_t>(num_blocks_, 1),
                               nullptr));
}

void CUHDGPUDecoderMemory::free() {
    if (!is_allocated_) return;
    MANS_CHECK(mans_free(workspace_));
    MANS_CHECK(mans_free(sync_info_));
    MANS_CHECK(mans_free(output_sizes_));
    MANS_CHECK(mans_free(synced_device_));
    MANS_CHECK(mans_host_free_pinned(synced_host_));
    workspace_ = nullptr;
    sync_info_ = nullptr;
    output_sizes_ = nullptr;
    synced_device_ = nullptr;
    synced_host_ = nullptr;
    have_last_ = false;
    is_allocated_ = false;
}

CUHDSubsequenceSyncPoint* CUHDGPUDecoderMemory::get_sync_info() {
    fill();
    return sync_info_;
}
std::uint32_t* CUHDGPUDecoderMemory::get_output_sizes() {
    fill();
    return output_sizes_;
}
std::uint8_t* CUHDGPUDecoderMemory::get_sequence_synced_device() {
    fill();
    return synced_device_;
}
std::uint8_t* CUHDGPUDecoderMemory::get_sequence_synced_host() { return synced_host_; }

// ------------------------------------------------------------------ util

std::pair<std::string, size_t> CUHDUtil::time(std::string s, std::function<void()> f) {
    auto t0 = std::chrono::steady_clock::now();
    f();
    MANS_CHECK(mans_sync(nullptr));
    auto t1 = std::chrono::steady_clock::now();
    return {s, (size_t)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count()};
}

bool CUHDUtil::equals(SYMBOL_TYPE* a, SYMBOL_TYPE* b, size_t size) {
    for (size_t i = 0; i < size; ++i) {
        if (a[i] != b[i]) {
            std::printf("mismatch at: %zu\n", i);
            return false;
        }
    }
    return true;
}

size_t CUHDUtil::optimal_subsequence_size(size_t input_size, size_t output_size, size_t pref,
                                          size_t device_pref) {
    // the reference's heuristic (src/multians_util.cc:74-86): halve the preferred
    // size while the subsequence count per output word is below 1/device_pref
    float device = 1.0f / (float)device_pref;
    size_t cur = pref;
    while (cur > 1 && (float)SDIV(SDIV(input_size, 4), cur) / (float)SDIV(output_size, 4) < device)
        cur = SDIV(cur, 2);
    return cur;
}

}  // namespace cuhd
