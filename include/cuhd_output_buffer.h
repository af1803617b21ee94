// cuhd_output_buffer.h -- decoded symbols in decode order.
// Drop-in for the reference's CUHDOutputBuffer (include/cuhd_output_buffer.h:19-39).
#ifndef CUHD_OUTPUT_BUFFER_
#define CUHD_OUTPUT_BUFFER_

#include <cstddef>
#include <memory>

#include "cuhd_constants.h"
#include "cuhd_definitions.h"

class CUHDOutputBuffer {
public:
    explicit CUHDOutputBuffer(size_t size);
    std::shared_ptr<SYMBOL_TYPE[]>& get_decompressed_data();
    void reverse();   // decode order -> original order
    size_t get_uncompressed_size();
    size_t get_symbol_size();

private:
    size_t uncompressed_size_;
    const size_t symbol_size_ = sizeof(SYMBOL_TYPE);
    cuhd_out_buf(SYMBOL_TYPE, buffer_);
};

#endif
