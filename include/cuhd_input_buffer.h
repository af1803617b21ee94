// cuhd_input_buffer.h -- compressed stream in decoder order.
// Drop-in for the reference's CUHDInputBuffer (include/cuhd_input_buffer.h:19-48,
// src/multians_input_buffer.cc:15-30): the constructor takes units in encoder order,
// stores them reversed and pads 4 zero units.
#ifndef CUHD_INPUT_BUFFER_
#define CUHD_INPUT_BUFFER_

#include <cstddef>
#include <memory>

#include "cuhd_constants.h"
#include "cuhd_definitions.h"

class CUHDInputBuffer {
public:
    CUHDInputBuffer(UNIT_TYPE* buffer, size_t size, size_t first_bit, size_t first_state);
    UNIT_TYPE* get_compressed_data();
    size_t get_first_bit();
    size_t get_first_state();
    size_t get_compressed_size();
    size_t get_unit_size();

private:
    size_t first_bit_;
    size_t first_state_;
    size_t compressed_size_;
    const size_t unit_size_ = sizeof(UNIT_TYPE);
    cuhd_buf(UNIT_TYPE, buffer_);
};

#endif
