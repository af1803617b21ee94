// ans_encoder.h -- single-state tANS encoder producing the decoder's input format.
// Drop-in for the reference's ANSEncoder (include/ans_encoder.h:30-45,
// src/ans_encoder.cc:74-97); the final unit's don't-care bits are zero.
#ifndef ANS_ENCODER_
#define ANS_ENCODER_

#include <memory>

#include "ans_encoder_table.h"
#include "ans_table_generator.h"
#include "cuhd_constants.h"
#include "cuhd_input_buffer.h"

struct Decoder_Info {
    UNIT_TYPE state;
    size_t bit;
    size_t size;   // encoded size in units
};

class ANSEncoder {
public:
    static std::shared_ptr<CUHDInputBuffer> encode(SYMBOL_TYPE* in, size_t size_in,
                                                   std::shared_ptr<ANSEncoderTable> encoder_table);
};

#endif
