// ans_encoder_table.h -- encoder table (reference include/ans_encoder_table.h:18-43).
#ifndef ANS_ENCODER_TABLE_
#define ANS_ENCODER_TABLE_

#include <cstddef>
#include <memory>
#include <vector>

#include "cuhd_constants.h"

struct ANSEncoderTable {
    struct ANSEncoderTableItem {
        std::uint32_t next_state;     // state after encoding, in [L, 2L)
        std::uint32_t code_sequence;  // bits emitted (MSB-first)
        std::uint32_t code_length;    // number of bits emitted
        SYMBOL_TYPE symbol;
    };
    size_t max_number_of_symbols;
    size_t number_of_symbols;
    size_t number_of_states;
    std::vector<std::vector<ANSEncoderTableItem>> table;   // [symbol][state - L]
};

#endif
