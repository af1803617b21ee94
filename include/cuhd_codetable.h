// cuhd_codetable.h -- decoder table, L rows of 4 bytes.
// Drop-in for the reference's CUHDCodetable (include/cuhd_codetable.h:20-45); the row
// layout {u16 next_state, u8 symbol, u8 min_num_bits} is the wire format the GPU reads.
#ifndef CUHD_CODETABLE_
#define CUHD_CODETABLE_

#include <cstddef>
#include <cstring>
#include <memory>

#include "cuhd_constants.h"
#include "cuhd_definitions.h"

struct CUHDCodetableItem {
    std::uint16_t next_state;
    std::uint8_t symbol;
    std::uint8_t min_num_bits;
};
static_assert(sizeof(CUHDCodetableItem) == 4, "table rows are 4 bytes");

class CUHDCodetable {
public:
    explicit CUHDCodetable(size_t num_entries);
    size_t get_size();
    size_t get_num_entries();
    size_t get_max_codeword_length();
    CUHDCodetableItem* get();

private:
    size_t size_;
    size_t num_entries_;
    cuhd_buf(CUHDCodetableItem, table_);
};

#endif
