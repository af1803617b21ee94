// cuhd_constants.h -- type widths of the stream format.
// Same names and widths as the reference (include/cuhd_constants.h:14-27), which the
// drop-in classes and callers shaped like src/main.cc rely on.
#ifndef CUHD_CONSTANTS_
#define CUHD_CONSTANTS_

#include <cstdint>

// longest codeword the reference expects (its value; ignored by this decoder)
#define MAX_CODEWORD_LENGTH 11
#define UNIT_TYPE std::uint32_t
#define STATE_TYPE std::uint32_t
#define SYMBOL_TYPE std::uint8_t
#define BIT_COUNT_TYPE std::uint8_t

#endif
