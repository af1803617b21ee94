// mans_internal.h -- helpers shared by the library's translation units.
#ifndef MANS_INTERNAL_H_
#define MANS_INTERNAL_H_

#include "../../include/mans_c_api.h"

// Record a message for mans_last_error() on this thread; returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Rounds the fixup kernels (decoder and encoder) may run per stream before they report
// non-convergence (mans_set_sync_round_limit; 0 = the proven bound, n_chunks + 1).
uint32_t sync_round_limit();

#endif
