// mans_internal.h -- helpers shared by the library's translation units.
#ifndef MANS_INTERNAL_H_
#define MANS_INTERNAL_H_

#include "../../include/mans_c_api.h"

// Record a message for mans_last_error() on this thread; returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#endif
