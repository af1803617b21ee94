// cuhd_util.h -- timing and comparison helpers (reference include/cuhd_util.h:20-44).
// time() reports microseconds at microsecond resolution (the reference truncates to
// whole milliseconds, src/multians_util.cc:59).
#ifndef CUHD_UTIL_
#define CUHD_UTIL_

#include <chrono>
#include <cstddef>
#include <functional>
#include <string>
#include <utility>

#include "cuhd_constants.h"

namespace cuhd {
class CUHDUtil {
public:
    // runs f and returns (label, elapsed microseconds); device work enqueued by f is
    // included (the device is synchronised before the clock stops)
    static std::pair<std::string, size_t> time(std::string s, std::function<void()> f);
#define TIMER_START(vec, label) vec.push_back(cuhd::CUHDUtil::time(label, [&]() {
#define TIMER_STOP }));
    static bool equals(SYMBOL_TYPE* a, SYMBOL_TYPE* b, size_t size);
#define SDIV(n, m) (((n) + (m) - 1) / (m))
    static size_t optimal_subsequence_size(size_t input_size, size_t output_size, size_t pref,
                                           size_t device_pref);
};
}  // namespace cuhd

#endif
