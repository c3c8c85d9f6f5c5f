// Shared host-side helpers: error reporting, HIP checks, a small parallel-for.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wcsph_abi.h"

namespace wcsph {

void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define WCSPH_HIP(call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            throw ::wcsph::Error{WCSPH_EHIP, std::string(#call) + ": " +             \
                                                 hipGetErrorString(e_)};             \
    } while (0)

#define WCSPH_REQUIRE(cond, code, msg)                                               \
    do {                                                                             \
        if (!(cond)) throw ::wcsph::Error{(code), (msg)};                            \
    } while (0)

// Run fn and convert exceptions into a return code plus wcsph_last_error().
template <class F>
int guarded(F&& fn) {
    try {
        fn();
        return WCSPH_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return WCSPH_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return WCSPH_EINVAL;
    }
}

// Static partition of [0, n) over up to `threads` host threads.
template <class F>
void parallel_for(int64_t n, F&& fn, int threads = 0) {
    if (threads <= 0) threads = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (n < 4096 || threads == 1) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&fn, b, e] {
            for (int64_t i = b; i < e; ++i) fn(i);
        });
    }
    for (auto& th : pool) th.join();
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace wcsph
