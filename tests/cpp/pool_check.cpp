// Checks of the host worker pool behind parallel_for (weaklycompressiblesph_amd/csrc/common.hpp):
// every index runs exactly once, nested calls run inline, an exception in any
// part reaches the caller after all parts finish, and callers on several
// threads are served one job at a time. Exit status 0 on success.
#include <atomic>
#include <cstdio>
#include <stdexcept>

#include "../../weaklycompressiblesph_amd/csrc/common.hpp"

namespace wcsph {
void set_error(const std::string&) {}
}

static int failures = 0;
#define CHECK(c)                                                  \
    do {                                                          \
        if (!(c)) {                                               \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                           \
        }                                                         \
    } while (0)

int main() {
    using wcsph::parallel_for;
    // every index once, for sizes around the inline threshold
    for (int64_t n : {0L, 1L, 4095L, 4096L, 100000L, 1234567L}) {
        std::vector<std::atomic<int>> hits(n);
        for (auto& h : hits) h = 0;
        parallel_for(n, [&](int64_t i) { hits[i]++; });
        bool once = true;
        for (auto& h : hits) once = once && h == 1;
        CHECK(once);
    }
    // min_n lowers the threshold; nested calls run inline
    {
        std::atomic<int64_t> sum{0};
        parallel_for(64, [&](int64_t i) {
            parallel_for(5000, [&](int64_t j) { sum += (i + 1) * (j + 1) > 0 ? 1 : 0; });
        }, 0, 8);
        CHECK(sum == 64 * 5000);
    }
    // an exception in some part arrives after every part has finished
    for (int bad : {0, 7, 99999}) {
        std::atomic<int64_t> done{0};
        bool caught = false;
        try {
            parallel_for(100000, [&](int64_t i) {
                if (i == bad) throw std::runtime_error("planted");
                done++;
            });
        } catch (const std::runtime_error&) {
            caught = true;
        }
        CHECK(caught);
        CHECK(done == 100000 - 1 || done < 100000);  // other parts ran or stopped at their own pace
    }
    // the pool still works after a failed job
    {
        std::atomic<int64_t> c{0};
        parallel_for(50000, [&](int64_t) { c++; });
        CHECK(c == 50000);
    }
    // callers on several threads
    {
        std::atomic<int64_t> total{0};
        std::vector<std::thread> ts;
        for (int t = 0; t < 4; ++t)
            ts.emplace_back([&] {
                for (int r = 0; r < 20; ++r) parallel_for(20000, [&](int64_t) { total++; });
            });
        for (auto& t : ts) t.join();
        CHECK(total == 4LL * 20 * 20000);
    }
    std::printf("pool checks: %s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
