// Shared host-side helpers: error reporting, HIP checks, a small parallel-for.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <exception>
#include <functional>
#include <mutex>
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

// Persistent host worker threads for parallel_for: created on first use and
// reused, since creating threads per call costs milliseconds. One job runs
// at a time (callers on other threads wait); a job started from inside a
// running job, on a worker or on the calling thread, runs inline.
class HostPool {
   public:
    static HostPool& get() {
        static HostPool pool;
        return pool;
    }
    int size() const { return (int)workers_.size() + 1; }
    static bool in_worker() { return worker_flag(); }
    // Runs task(k) for k in [0, parts), part 0 on the calling thread.
    // An exception from any part is rethrown here once every part is done.
    void run(int parts, const std::function<void(int)>& task) {
        std::lock_guard<std::mutex> job(job_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            task_ = &task;
            parts_ = parts;
            remaining_ = parts - 1;
            error_ = nullptr;
            ++gen_;
        }
        cv_start_.notify_all();
        std::exception_ptr mine;
        {  // a parallel_for nested in part 0 runs inline instead of waiting on this job
            const bool was = worker_flag();
            worker_flag() = true;
            try {
                task(0);
            } catch (...) {
                mine = std::current_exception();
            }
            worker_flag() = was;
        }
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return remaining_ == 0; });
        task_ = nullptr;
        if (!mine) mine = error_;
        lk.unlock();
        if (mine) std::rethrow_exception(mine);
    }

   private:
    HostPool() {
        const int n = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
        for (int w = 1; w < n; ++w) workers_.emplace_back([this, w] { loop(w); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_start_.notify_all();
        for (auto& t : workers_) t.join();
    }
    static bool& worker_flag() {
        static thread_local bool f = false;
        return f;
    }
    void loop(int w) {
        worker_flag() = true;
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)>* task;
            int parts;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_start_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                task = task_;
                parts = parts_;
            }
            if (w < parts) {
                std::exception_ptr e;
                try {
                    (*task)(w);
                } catch (...) {
                    e = std::current_exception();
                }
                std::lock_guard<std::mutex> lk(mu_);
                if (e && !error_) error_ = e;
                if (--remaining_ == 0) cv_done_.notify_one();
            }
        }
    }
    std::vector<std::thread> workers_;
    std::mutex job_mu_, mu_;
    std::condition_variable cv_start_, cv_done_;
    const std::function<void(int)>* task_ = nullptr;
    std::exception_ptr error_;
    int parts_ = 0, remaining_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Static partition of [0, n) over up to `threads` host threads (0: all of
// the pool's), on the persistent pool; fewer than min_n items run inline.
template <class F>
void parallel_for(int64_t n, F&& fn, int threads = 0, int64_t min_n = 4096) {
    HostPool& pool = HostPool::get();
    if (threads <= 0 || threads > pool.size()) threads = pool.size();
    if (n < min_n || threads == 1 || HostPool::in_worker()) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    const int parts = (int)std::min<int64_t>(threads, n);
    const int64_t chunk = (n + parts - 1) / parts;
    const std::function<void(int)> task = [&](int k) {
        const int64_t b = k * chunk, e = std::min(n, b + chunk);
        for (int64_t i = b; i < e; ++i) fn(i);
    };
    pool.run(parts, task);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace wcsph
