// Host-side thread pool of the upload path (no device code).
#pragma once

#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

namespace hbt {

// Host threads of the upload path, started once per process: a per-buffer
// upload is a few milliseconds of host work, and starting threads for each
// of its passes would cost as much as the passes themselves.  run() hands
// tasks 0..nt-1 to the workers and the caller; while the pool is busy (a
// nested call, another thread's upload: the server's concurrent requests)
// or in a forked child, the call starts threads of its own instead.
class HostPool {
  public:
    static HostPool &get() {
        static HostPool pool;
        return pool;
    }
    int threads() const { return (int)workers_.size() + 1; }
    void run(int nt, const std::function<void(int)> &fn) {
        std::unique_lock<std::mutex> busy(run_mu_, std::try_to_lock);
        if (!busy.owns_lock() || getpid() != pid_ || workers_.empty()) {
            std::vector<std::thread> th;
            for (int t = 1; t < nt; t++) th.emplace_back(fn, t);
            fn(0);
            for (auto &x : th) x.join();
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = &fn;
            ntask_ = nt;
            next_ = 0;
            left_ = nt;
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [&] { return left_ == 0; });
        fn_ = nullptr;
        // a task that threw: every other task has finished, so fn is no
        // longer referenced when the exception leaves
        if (error_) std::rethrow_exception(std::exchange(error_, nullptr));
    }

  private:
    HostPool() : pid_(getpid()) {
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        for (unsigned t = 1; t < hw; t++) workers_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
            gen_++;
        }
        cv_.notify_all();
        for (auto &t : workers_)
            if (t.joinable()) t.join();
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
        }
    }
    // takes tasks of the current run until none is left
    void work() {
        for (;;) {
            const std::function<void(int)> *f;
            int t;
            {
                std::lock_guard<std::mutex> g(mu_);
                if (!fn_ || next_ >= ntask_) return;
                t = next_++;
                f = fn_;
            }
            std::exception_ptr e;
            try {
                (*f)(t);
            } catch (...) {
                e = std::current_exception();
            }
            std::lock_guard<std::mutex> g(mu_);
            if (e && !error_) error_ = e;
            if (--left_ == 0) done_.notify_all();
        }
    }
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> workers_;
    const std::function<void(int)> *fn_ = nullptr;
    int ntask_ = 0, next_ = 0, left_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::exception_ptr error_;
    const pid_t pid_;
};

// Runs fn(begin, end) over [0, n) split into contiguous ranges on up to 16
// host threads (one when n is small).
template <typename F>
inline void parallel_ranges(long long n, long long min_per_thread, F fn) {
    HostPool &pool = HostPool::get();
    const long long nt =
        std::max(1ll, std::min<long long>(pool.threads(), n / std::max(1ll, min_per_thread)));
    if (nt <= 1) {
        fn(0ll, n);
        return;
    }
    pool.run((int)nt, [&](int t) { fn(n * t / nt, n * (t + 1) / nt); });
}

}  // namespace hbt
