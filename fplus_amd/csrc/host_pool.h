// host_pool.h -- a small pool of host threads for the driver's parallel host loops: run(n, fn) calls fn(i, tid) for
// every i < n across the pool (the calling thread takes part as tid 0) and returns when all are done.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace fpa {

class HostPool {
public:
    explicit HostPool(int max_threads = 16)
    {
        const int hc = (int)std::thread::hardware_concurrency();
        nt_ = std::max(1, std::min(max_threads, hc > 0 ? hc : 1));
        for (int t = 1; t < nt_; t++) workers_.emplace_back([this, t]() { loop(t); });
    }
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread& w : workers_) w.join();
    }
    HostPool(const HostPool&) = delete;
    HostPool& operator=(const HostPool&) = delete;
    int size() const { return nt_; }
    void run(int64_t n, const std::function<void(int64_t, int)>& fn)
    {
        if (n <= 0) return;
        if (nt_ == 1 || n == 1) {
            for (int64_t i = 0; i < n; i++) fn(i, 0);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            busy_ = nt_ - 1;
            err_ = nullptr;
            gen_++;
        }
        cv_.notify_all();
        work(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this]() { return busy_ == 0; });
        fn_ = nullptr;
        if (err_) std::rethrow_exception(err_);
    }

private:
    void work(int tid)
    {
        for (;;) {
            const int64_t i = next_.fetch_add(1);
            if (i >= n_) break;
            try {
                (*fn_)(i, tid);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu_);
                if (!err_) err_ = std::current_exception();
                next_.store(n_);
            }
        }
    }
    void loop(int tid)
    {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&]() { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work(tid);
            std::lock_guard<std::mutex> lk(mu_);
            if (--busy_ == 0) done_cv_.notify_one();
        }
    }
    int nt_ = 1;
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int64_t, int)>* fn_ = nullptr;
    int64_t n_ = 0;
    std::atomic<int64_t> next_{0};
    int busy_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::exception_ptr err_;
};

}  // namespace fpa
