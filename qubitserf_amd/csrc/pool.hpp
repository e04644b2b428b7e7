// A process-wide pool of host worker threads, started on first use and kept until the
// process ends. The batched rounds split small host jobs (a level's decomposition, a
// batch's solve tables, its matroid partitions, its problem builds) many times per batch,
// and starting and joining a thread per worker and job cost more than those jobs. A child
// process made by fork has none of the parent's threads: it starts its own on first use
// and never joins the parent's.
#pragma once

#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace qsa {

class HostPool {
  public:
    static HostPool& get() {
        static HostPool pool;
        return pool;
    }

    // fn(i) for every i in [0, n), on the calling thread and up to workers - 1 pool threads.
    // Every index runs even when some throw; the first exception (lowest index) is rethrown
    // afterwards. One job at a time: concurrent callers wait their turn, and fn must not
    // call run itself.
    void run(std::size_t n, std::size_t workers, const std::function<void(std::size_t)>& fn) {
        workers = std::min(workers, n);
        if (workers <= 1) {
            std::exception_ptr first;
            for (std::size_t i = 0; i < n; ++i) {
                try {
                    fn(i);
                } catch (...) {
                    if (!first) first = std::current_exception();
                }
            }
            if (first) std::rethrow_exception(first);
            return;
        }
        std::lock_guard<std::mutex> call(call_mu_);
        std::vector<std::exception_ptr> errs(n);
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (owner_ != getpid()) {   // first use, or a forked child: the threads are not ours
                threads_ = new std::vector<std::thread>();   // (a parent's are left, never joined)
                owner_ = getpid();
            }
            while (threads_->size() + 1 < workers) threads_->emplace_back([this] { loop(); });
            fn_ = &fn;
            errs_ = &errs;
            n_ = n;
            next_.store(0);
            helpers_ = workers - 1;
            left_ = workers - 1;
            ++gen_;
        }
        cv_.notify_all();
        drain();
        {
            std::unique_lock<std::mutex> lk(mu_);
            done_.wait(lk, [this] { return left_ == 0; });
            fn_ = nullptr;
            errs_ = nullptr;
        }
        for (auto& e : errs)
            if (e) std::rethrow_exception(e);
    }

    ~HostPool() {
        if (owner_ != getpid()) return;   // no threads here, or only a parent's
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : *threads_) t.join();
        delete threads_;
    }

  private:
    HostPool() = default;

    void drain() {
        for (std::size_t i; (i = next_++) < n_;) {
            try {
                (*fn_)(i);
            } catch (...) {
                (*errs_)[i] = std::current_exception();
            }
        }
    }

    void loop() {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || (gen_ != seen && helpers_ > 0); });
                if (stop_) return;
                seen = gen_;
                --helpers_;   // this thread takes one helper slot of the job
            }
            drain();
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--left_ == 0) done_.notify_one();
            }
        }
    }

    std::mutex call_mu_;                // one job at a time
    std::mutex mu_;
    std::condition_variable cv_, done_;
    std::vector<std::thread>* threads_ = nullptr;
    pid_t owner_ = 0;                   // the process whose threads threads_ holds
    const std::function<void(std::size_t)>* fn_ = nullptr;
    std::vector<std::exception_ptr>* errs_ = nullptr;
    std::size_t n_ = 0;
    std::atomic<std::size_t> next_{0};
    std::size_t helpers_ = 0, left_ = 0;   // slots still to take, helpers still running
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

// Worker count for a host job of n items: up to 16 threads (the GPU box's CPU share),
// a single thread below `min_items`.
inline std::size_t host_workers(std::size_t n, std::size_t min_items) {
    if (n < min_items) return 1;
    return std::min<std::size_t>(16, std::max(1u, std::thread::hardware_concurrency()));
}

}  // namespace qsa
