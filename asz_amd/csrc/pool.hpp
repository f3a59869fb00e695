// A small fork-join pool of host threads for the host-side scoring and
// lossless-size work of plan generation and family selection: run(n, f)
// calls f(0) .. f(n-1) on the workers and the calling thread and returns when
// all are done.  The first exception thrown by a task is rethrown by run().
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace asz {

class TaskPool {
public:
    explicit TaskPool(unsigned workers) {
        for (unsigned i = 0; i < workers; i++) th_.emplace_back([this] { loop(); });
    }
    ~TaskPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    TaskPool(const TaskPool&) = delete;
    TaskPool& operator=(const TaskPool&) = delete;

    void run(size_t n, const std::function<void(size_t)>& f) {
        if (n == 0) return;
        if (n == 1 || th_.empty()) {
            for (size_t i = 0; i < n; i++) f(i);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &f;
            n_ = n;
            next_.store(0);
            done_ = 0;
            err_ = nullptr;
            gen_++;
        }
        cv_.notify_all();
        work(&f, n);
        std::unique_lock<std::mutex> lk(mu_);
        // every worker that joined this batch has left it before the next
        // batch can reset the shared index
        done_cv_.wait(lk, [&] { return done_ == n_ && active_ == 0; });
        fn_ = nullptr;
        if (err_) std::rethrow_exception(err_);
    }

    // Asynchronous form: the workers alone run f(0) .. f(n-1) (f is copied);
    // wait() blocks until they are done and rethrows the first exception.
    // No other batch may be started in between.
    void start(size_t n, std::function<void(size_t)> f) {
        owned_ = std::move(f);
        if (th_.empty()) {
            // no workers: run inline
            for (size_t i = 0; i < n; i++) owned_(i);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &owned_;
            n_ = n;
            next_.store(0);
            done_ = 0;
            err_ = nullptr;
            gen_++;
        }
        cv_.notify_all();
    }
    void wait() {
        if (th_.empty()) return;
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return fn_ == nullptr || (done_ == n_ && active_ == 0); });
        fn_ = nullptr;
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }
    unsigned workers() const { return (unsigned)th_.size(); }

    static unsigned default_workers() {
        const unsigned hc = std::thread::hardware_concurrency();
        return std::max(1u, std::min(hc ? hc : 1u, 8u) - 1);
    }

private:
    void work(const std::function<void(size_t)>* f, size_t n) {
        for (;;) {
            const size_t i = next_.fetch_add(1);
            if (i >= n) return;
            try {
                (*f)(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu_);
                if (!err_) err_ = std::current_exception();
            }
            std::lock_guard<std::mutex> lk(mu_);
            if (++done_ == n) done_cv_.notify_all();
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(size_t)>* f;
            size_t n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || (gen_ != seen && fn_ != nullptr); });
                if (stop_) return;
                seen = gen_;
                f = fn_;
                n = n_;
                active_++;
            }
            work(f, n);
            std::lock_guard<std::mutex> lk(mu_);
            active_--;
            done_cv_.notify_all();
        }
    }

    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)>* fn_ = nullptr;
    std::function<void(size_t)> owned_;
    size_t n_ = 0, done_ = 0, active_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::exception_ptr err_;
};

}  // namespace asz
