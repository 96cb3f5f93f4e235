// A small fixed pool of host threads for the per-query prologues of large
// batches.  run(n, min_chunk, fn) splits [0, n) into contiguous chunks, calls
// fn(lo, hi, chunk_index) for each (the caller takes part) and returns the
// number of chunks; chunk c covers a range that precedes chunk c + 1, so
// results gathered per chunk and concatenated in chunk order are in query
// order whatever the thread timing.  One parallel region runs at a time; a
// caller that finds the pool busy runs its chunks itself.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace pq {

class ChunkPool {
public:
    static ChunkPool& get() {
        static ChunkPool* p = new ChunkPool();   // never destroyed: workers may outlive static teardown
        return *p;
    }
    unsigned threads() const { return nworkers_ + 1; }

    unsigned run(size_t n, size_t min_chunk, const std::function<void(size_t, size_t, unsigned)>& fn) {
        const unsigned want = (unsigned)std::max<size_t>(1, std::min<size_t>(threads(), n / std::max<size_t>(min_chunk, 1)));
        if (want <= 1) { fn(0, n, 0); return 1; }
        std::unique_lock<std::mutex> region(region_mu_, std::try_to_lock);
        if (!region.owns_lock()) { fn(0, n, 0); return 1; }
        auto task = std::make_shared<Task>();
        task->fn = &fn; task->n = n; task->chunks = want;
        {
            std::lock_guard<std::mutex> lk(mu_);
            current_ = task;
        }
        cv_.notify_all();
        work(*task);
        {
            std::unique_lock<std::mutex> lk(task->mu);
            task->cv.wait(lk, [&] { return task->done == task->chunks; });
        }
        std::lock_guard<std::mutex> lk(mu_);
        if (current_ == task) current_.reset();
        return want;
    }

private:
    struct Task {
        const std::function<void(size_t, size_t, unsigned)>* fn = nullptr;
        size_t n = 0;
        unsigned chunks = 0;
        std::atomic<unsigned> next{0};
        std::mutex mu;
        std::condition_variable cv;
        unsigned done = 0;
    };
    ChunkPool() {
        unsigned hw = std::thread::hardware_concurrency();
        nworkers_ = std::min<unsigned>(hw ? hw : 1, 16) - 1;
        for (unsigned i = 0; i < nworkers_; ++i) std::thread([this] { loop(); }).detach();
    }
    static void work(Task& t) {
        for (;;) {
            const unsigned c = t.next.fetch_add(1);
            if (c >= t.chunks) return;
            (*t.fn)(t.n * c / t.chunks, t.n * (c + 1) / t.chunks, c);
            std::lock_guard<std::mutex> lk(t.mu);
            if (++t.done == t.chunks) t.cv.notify_all();
        }
    }
    void loop() {
        std::shared_ptr<Task> last;
        for (;;) {
            std::shared_ptr<Task> t;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return current_ && current_ != last; });
                t = current_;
            }
            last = t;
            work(*t);
        }
    }
    unsigned nworkers_ = 0;
    std::mutex region_mu_, mu_;
    std::condition_variable cv_;
    std::shared_ptr<Task> current_;
};

}  // namespace pq
