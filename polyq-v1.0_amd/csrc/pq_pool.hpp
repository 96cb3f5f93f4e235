// A small fixed pool of host threads for the per-query prologues of batches.
// run(n, min_chunk, max_chunks, fn) splits [0, n) into contiguous chunks, calls
// fn(lo, hi, chunk_index) for each (the caller takes part) and returns the
// number of chunks; chunk c covers a range that precedes chunk c + 1, so
// results gathered per chunk and concatenated in chunk order are in query
// order whatever the thread timing.  One parallel region runs at a time; a
// caller that finds the pool busy runs its chunks itself.  The first few
// workers poll for a new region for a short while after each one before they
// block, so a caller that issues small batches back to back does not pay a
// thread wake-up per batch; the others block at once.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
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

    // Workers that poll for new regions (the rest block between regions).
    static constexpr unsigned kHot = 3;

    // At most max_chunks chunks of at least min_chunk items each.  More
    // chunks than threads let workers that wake late still take a share.
    unsigned run(size_t n, size_t min_chunk, unsigned max_chunks, const std::function<void(size_t, size_t, unsigned)>& fn) {
        const size_t cap = std::max<size_t>(1, max_chunks);
        const unsigned want = (unsigned)std::max<size_t>(1, std::min<size_t>(cap, n / std::max<size_t>(min_chunk, 1)));
        if (want <= 1) { fn(0, n, 0); return 1; }
        std::unique_lock<std::mutex> region(region_mu_, std::try_to_lock);
        if (!region.owns_lock()) { fn(0, n, 0); return 1; }
        auto task = std::make_shared<Task>();
        task->fn = &fn; task->n = n; task->chunks = want;
        {
            std::lock_guard<std::mutex> lk(mu_);
            current_ = task;
            gen_.fetch_add(1, std::memory_order_release);
        }
        // Polling workers see the new generation by themselves (a worker
        // counted here checks the generation again before it blocks), so
        // only the chunks they cannot cover wake a blocked worker: in steady
        // back-to-back use no wake-up system call is made at all.
        const unsigned helpers = std::min<unsigned>(want - 1, nworkers_), polling = polling_.load(std::memory_order_acquire);
        for (unsigned i = polling; i < helpers; ++i) cv_.notify_one();
        work(*task);
        while (task->done.load(std::memory_order_acquire) != task->chunks) std::this_thread::yield();
        std::lock_guard<std::mutex> lk(mu_);
        if (current_ == task) current_.reset();
        return want;
    }

private:
    struct Task {
        const std::function<void(size_t, size_t, unsigned)>* fn = nullptr;
        size_t n = 0;
        unsigned chunks = 0;
        std::atomic<unsigned> next{0}, done{0};
    };
    // POLYQ_AMD_HOST_THREADS sets the pool size (caller included), default
    // min(16, hardware threads); 1 runs every prologue on the calling thread.
    ChunkPool() {
        unsigned hw = std::thread::hardware_concurrency();
        unsigned n = std::min<unsigned>(hw ? hw : 1, 16);
        if (const char* e = std::getenv("POLYQ_AMD_HOST_THREADS")) {
            const long v = std::strtol(e, nullptr, 10);
            if (v >= 1 && v <= 256) n = (unsigned)v;
        }
        nworkers_ = n - 1;
        for (unsigned i = 0; i < nworkers_; ++i) std::thread([this, i] { loop(i < kHot); }).detach();
    }
    static void work(Task& t) {
        for (;;) {
            const unsigned c = t.next.fetch_add(1);
            if (c >= t.chunks) return;
            (*t.fn)(t.n * c / t.chunks, t.n * (c + 1) / t.chunks, c);
            t.done.fetch_add(1, std::memory_order_release);
        }
    }
    void loop(bool hot) {
        uint64_t seen = 0;
        for (;;) {
            // hot workers poll for up to 200 us, then block
            if (hot) {
                polling_.fetch_add(1, std::memory_order_acq_rel);
                const auto t0 = std::chrono::steady_clock::now();
                while (gen_.load(std::memory_order_acquire) == seen &&
                       std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(200))
                    std::this_thread::yield();
                polling_.fetch_sub(1, std::memory_order_acq_rel);
            }
            std::shared_ptr<Task> t;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_.load(std::memory_order_relaxed) != seen; });
                seen = gen_.load(std::memory_order_relaxed);
                t = current_;
            }
            if (t) work(*t);
        }
    }
    unsigned nworkers_ = 0;
    std::mutex region_mu_, mu_;
    std::condition_variable cv_;
    std::atomic<uint64_t> gen_{0};
    std::atomic<unsigned> polling_{0};   // hot workers in their polling loop
    std::shared_ptr<Task> current_;
};

}  // namespace pq
