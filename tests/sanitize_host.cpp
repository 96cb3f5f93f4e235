// Host-side concurrency driver for the ThreadSanitizer run (tests/sanitize.sh):
// the thread pool of the batch prologues, concurrent simulator calls that
// build and evict marginal engines, and the request loop's host paths (content
// keys, worker builds, admission, host LRU), all without a device.  Built with
// -fsanitize=thread against a library whose host code is built the same way;
// the run fails on any report (TSAN_OPTIONS=halt_on_error=1).
#include "../include/polyq_amd.h"
#include "../polyq-v1.0_amd/csrc/pq_pool.hpp"
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

int fail(const char* what) {
    std::fprintf(stderr, "sanitize_host: %s: %s\n", what, pq_last_error());
    return 1;
}

// A small Clifford+T circuit: H layer, CZ/CX ring, T on every other qubit,
// H layer; `seed` varies the T placement so circuits differ in content.
struct Gates {
    std::vector<uint8_t> op;
    std::vector<uint32_t> q0, q1, ph;
    void add(uint8_t o, uint32_t a, uint32_t b = 0) { op.push_back(o); q0.push_back(a); q1.push_back(b); ph.push_back(0); }
};
Gates circuit(uint32_t n, uint32_t seed) {
    Gates g;
    for (uint32_t q = 0; q < n; ++q) g.add(PQ_H, q);
    for (uint32_t q = 0; q < n; ++q) g.add((q + seed) % 2 ? PQ_CZ : PQ_CX, q, (q + 1) % n);
    for (uint32_t q = seed % 2; q < n; q += 2) g.add(PQ_T, q);
    for (uint32_t q = 0; q < n; ++q) g.add(PQ_CZ, q, (q + 1 + seed % 3) % n);
    for (uint32_t q = 0; q < n; ++q) g.add(PQ_H, q);
    return g;
}

int pool_stress() {
    pq::ChunkPool& pool = pq::ChunkPool::get();
    std::atomic<int> bad{0};
    std::vector<std::thread> callers;
    for (int c = 0; c < 6; ++c)
        callers.emplace_back([&, c] {
            for (int rep = 0; rep < 200; ++rep) {
                const size_t n = 64 + 37 * (size_t)((rep + c) % 11);
                std::vector<uint64_t> part(pool.threads(), 0);
                const unsigned nch = pool.run(n, 8, pool.threads(), [&](size_t lo, size_t hi, unsigned k) {
                    uint64_t s = 0;
                    for (size_t i = lo; i < hi; ++i) s += i;
                    part[k] += s;
                });
                uint64_t tot = 0;
                for (unsigned k = 0; k < nch; ++k) tot += part[k];
                if (tot != (uint64_t)n * (n - 1) / 2) bad++;
            }
        });
    for (auto& t : callers) t.join();
    if (bad) { std::fprintf(stderr, "sanitize_host: pool sums wrong %d times\n", bad.load()); return 1; }
    return 0;
}

int sim_stress() {
    // concurrent host-only calls on one simulator: marginal engines are built
    // and, past the cache size, evicted while other threads hold theirs
    const Gates g = circuit(10, 1);
    pq_sim* sim = nullptr;
    if (pq_sim_create(10, g.op.size(), g.op.data(), g.q0.data(), g.q1.data(), g.ph.data(), PQ_CACHED, 0, &sim)) return fail("pq_sim_create");
    std::atomic<int> bad{0};
    std::vector<std::thread> th;
    for (int c = 0; c < 4; ++c)
        th.emplace_back([&, c] {
            uint64_t st[7];
            for (uint32_t k = 0; k < 40; ++k) {
                const uint32_t qs[2] = {(k + (uint32_t)c) % 10, (k * 3 + 1) % 10};
                if (qs[0] == qs[1]) continue;
                if (pq_sim_marginal_stats(sim, qs, 2, st)) bad++;
                uint64_t xs[4] = {k, k + 1, 3 * k, 7}, ys[4] = {0, k, 5, 1023}, nb = 0;
                if (pq_sim_prologue_selfcheck(sim, 4, xs, ys, &nb) || nb) bad++;
            }
        });
    for (auto& t : th) t.join();
    pq_sim_destroy(sim);
    if (bad) { std::fprintf(stderr, "sanitize_host: simulator calls failed %d times\n", bad.load()); return 1; }
    return 0;
}

int server_stress() {
    // requests from several threads on circuits with repeated and new
    // content, under a host budget small enough for the host LRU to drop
    // idle builds; without a device every request fails at upload, with one
    // they complete: either way each ticket ends
    pq_server* srv = nullptr;
    if (pq_server_create(1ull << 30, 4, PQ_FORM_Z8, &srv)) return fail("pq_server_create");
    pq_server_set_host_budget(srv, 1);
    std::atomic<int> lost{0};
    std::vector<std::thread> th;
    for (int c = 0; c < 4; ++c)
        th.emplace_back([&, c] {
            std::vector<uint64_t> tickets;
            for (uint32_t k = 0; k < 30; ++k) {
                const Gates g = circuit(8, (k + (uint32_t)c) % 5);
                const uint32_t mq[2] = {0, 3};
                const uint8_t mo[2] = {1, 0};
                const int kind = (int)(k % 3);
                uint64_t t = 0;
                if (pq_server_submit(srv, 8, g.op.size(), g.op.data(), g.q0.data(), g.q1.data(), g.ph.data(), PQ_CACHED, kind, k % 256,
                                     (k * 7) % 256, mq, mo, kind == 2 ? 2 : 0, &t) == 0)
                    tickets.push_back(t);
            }
            for (uint64_t t : tickets) {
                double out[2], lat;
                if (pq_server_wait(srv, t, 60.0, out, &lat) == 1) lost++;
            }
        });
    for (auto& t : th) t.join();
    double st[10], hs[3];
    pq_server_stats(srv, st);
    pq_server_host_stats(srv, hs);
    pq_server_destroy(srv);
    std::printf("server: completed %.0f failed %.0f builds %.0f host drops %.0f\n", st[0], st[1], st[5], hs[2]);
    if (lost) { std::fprintf(stderr, "sanitize_host: %d requests never ended\n", lost.load()); return 1; }
    if (st[0] + st[1] != 120) { std::fprintf(stderr, "sanitize_host: %.0f of 120 requests accounted\n", st[0] + st[1]); return 1; }
    return 0;
}

}  // namespace

int main() {
    if (pool_stress()) return 1;
    std::printf("pool: ok\n");
    if (sim_stress()) return 1;
    std::printf("simulator: ok\n");
    if (server_stress()) return 1;
    std::printf("server: ok\n");
    return 0;
}
