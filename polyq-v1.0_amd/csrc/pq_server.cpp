// Request loop of the branch evaluator: resident engines in HBM, admission by
// device memory, continuous batching.
//
//  * Circuits are keyed by content (a hash, confirmed by comparing the gate
//    list); an engine set is compiled and built on the host once
//    (src/sim.rs:144-148 + src/rank.rs:88-123), uploaded and kept resident
//    until evicted (LRU, only when the HBM budget is exceeded and the entry
//    has no request in flight).  Host builds are bounded too: past the host
//    byte budget the least recently used idle circuits are dropped and are
//    rebuilt if they come back.
//  * Worker threads run every request's host prologue (boundary, prepare,
//    Z8 query bits) and any build it needs, off the dispatch path.
//  * The dispatcher keeps up to two batches in flight on separate streams:
//    requests that become ready while one batch runs are launched as the next
//    batch immediately and start on the device as the first one drains.
//  * A request's result does not depend on the batch it lands in: its items
//    and reduction order depend only on its own branch dimension.
#include "pq_api_internal.hpp"
#include "../../include/polyq_amd.h"
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>
#include <algorithm>
#include <cmath>

namespace pq {
namespace {

double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// One circuit: the host build is kept while the entry is in the table (host
// LRU), the device images only while resident.  `pinned` counts requests whose
// device jobs exist (prepared, queued or running); only unpinned entries are
// evictable, and requests still waiting for a worker pin nothing.
struct Entry {
    uint64_t key = 0;
    Circuit circ;              // the content the key stands for
    int strategy = 0;
    bool in_table = true;      // false once the host LRU dropped it
    size_t host_bytes = 0;     // host images accounted against the host budget
    std::mutex build_mu;
    bool built = false;        // sim exists (its amplitude engines may not yet)
    std::string err;           // the amplitude engine build failed
    std::unique_ptr<Sim> sim;
    size_t bytes = 0;          // device bytes accounted while resident
    bool resident = false;
    uint64_t last_use = 0;
    int pinned = 0;
};

struct Req {
    uint64_t ticket = 0;
    int kind = 0;                  // 0 amplitude, 1 probability, 2 marginal
    uint64_t x = 0, y = 0;
    std::vector<uint32_t> mq;
    std::vector<uint8_t> mo;
    std::shared_ptr<Entry> entry;
    Circuit circ;
    int strategy = PQ_CACHED;
    double t_submit = 0, t_done = 0;
    // phase stamps: worker start, host build done, engines resident, jobs
    // ready, batch launched (the submit fast path sets the first three to
    // t_submit)
    double t_start = 0, t_built = 0, t_res = 0, t_ready = 0, t_launch = 0;
    std::vector<Z8Job> jobs;       // prepared device work
    std::shared_ptr<const RankEng> keep;   // the engine set the jobs point into (marginals)
    std::complex<double> val{0, 0};
    int status = 0;                // 0 pending, 1 done, -1 failed
    std::string err;
};

bool same_circuit(const Circuit& a, const Circuit& b) {
    if (a.n != b.n || a.gates.size() != b.gates.size()) return false;
    for (size_t i = 0; i < a.gates.size(); ++i) {
        const Gate &x = a.gates[i], &y = b.gates[i];
        if (x.op != y.op || x.a != y.a || x.b != y.b || x.ph != y.ph) return false;
    }
    return true;
}

uint64_t circuit_key(const Circuit& c, int strategy) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { for (int i = 0; i < 8; ++i) { h ^= (v >> (8 * i)) & 0xFF; h *= 1099511628211ull; } };
    mix(c.n); mix((uint64_t)strategy);
    for (auto& g : c.gates) { mix(g.op); mix(g.a); mix(g.b); mix(g.ph); }
    return h;
}

}  // namespace

struct Server {
    uint64_t budget;
    uint64_t host_budget = 8ull << 30;   // host bytes of built engine images kept in the table
    std::atomic<uint64_t> host_resident{0};
    uint64_t host_drops = 0;
    // requests whose jobs are queued for or running on the device: they
    // finish (and unpin) without further admission
    uint64_t inflight = 0;
    int form;
    int walk;
    std::atomic<uint64_t> resident{0};
    std::mutex mu;
    std::condition_variable cv_work, cv_ready, cv_done, cv_mem;
    std::unordered_map<uint64_t, std::shared_ptr<Entry>> entries;
    std::deque<std::shared_ptr<Req>> work_q, ready_q;
    std::unordered_map<uint64_t, std::shared_ptr<Req>> reqs;
    std::vector<std::thread> workers;
    std::thread dispatcher;
    bool stop = false;
    std::atomic<size_t> nready{0};   // requests in ready_q (read by the dispatcher's poll without the lock)
    uint64_t next_ticket = 1, use_clock = 0;
    // statistics
    uint64_t completed = 0, failed = 0, batches = 0, evictions = 0, builds = 0, max_batch = 0, overlapped = 0;
    // phase durations (s) of the last kPhaseLog completed requests: queue,
    // build, upload, prologue, wait for launch, device and finish
    static constexpr size_t kPhaseLog = 1 << 16;
    std::vector<std::array<float, 6>> phase_log;
    size_t phase_next = 0;
    double timed_kernel_ms = 0;      // walk-kernel time of the timed batches
    uint64_t timed_batches = 0, launched = 0;

    Server(uint64_t b, unsigned nworkers, int f) : budget(b), form(f), walk(walk_of_form(f)) {
        for (unsigned i = 0; i < std::max(1u, nworkers); ++i) workers.emplace_back([this] { worker(); });
        dispatcher = std::thread([this] { dispatch(); });
    }
    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all(); cv_ready.notify_all(); cv_mem.notify_all(); cv_done.notify_all();
        for (auto& t : workers) t.join();
        dispatcher.join();
    }

    // ---- admission: make room for `bytes` of engine images in the HBM budget.
    // Called with the requesting entry `self` pinned, so it is never its own
    // victim.  Gives up (false) when the request can never fit (its own
    // resident images plus `bytes` exceed the budget), when every other
    // resident entry is pinned and no queued or running request will release
    // a pin, or after kAdmitWait seconds of back-pressure.
    static constexpr double kAdmitWait = 30.0;
    bool admit(size_t bytes, const Entry* self) {
        std::unique_lock<std::mutex> lk(mu);
        const double t0 = now_s();
        for (;;) {
            if (self->bytes + bytes > budget || stop) return false;
            if (resident.load() + bytes <= budget) { resident += bytes; return true; }
            // evict the least recently used resident entry that nothing pins
            std::shared_ptr<Entry> victim;
            for (auto& kv : entries) {
                auto& e = kv.second;
                if (!e->resident || e->pinned > 0 || e.get() == self) continue;
                if (!victim || e->last_use < victim->last_use) victim = e;
            }
            if (victim) {
                drop_device_locked(*victim);
                evictions++;
                continue;
            }
            if (inflight == 0 || now_s() - t0 > kAdmitWait) return false;
            cv_mem.wait_for(lk, std::chrono::milliseconds(5));   // back-pressure until a batch retires
        }
    }

    void drop_device_locked(Entry& e) {     // with mu held, e unpinned
        if (e.sim) e.sim->drop_device();
        resident -= e.bytes;
        e.bytes = 0;
        e.resident = false;
    }

    // Host LRU: drop the least recently used idle entries from the table while
    // the host images exceed the host budget.  A dropped entry lives on while
    // requests still hold it; its device images go when its last pin does.
    void trim_host_locked(const Entry* keep) {
        while (host_resident.load() > host_budget) {
            std::shared_ptr<Entry> victim;
            for (auto& kv : entries) {
                auto& e = kv.second;
                if (e.get() == keep || e->pinned > 0 || !e->built) continue;
                if (!victim || e->last_use < victim->last_use) victim = e;
            }
            if (!victim) return;
            if (victim->resident) drop_device_locked(*victim);
            host_resident -= victim->host_bytes;
            victim->host_bytes = 0;
            victim->in_table = false;
            entries.erase(victim->key);
            host_drops++;
        }
    }

    // The entry's host images as they are now (engine set and cached marginal
    // sets), accounted against the host budget.
    void account_host(const std::shared_ptr<Entry>& E) {
        size_t hb = 0;
        {
            std::lock_guard<std::mutex> bl(E->build_mu);
            if (E->sim) hb = E->sim->host_bytes();
        }
        std::lock_guard<std::mutex> lk(mu);
        if (!E->in_table) return;
        host_resident += hb;
        host_resident -= E->host_bytes;
        E->host_bytes = hb;
        trim_host_locked(E.get());
    }

    // Host build once per circuit (src/sim.rs:144-148).  The amplitude engine
    // set (src/rank.rs:88-123) is built by the first request that needs it:
    // a marginal request needs only its sandwich engines, so a new circuit
    // whose first request is a marginal waits for those alone.  A failed
    // amplitude build is kept in err and fails the amplitude and probability
    // requests of the circuit; marginals go on.  True when this call built
    // something.
    bool build_entry(const std::shared_ptr<Entry>& E, const Req& r) {
        std::lock_guard<std::mutex> bl(E->build_mu);
        bool did = false;
        if (!E->built) {
            auto s = std::make_unique<Sim>();
            s->circ = r.circ;
            s->strategy = r.strategy;
            E->sim = std::move(s);
            std::lock_guard<std::mutex> lk(mu);
            E->built = true;
            builds++;
            did = true;
        }
        if (r.kind != 2 && !E->sim->eng && E->err.empty()) {
            try {
                Sim& S = *E->sim;
                PathSum ps = compile(S.circ);
                auto eng = build_rank(ps, ps.io_vars(), S.strategy, 0);
                eng->bprog = boundary_prog(ps, *eng->host[0]);
                eng->has_bprog = true;
                std::lock_guard<std::mutex> lk(S.mu);
                S.ps = std::move(ps);
                S.eng = eng;
                did = true;
            } catch (std::exception& e) {
                E->err = e.what();
            }
        }
        return did;
    }

    static size_t estimate_bytes(const RankEng& R) { return R.image_bytes; }

    // Upload the engine set `R` of entry E if it is not on the device,
    // within the budget.  E is pinned by the caller.
    void make_resident(const std::shared_ptr<Entry>& E, const RankEng& R) {
        std::lock_guard<std::mutex> bl(E->build_mu);
        if (R.on_device()) return;
        const size_t est = estimate_bytes(R);
        if (!admit(est, E.get())) throw std::runtime_error("engine does not fit the HBM budget");
        try {
            R.ensure_device();
        } catch (...) {
            resident -= est;
            throw;
        }
        // the entry's device bytes as they are now (marginal sets the
        // simulator's cache dropped no longer count)
        const size_t now_bytes = E->sim->device_bytes();
        std::lock_guard<std::mutex> lk(mu);
        resident -= est;
        resident += now_bytes;
        resident -= E->bytes;
        E->bytes = now_bytes;
        E->resident = true;
    }

    void unpin(Req& r) {   // with mu held
        if (!r.entry) return;
        Entry& e = *r.entry;
        e.pinned--;
        // an entry the host LRU dropped leaves the device with its last pin
        if (e.pinned == 0 && !e.in_table && e.resident) drop_device_locked(e);
        r.entry.reset();
        r.keep.reset();
    }

    // Walk length and size check of a request's device jobs.
    void finish_jobs(Req& r) {
        for (auto& J : r.jobs) {
            const uint32_t d = dev_z8(*J.eng).d;
            if (d > 40) throw std::runtime_error("branch dimension above 40: request rejected");
            J.mid_bits = dev_mid_bits(*J.eng, walk);
        }
    }

    void worker() {
        for (;;) {
            std::shared_ptr<Req> r;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !work_q.empty(); });
                if (stop) return;
                r = work_q.front(); work_q.pop_front();
                r->entry->pinned++;
                r->entry->last_use = ++use_clock;
            }
            r->t_start = now_s();
            try {
                const bool built_now = build_entry(r->entry, *r);
                if (r->kind != 2 && !r->entry->err.empty()) throw std::runtime_error(r->entry->err);
                Sim& S = *r->entry->sim;
                std::vector<std::pair<uint32_t, uint8_t>> asg;
                std::vector<size_t> owner;
                if (r->kind == 2) {
                    r->keep = S.marginal_engine(r->mq);
                    account_host(r->entry);
                    r->t_built = now_s();
                    make_resident(r->entry, *r->keep);
                    r->t_res = now_s();
                    if (S.marginal_assignment(r->x, r->mq.data(), r->mo.data(), r->mq.size(), asg)) add_jobs(*r->keep, asg, walk, r->jobs, owner, 0);
                } else {
                    if (built_now) account_host(r->entry);
                    r->t_built = now_s();
                    make_resident(r->entry, *S.eng);
                    S.eng->ensure_device();
                    r->t_res = now_s();
                    add_query_jobs_xy(*S.eng, r->x, r->y, walk, r->jobs, owner, 0);
                }
                finish_jobs(*r);
                r->t_ready = now_s();
                std::lock_guard<std::mutex> lk(mu);
                ready_q.push_back(r);
                nready++;
                inflight++;
            } catch (std::exception& e) {
                std::lock_guard<std::mutex> lk(mu);
                r->status = -1; r->err = e.what(); r->t_done = now_s();
                r->jobs.clear();
                unpin(*r);
                failed++;
                cv_done.notify_all();
                cv_mem.notify_all();
                continue;
            }
            cv_ready.notify_one();
        }
    }

    void log_phases(const Req& r) {     // with mu held
        const double t_launch = r.t_launch > 0 ? r.t_launch : r.t_ready;
        const std::array<float, 6> ph{(float)(r.t_start - r.t_submit), (float)(r.t_built - r.t_start), (float)(r.t_res - r.t_built),
                                      (float)(r.t_ready - r.t_res), (float)(t_launch - r.t_ready), (float)(r.t_done - t_launch)};
        if (phase_log.size() < kPhaseLog) phase_log.push_back(ph);
        else phase_log[phase_next] = ph;
        phase_next = (phase_next + 1) % kPhaseLog;
    }

    void complete(std::vector<std::shared_ptr<Req>>& batch, const std::vector<std::complex<double>>& res) {
        const double t = now_s();
        std::lock_guard<std::mutex> lk(mu);
        size_t k = 0;
        for (auto& r : batch) {
            std::complex<double> a(0, 0);
            for (size_t j = 0; j < r->jobs.size(); ++j) a += res[k++];
            if (r->kind == 0) r->val = a;
            else if (r->kind == 1) r->val = {std::norm(a), 0};
            else r->val = {std::max(a.real(), 0.0), 0};        // sim.rs:230
            r->status = 1; r->t_done = t;
            log_phases(*r);
            r->jobs.clear();
            unpin(*r);
            inflight--;
            completed++;
        }
        cv_done.notify_all();
        cv_mem.notify_all();
    }

    // Batches in flight at once, each on a slot of its own (stream, events,
    // buffers): requests that arrive while batches run are launched beside
    // them as long as a slot is free.
    static constexpr int kSlots = 4;

    void dispatch() {
        std::shared_ptr<Z8Slot> slots[kSlots];
        std::vector<std::shared_ptr<Req>> held[kSlots];
        std::deque<int> order;       // in-flight slots, oldest first
        try {
            for (auto& sl : slots) sl = dev_slot_create();
        } catch (std::exception&) {
            // no device: requests fail when they reach the dispatcher
            for (auto& sl : slots) sl.reset();
        }
        auto retire = [&](int s) {
            std::vector<std::complex<double>> res;
            Z8Timing T;
            dev_z8_wait(*slots[s], res, &T);
            if (std::isfinite(T.kernel_ms)) {
                std::lock_guard<std::mutex> lk(mu);
                timed_kernel_ms += T.kernel_ms;
                timed_batches++;
            }
            complete(held[s], res);
            held[s].clear();
        };
        for (;;) {
            std::vector<std::shared_ptr<Req>> batch;
            const int inflight = (int)order.size();
            {
                std::unique_lock<std::mutex> lk(mu);
                if (inflight == 0) cv_ready.wait(lk, [&] { return stop || !ready_q.empty(); });
                if (stop && ready_q.empty() && inflight == 0) return;
                if (inflight < kSlots) {
                    while (!ready_q.empty()) { batch.push_back(ready_q.front()); ready_q.pop_front(); }
                    nready -= batch.size();
                }
            }
            if (!batch.empty()) {
                if (!slots[0]) {
                    std::lock_guard<std::mutex> lk(mu);
                    for (auto& r : batch) { r->status = -1; r->err = "no HIP device available"; r->t_done = now_s(); r->jobs.clear(); unpin(*r); this->inflight--; failed++; }
                    cv_done.notify_all();
                    continue;
                }
                int s = 0;                   // a free slot (fewer than kSlots are in flight here)
                while (std::find(order.begin(), order.end(), s) != order.end()) ++s;
                std::vector<Z8Job> jobs;
                for (auto& r : batch) for (auto& J : r->jobs) jobs.push_back(J);
                const double tl = now_s();
                for (auto& r : batch) r->t_launch = tl;
                if (jobs.empty()) { complete(batch, {}); continue; }
                // timing events on every 8th batch only: each event is a
                // queue barrier that lengthens a short batch
                dev_z8_launch(*slots[s], jobs, walk, (launched++ & 7) == 0);
                held[s] = std::move(batch);
                if (!order.empty()) overlapped++;
                order.push_back(s);
                std::lock_guard<std::mutex> lk(mu);
                batches++;
                max_batch = std::max<uint64_t>(max_batch, held[s].size());
            }
            if (order.empty()) continue;
            // retire every finished batch, oldest first
            bool retired = false;
            for (size_t i = 0; i < order.size();) {
                if (dev_z8_ready(*slots[order[i]])) { retire(order[i]); order.erase(order.begin() + (long)i); retired = true; }
                else ++i;
            }
            if (retired) continue;
            if ((int)order.size() == kSlots) {
                // every slot busy: wait for the oldest
                retire(order.front());
                order.pop_front();
                continue;
            }
            // poll the oldest batch without sleeping for up to 2 ms (a batch
            // of a few requests takes tens of microseconds), going back to
            // launch as soon as a request is ready; after that, wait in
            // 200 us steps
            const int s = order.front();
            bool done = false;
            const double t0 = now_s();
            while (!(done = dev_z8_ready(*slots[s]))) {
                if (nready.load(std::memory_order_acquire) > 0 || now_s() - t0 > 2e-3) break;
                std::this_thread::yield();
            }
            if (done) {
                retire(s);
                order.pop_front();
                continue;
            }
            std::unique_lock<std::mutex> lk(mu);
            if (ready_q.empty() && !stop) cv_ready.wait_for(lk, std::chrono::microseconds(200));
        }
    }

    uint64_t submit(const Circuit& c, int strategy, int kind, uint64_t x, uint64_t y, const uint32_t* mq, const uint8_t* mo, size_t k) {
        auto r = std::make_shared<Req>();
        r->kind = kind; r->x = x; r->y = y; r->circ = c; r->strategy = strategy;
        r->mq.assign(mq, mq + k); r->mo.assign(mo, mo + k);
        r->t_submit = now_s();
        const uint64_t key = circuit_key(c, strategy);
        // An amplitude or probability request on a circuit whose engines are
        // built and on the device takes its prologue (about 0.2 us) here, on
        // the submitting thread, and goes straight to the dispatcher; the
        // rest go through a worker (host build, upload, marginal engines).
        bool here = false;
        {
            std::lock_guard<std::mutex> lk(mu);
            // open addressing on the 64-bit key: a colliding circuit with
            // other content takes the next free key
            uint64_t k = key;
            auto it = entries.find(k);
            while (it != entries.end() && !(it->second->strategy == strategy && same_circuit(it->second->circ, c))) it = entries.find(++k);
            if (it == entries.end()) {
                auto e = std::make_shared<Entry>();
                e->key = k;
                e->circ = c;
                e->strategy = strategy;
                it = entries.emplace(k, e).first;
            }
            r->entry = it->second;
            r->ticket = next_ticket++;
            reqs[r->ticket] = r;
            Entry& E = *r->entry;
            // try_lock: the build lock is taken before mu elsewhere.  The
            // build state and the device images are read under it; a pinned
            // entry is never evicted, so the images stay after it is released.
            if (kind != 2 && E.build_mu.try_lock()) {
                if (E.built && E.err.empty()) {
                    here = E.sim->eng && E.sim->eng->on_device();
                }
                E.build_mu.unlock();
                if (here) { E.pinned++; E.last_use = ++use_clock; }
            }
            if (!here) work_q.push_back(r);
        }
        if (!here) {
            cv_work.notify_one();
            return r->ticket;
        }
        r->t_start = r->t_built = r->t_res = r->t_submit;
        try {
            std::vector<size_t> owner;
            add_query_jobs_xy(*r->entry->sim->eng, r->x, r->y, walk, r->jobs, owner, 0);
            finish_jobs(*r);
            r->t_ready = now_s();
        } catch (std::exception& e) {
            std::lock_guard<std::mutex> lk(mu);
            r->status = -1; r->err = e.what(); r->t_done = now_s();
            r->jobs.clear();
            unpin(*r);
            failed++;
            cv_done.notify_all();
            cv_mem.notify_all();
            return r->ticket;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            ready_q.push_back(r);
            nready++;
            inflight++;
        }
        cv_ready.notify_one();
        return r->ticket;
    }

    int wait(uint64_t ticket, double timeout_s, double* out, double* latency_ms) {
        std::unique_lock<std::mutex> lk(mu);
        auto it = reqs.find(ticket);
        if (it == reqs.end()) throw std::invalid_argument("unknown ticket");
        auto r = it->second;
        auto pred = [&] { return r->status != 0; };
        if (timeout_s < 0) cv_done.wait(lk, pred);
        else if (!cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return 1;
        reqs.erase(it);
        if (latency_ms) *latency_ms = 1e3 * (r->t_done - r->t_submit);
        if (r->status < 0) throw std::runtime_error(r->err);
        out[0] = r->val.real(); out[1] = r->val.imag();
        return 0;
    }
};

}  // namespace pq

using namespace pq;
struct pq_server { std::unique_ptr<Server> s; };

extern "C" {

int pq_server_create(uint64_t hbm_budget_bytes, uint32_t workers, int form, pq_server** out) {
    try {
        (void)walk_of_form(form);   // the server runs the batched forms only
        auto h = std::make_unique<pq_server>();
        h->s = std::make_unique<Server>(hbm_budget_bytes, workers, form);
        *out = h.release();
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

void pq_server_destroy(pq_server* h) { delete h; }

int pq_server_submit(pq_server* h, uint32_t nq, size_t ng, const uint8_t* ops, const uint32_t* q0, const uint32_t* q1, const uint32_t* ph,
                     int strategy, int kind, uint64_t x, uint64_t y, const uint32_t* mq, const uint8_t* mo, size_t k, uint64_t* ticket) {
    try {
        if (kind < 0 || kind > 2) throw std::invalid_argument("kind must be 0 (amplitude), 1 (probability) or 2 (marginal)");
        if (nq > 64) throw std::invalid_argument("at most 64 qubits");
        Circuit c; c.n = nq;
        for (size_t i = 0; i < ng; ++i) c.add(ops[i], q0[i], q1 ? q1[i] : 0, ph ? ph[i] : 0);
        if ((nq < 64 && ((x >> nq) || (y >> nq)))) throw std::invalid_argument("basis index has bits beyond the qubit count");
        for (size_t i = 0; i < k; ++i) if (mq[i] >= c.n) throw std::invalid_argument("measured qubit out of range");
        *ticket = h->s->submit(c, strategy, kind, x, y, mq, mo, k);
        return 0;
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

int pq_server_wait(pq_server* h, uint64_t ticket, double timeout_s, double* out, double* latency_ms) {
    try {
        return h->s->wait(ticket, timeout_s, out, latency_ms);
    } catch (std::exception& e) { g_err = e.what(); return -1; }
}

int pq_server_stats(pq_server* h, double* o) {
    Server& S = *h->s;
    std::lock_guard<std::mutex> lk(S.mu);
    o[0] = (double)S.completed; o[1] = (double)S.failed; o[2] = (double)S.batches; o[3] = (double)S.evictions;
    o[4] = (double)S.resident.load(); o[5] = (double)S.builds; o[6] = (double)S.max_batch; o[7] = (double)S.overlapped;
    o[8] = S.timed_batches ? S.timed_kernel_ms * (double)S.batches / (double)S.timed_batches : 0.0;
    o[9] = (double)S.entries.size();
    return 0;
}

int pq_server_phase_stats(pq_server* h, double* o) {
    Server& S = *h->s;
    std::vector<std::array<float, 6>> log;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        log = S.phase_log;
    }
    for (int k = 0; k < 13; ++k) o[k] = 0.0;
    o[0] = (double)log.size();
    if (log.empty()) return 0;
    std::vector<std::pair<double, size_t>> tot(log.size());
    for (size_t i = 0; i < log.size(); ++i) {
        double t = 0;
        for (int k = 0; k < 6; ++k) t += log[i][k];
        tot[i] = {t, i};
    }
    std::sort(tot.begin(), tot.end());
    const size_t ntail = std::max<size_t>(1, log.size() / 100);
    for (size_t i = 0; i < log.size(); ++i)
        for (int k = 0; k < 6; ++k) o[1 + k] += 1e3 * log[i][k] / (double)log.size();
    for (size_t j = log.size() - ntail; j < log.size(); ++j)
        for (int k = 0; k < 6; ++k) o[7 + k] += 1e3 * log[tot[j].second][k] / (double)ntail;
    return 0;
}

int pq_server_set_host_budget(pq_server* h, uint64_t bytes) {
    Server& S = *h->s;
    std::lock_guard<std::mutex> lk(S.mu);
    S.host_budget = bytes;
    S.trim_host_locked(nullptr);
    return 0;
}

int pq_server_host_stats(pq_server* h, double* o) {
    Server& S = *h->s;
    std::lock_guard<std::mutex> lk(S.mu);
    o[0] = (double)S.host_resident.load(); o[1] = (double)S.host_budget; o[2] = (double)S.host_drops;
    return 0;
}

}  // extern "C"
