// Simulator / RankEngine mirror (src/sim.rs, src/rank.rs) with the GPU branch
// evaluator as the strategy backend, and the C ABI declared in
// include/polyq_amd.h.
#include "pq_host.hpp"
#include "pq_device.hpp"
#include "pq_api_internal.hpp"
#include "pq_pool.hpp"
#include "../../include/polyq_amd.h"
#include <mutex>
#include <map>
#include <cstring>
#include <algorithm>
#include <exception>

namespace pq {

thread_local std::string g_err;
thread_local double g_timing[5] = {0, 0, 0, 0, 0};

std::shared_ptr<RankEng> build_rank(const PathSum& ps, const std::vector<uint32_t>& fixed, int strategy, uint32_t pairs) {
    auto R = std::make_shared<RankEng>();
    auto push = [&](std::shared_ptr<HostEngine> he, std::shared_ptr<Z8Image> z, std::shared_ptr<SegImage> g) {
        R->image_bytes += dev_engine_image_bytes(*he, *z, *g);
        R->host.push_back(he);
        R->z8.push_back(z);
        R->seg.push_back(g);
    };
    auto add = [&](std::shared_ptr<HostEngine> he) {
        push(he, std::make_shared<Z8Image>(build_z8_image(*he)), std::make_shared<SegImage>(build_seg_image(*he)));
    };
    uint32_t nt = 0;
    auto cands = pair_candidates(ps, fixed, nt);
    uint32_t p = 0;
    if (strategy == PQ_AUTO || strategy == PQ_GPU_AUTO) {
        // rank.rs:115-123: base engine, cost model, rebuild with p pairs if p > 0.
        // A base engine beyond the device's term limit still has its stats,
        // and the pair engines may be within the limit.
        std::shared_ptr<HostEngine> base;
        EngineStats st;
        try {
            base = build_engine(ps, fixed);
            st = base->stats();
        } catch (const TermLimit& e) {
            st = e.st;
        }
        p = strategy == PQ_AUTO ? best_pairs(st, (uint32_t)cands.size(), nt)
                                : best_pairs_gpu(st, (uint32_t)cands.size(), nt, kGpuBranchSeconds);
        if (p == 0) {
            if (!base) throw TermLimit(st);
            add(base);
            R->pairs = 0;
        }
    } else if (strategy == PQ_CACHED) {
        p = 0;
        add(build_engine(ps, fixed));
    } else if (strategy == PQ_PAIRS) {
        p = std::min<uint32_t>(std::min<uint32_t>(pairs, (uint32_t)cands.size()), kMaxPairs);
        if (p == 0) add(build_engine(ps, fixed));
    } else {
        throw std::invalid_argument("strategy must be PQ_AUTO, PQ_CACHED, PQ_PAIRS or PQ_GPU_AUTO");
    }
    if (R->host.empty()) {
        // rank.rs:88-103: one engine per substitution pattern of the chosen
        // pairs.  The patterns' engines and walk images are independent: they
        // are built on the host pool and kept in pattern order.
        const size_t np = size_t(1) << p;
        std::vector<std::shared_ptr<HostEngine>> hs(np);
        std::vector<std::shared_ptr<Z8Image>> zs(np);
        std::vector<std::shared_ptr<SegImage>> gs(np);
        std::mutex emu;
        std::exception_ptr err;
        ChunkPool::get().run(np, 1, 64, [&](size_t lo, size_t hi, unsigned) {
            try {
                for (size_t pattern = lo; pattern < hi; ++pattern) {
                    PathSum b = ps;
                    for (uint32_t i = 0; i < p; ++i) b.substitute(cands[i].first, cands[i].second, (pattern >> i) & 1);
                    hs[pattern] = build_engine(b, fixed);
                    zs[pattern] = std::make_shared<Z8Image>(build_z8_image(*hs[pattern]));
                    gs[pattern] = std::make_shared<SegImage>(build_seg_image(*hs[pattern]));
                }
            } catch (...) {
                std::lock_guard<std::mutex> lk(emu);
                if (!err) err = std::current_exception();
            }
        });
        if (err) std::rethrow_exception(err);
        for (size_t pattern = 0; pattern < np; ++pattern) push(hs[pattern], zs[pattern], gs[pattern]);
        R->pairs = p;
    }
    return R;
}

// Append the device jobs of one (engine set, assignment) query: one job per
// branch engine whose prologue does not vanish.  Only the query record of the
// given walk form is built.  The engine set must already be on the device.
void add_query_jobs(const RankEng& R, const std::vector<std::pair<uint32_t, uint8_t>>& asg, int walk, std::vector<Z8Job>& jobs,
                    std::vector<size_t>& owner, size_t qid) {
    thread_local Prepared P;
    for (size_t b = 0; b < R.host.size(); ++b) {
        if (!prepare_assignment(*R.host[b], asg, P)) continue;
        Z8Job J;
        J.eng = R.dev[b].get();
        if (walk == kWalkSeg) J.sq = seg_query(*R.host[b], *R.seg[b], P);
        else J.q = z8_query(*R.host[b], *R.z8[b], P);
        J.scale = P.scale;
        J.mid_bits = 0;
        jobs.push_back(J);
        owner.push_back(qid);
    }
}

// Same for <y|C|x> on amplitude engines, straight from the basis indices.
void add_query_jobs_xy(const RankEng& R, uint64_t x, uint64_t y, int walk, std::vector<Z8Job>& jobs, std::vector<size_t>& owner, size_t qid) {
    thread_local Prepared P;
    for (size_t b = 0; b < R.host.size(); ++b) {
        bool contradictory = false;
        if (!prepare_bits(*R.host[b], R.bprog, x, y, P, &contradictory)) {
            if (contradictory) return;          // exactly 0 on every branch engine
            continue;
        }
        Z8Job J;
        J.eng = R.dev[b].get();
        if (walk == kWalkSeg) J.sq = seg_query(*R.host[b], *R.seg[b], P);
        else J.q = z8_query(*R.host[b], *R.z8[b], P);
        J.scale = P.scale;
        J.mid_bits = 0;
        jobs.push_back(J);
        owner.push_back(qid);
    }
}

void add_jobs(const RankEng& R, const std::vector<std::pair<uint32_t, uint8_t>>& asg, int walk, std::vector<Z8Job>& jobs,
              std::vector<size_t>& owner, size_t qid) {
    R.ensure_device();
    add_query_jobs(R, asg, walk, jobs, owner, qid);
}

// Jobs of n queries <ys[i]|C_i|xs[i]>, query i on the amplitude engine set
// rank_of(i).  Large batches run their prologues on the host thread pool; the
// jobs come out in query order either way.
void build_jobs(size_t n, int walk, const std::function<const RankEng*(size_t)>& rank_of, const std::function<void(size_t)>& check,
                const uint64_t* xs, const uint64_t* ys, std::vector<Z8Job>& jobs, std::vector<size_t>& owner) {
    const RankEng* last = nullptr;
    for (size_t i = 0; i < n; ++i) {
        check(i);                       // every index first: a bad one fails the call before any work
        const RankEng* R = rank_of(i);
        if (R != last) {
            if (!R->has_bprog) throw std::logic_error("engine set without a boundary map");
            R->ensure_device();
            last = R;
        }
    }
    // A prologue takes about 0.15 us and handing chunks to the pool costs a
    // few microseconds, so batches below kInline queries stay on the calling
    // thread.  Larger ones go out in chunks of at least 8 queries; below 256
    // queries only to the caller and the pool's polling workers, so no
    // blocked worker has to be woken.
    constexpr size_t kChunk = 8, kInline = 48;
    jobs.clear(); owner.clear();       // the jobs of these n queries only, whatever the vectors held
    if (n < kInline) {
        for (size_t i = 0; i < n; ++i) add_query_jobs_xy(*rank_of(i), xs[i], ys[i], walk, jobs, owner, i);
        return;
    }
    ChunkPool& pool = ChunkPool::get();
    // large batches in four chunks per thread: workers that wake late (a
    // blocked worker takes tens of microseconds) still find work
    const unsigned maxch = n < 256 ? ChunkPool::kHot + 1 : 4 * pool.threads();
    std::vector<std::vector<Z8Job>> cj(maxch);
    std::vector<std::vector<size_t>> co(maxch);
    std::vector<std::string> err(maxch);
    const unsigned nch = pool.run(n, kChunk, maxch, [&](size_t lo, size_t hi, unsigned c) {
        try {
            for (size_t i = lo; i < hi; ++i) add_query_jobs_xy(*rank_of(i), xs[i], ys[i], walk, cj[c], co[c], i);
        } catch (std::exception& e) {
            err[c] = e.what();
        }
    });
    for (unsigned c = 0; c < nch; ++c) if (!err[c].empty()) throw std::invalid_argument(err[c]);
    if (nch == 1) { jobs.swap(cj[0]); owner.swap(co[0]); return; }
    size_t tot = 0;
    for (unsigned c = 0; c < nch; ++c) tot += cj[c].size();
    jobs.reserve(tot); owner.reserve(tot);
    for (unsigned c = 0; c < nch; ++c) {
        jobs.insert(jobs.end(), cj[c].begin(), cj[c].end());
        owner.insert(owner.end(), co[c].begin(), co[c].end());
    }
}

int walk_of_form(int form) {
    if (form == PQ_FORM_Z8) return kWalkSeg;
    if (form == PQ_FORM_BRANCH_WALK) return kWalkBranch;
    throw std::invalid_argument("batched evaluation uses PQ_FORM_Z8 or PQ_FORM_BRANCH_WALK");
}

void assign_mid_bits(std::vector<Z8Job>& jobs, int walk) {
    for (auto& J : jobs) J.mid_bits = dev_mid_bits(*J.eng, walk);
}

void run_jobs(std::vector<Z8Job>& jobs, const std::vector<size_t>& owner, size_t nq, int walk, std::vector<std::complex<double>>& out) {
    out.assign(nq, {0, 0});
    if (jobs.empty()) { for (double& x : g_timing) x = 0; return; }
    assign_mid_bits(jobs, walk);
    std::vector<std::complex<double>> r;
    Z8Timing T;
    dev_z8_batch(jobs, r, &T, walk);
    for (size_t i = 0; i < jobs.size(); ++i) out[owner[i]] += r[i];
    g_timing[0] = T.kernel_ms; g_timing[1] = T.total_ms; g_timing[2] = (double)T.branches; g_timing[3] = (double)T.items; g_timing[4] = T.grid;
}

std::complex<double> plain_eval(const RankEng& R, const std::vector<std::pair<uint32_t, uint8_t>>& asg) {
    std::complex<double> tot(0, 0);
    double kms = 0, kall = 0;
    uint64_t br = 0;
    R.ensure_device();
    for (size_t b = 0; b < R.host.size(); ++b) {
        Prepared P;
        if (!prepare_assignment(*R.host[b], asg, P)) continue;
        std::complex<double> s = dev_plain_sum(*R.dev[b], P, &kms);
        const HostEngine& E = *R.host[b];
        // engine.rs:560: phase * total * 2^{log2mag - h/2}
        tot += phase_complex(P.phase) * s * std::pow(2.0, E.log2mag - (double)E.nh / 2.0);
        kall += kms;
        br += 1ull << E.d;
    }
    g_timing[0] = kall; g_timing[1] = kall; g_timing[2] = (double)br; g_timing[3] = 0; g_timing[4] = 0;
    return tot;
}

bool Sim::assignment(uint64_t x, uint64_t y, std::vector<std::pair<uint32_t, uint8_t>>& asg) const {
    if (ps.nq > 64) throw std::invalid_argument(kWideCircuit);
    return ps.boundary_bits(x, y, asg);
}

std::shared_ptr<const RankEng> Sim::marginal_engine(const std::vector<uint32_t>& qubits) {
    std::vector<uint32_t> key(qubits);
    std::sort(key.begin(), key.end());
    key.erase(std::unique(key.begin(), key.end()), key.end());
    std::vector<uint32_t> f;
    const Sandwich* sp;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!sw) sw = std::make_unique<Sandwich>(compile_sandwich(circ));
        auto it = marg.find(key);
        if (it != marg.end()) {
            it->second.last_use = ++marg_clock;
            return it->second.r;
        }
        // Sandwich::fixed_for (sim.rs:95-103)
        f = sw->ps.io_vars();
        for (uint32_t q : qubits) if (std::find(f.begin(), f.end(), sw->junction[q]) == f.end()) f.push_back(sw->junction[q]);
        sp = sw.get();          // made once, never replaced
    }
    // built outside the lock: marginals of other qubit sets (and amplitude
    // requests) on this simulator go on meanwhile; a concurrent build of the
    // same set keeps the first one cached
    auto R = build_rank(sp->ps, f, strategy, pairs_opt);
    std::lock_guard<std::mutex> lk(mu);
    auto it = marg.find(key);
    if (it != marg.end()) {
        it->second.last_use = ++marg_clock;
        return it->second.r;
    }
    if (marg.size() >= kMargCap) {
        auto lru = marg.begin();
        for (auto i = marg.begin(); i != marg.end(); ++i) if (i->second.last_use < lru->second.last_use) lru = i;
        marg.erase(lru);
    }
    marg[key] = MargSet{R, ++marg_clock};
    return R;
}

size_t Sim::host_bytes() {
    std::lock_guard<std::mutex> lk(mu);
    size_t b = eng ? eng->image_bytes : 0;
    for (auto& kv : marg) b += kv.second.r->image_bytes;
    return b;
}

bool Sim::marginal_assignment(uint64_t x, const uint32_t* qubits, const uint8_t* outcomes, size_t k, std::vector<std::pair<uint32_t, uint8_t>>& a) {
    if (ps.nq > 64) throw std::invalid_argument(kWideCircuit);
    // Sandwich::assignment (sim.rs:106-124): x on both ends, the measured
    // outcomes on the junction variables
    std::vector<uint32_t> jv(k);
    std::vector<uint8_t> jval(k);
    for (size_t i = 0; i < k; ++i) { jv[i] = sw->junction[qubits[i]]; jval[i] = (uint8_t)((outcomes[i] ? 1 : 0) ^ sw->jneg[qubits[i]]); }
    return sw->ps.boundary_bits(x, x, a, k, jv.data(), jval.data());
}

bool Sim::marginal_assignment_bytes(const uint8_t* x, const uint32_t* qubits, const uint8_t* outcomes, size_t k,
                                    std::vector<std::pair<uint32_t, uint8_t>>& a) {
    std::vector<uint32_t> jv(k);
    std::vector<uint8_t> jval(k);
    for (size_t i = 0; i < k; ++i) { jv[i] = sw->junction[qubits[i]]; jval[i] = (uint8_t)((outcomes[i] ? 1 : 0) ^ sw->jneg[qubits[i]]); }
    return sw->ps.boundary(x, x, a, k, jv.data(), jval.data());
}

}  // namespace pq

using namespace pq;

struct pq_sim { Sim s; };
struct pq_graph { std::shared_ptr<Z8Graph> g; std::vector<size_t> owner; size_t nq = 0; std::vector<std::shared_ptr<RankEng>> keep; };

#define PQ_TRY(...)                                                                                   \
    try { __VA_ARGS__ } catch (std::exception & e) { g_err = e.what(); return -1; } catch (...) { g_err = "unknown error"; return -1; }

// The per-branch diagnostics take t-bit branch vectors as u64 words: engines
// with more than 64 terms are evaluated, not dumped.
static void diag_terms_ok(const HostEngine& E) {
    if (E.tdev != E.t) throw std::invalid_argument("per-branch diagnostics need at most 64 non-Clifford terms");
}

// A u64 basis index addresses at most 64 qubits: wider circuits take the
// byte-array entry points.
static bool qubits_ok(const pq_sim* h, uint64_t x) {
    if (h->s.ps.nq > 64) throw std::invalid_argument(kWideCircuit);
    return h->s.ps.nq == 64 || (x >> h->s.ps.nq) == 0;
}

// Sum over the branch engines of R for one (variable, value) assignment
// (RankEngine::eval_assignment, src/rank.rs:184-190), in the given form.
static std::complex<double> eval_assignment_form(const RankEng& R, const std::vector<std::pair<uint32_t, uint8_t>>& asg, int form) {
    if (form == PQ_FORM_PLAIN_FP64) return plain_eval(R, asg);
    std::vector<Z8Job> jobs;
    std::vector<size_t> owner;
    const int walk = walk_of_form(form);
    add_jobs(R, asg, walk, jobs, owner, 0);
    std::vector<std::complex<double>> res;
    run_jobs(jobs, owner, 1, walk, res);
    return res[0];
}

extern "C" {

const char* pq_last_error(void) { return g_err.c_str(); }

int pq_device_info(char* buf, size_t len) {
    PQ_TRY({
        if (dev_count() == 0) { g_err = "no HIP device"; return -1; }
        std::string s = dev_name();
        if (buf && len) { std::strncpy(buf, s.c_str(), len - 1); buf[len - 1] = 0; }
        return 0;
    })
}

int pq_sim_create(uint32_t nq, size_t ng, const uint8_t* ops, const uint32_t* q0, const uint32_t* q1, const uint32_t* ph, int strategy,
                  uint32_t pairs, pq_sim** out) {
    PQ_TRY({
        if (!out) throw std::invalid_argument("null output handle");
        auto h = std::make_unique<pq_sim>();
        h->s.circ.n = nq;
        for (size_t i = 0; i < ng; ++i) h->s.circ.add(ops[i], q0[i], q1 ? q1[i] : 0, ph ? ph[i] : 0);
        h->s.strategy = strategy; h->s.pairs_opt = pairs;
        h->s.ps = compile(h->s.circ);
        h->s.eng = build_rank(h->s.ps, h->s.ps.io_vars(), strategy, pairs);
        // the basis-index boundary map serves circuits of at most 64 qubits;
        // wider ones go through the (variable, value) assignment
        if (h->s.ps.nq <= 64) {
            h->s.eng->bprog = boundary_prog(h->s.ps, *h->s.eng->host[0]);
            h->s.eng->has_bprog = true;
        }
        *out = h.release();
        return 0;
    })
}

void pq_sim_destroy(pq_sim* h) { delete h; }

uint32_t pq_sim_num_qubits(const pq_sim* h) { return h ? h->s.ps.nq : 0; }

int pq_sim_stats(const pq_sim* h, uint64_t* o) {
    PQ_TRY({
        EngineStats s = h->s.eng->host[0]->stats();
        o[0] = s.free_vars; o[1] = s.planes; o[2] = s.lines; o[3] = s.radical; o[4] = s.terms; o[5] = s.branch_dim; o[6] = h->s.eng->pairs;
        return 0;
    })
}

int pq_sim_amplitudes(pq_sim* h, size_t n, const uint64_t* xs, const uint64_t* ys, int form, double* out) {
    PQ_TRY({
        std::vector<std::complex<double>> res(n, {0, 0});
        std::vector<std::pair<uint32_t, uint8_t>> asg;
        if (form == PQ_FORM_PLAIN_FP64) {
            for (size_t i = 0; i < n; ++i) {
                if (!qubits_ok(h, xs[i]) || !qubits_ok(h, ys[i])) throw std::invalid_argument("basis index has bits beyond the qubit count");
                if (!h->s.assignment(xs[i], ys[i], asg)) continue;
                res[i] = plain_eval(*h->s.eng, asg);
            }
        } else {
            const int walk = walk_of_form(form);
            std::vector<Z8Job> jobs;
            std::vector<size_t> owner;
            const RankEng* R = h->s.eng.get();
            build_jobs(n, walk, [&](size_t) { return R; }, [&](size_t i) {
                if (!qubits_ok(h, xs[i]) || !qubits_ok(h, ys[i])) throw std::invalid_argument("basis index has bits beyond the qubit count");
            }, xs, ys, jobs, owner);
            run_jobs(jobs, owner, n, walk, res);
        }
        for (size_t i = 0; i < n; ++i) { out[2 * i] = res[i].real(); out[2 * i + 1] = res[i].imag(); }
        return 0;
    })
}

int pq_sim_amplitudes_noalloc(pq_sim* h, size_t n, const uint64_t* xs, const uint64_t* ys, int form, double* out) {
    PQ_TRY({
        const int walk = walk_of_form(form);     // the batched forms only
        const RankEng* R = h->s.eng.get();
        if (!R->on_device())
            throw std::invalid_argument("engines are not resident on the device: evaluate once through an allocating entry point first");
        thread_local std::vector<Z8Job> jobs;
        thread_local std::vector<size_t> owner;
        thread_local std::vector<std::complex<double>> r, res;
        build_jobs(n, walk, [&](size_t) { return R; }, [&](size_t i) {
            if (!qubits_ok(h, xs[i]) || !qubits_ok(h, ys[i])) throw std::invalid_argument("basis index has bits beyond the qubit count");
        }, xs, ys, jobs, owner);
        res.assign(n, {0, 0});
        if (!jobs.empty()) {
            assign_mid_bits(jobs, walk);
            dev_z8_batch_noalloc(jobs, r, walk);
            for (size_t i = 0; i < jobs.size(); ++i) res[owner[i]] += r[i];
        }
        for (size_t i = 0; i < n; ++i) { out[2 * i] = res[i].real(); out[2 * i + 1] = res[i].imag(); }
        return 0;
    })
}

size_t pq_noalloc_scratch_bytes(void) { return dev_noalloc_scratch_bytes(); }

int pq_best_pairs(const uint64_t* st6, uint32_t avail, uint32_t tclass, int gpu_model) {
    PQ_TRY({
        const EngineStats st{st6[0], st6[1], st6[2], st6[3], st6[4], st6[5]};
        return (int)(gpu_model ? best_pairs_gpu(st, avail, tclass, kGpuBranchSeconds) : best_pairs(st, avail, tclass));
    })
}

int pq_sim_amplitude_index(pq_sim* h, uint64_t x, uint64_t y, int form, double* out) { return pq_sim_amplitudes(h, 1, &x, &y, form, out); }

int pq_sim_amplitude(pq_sim* h, const uint8_t* input, const uint8_t* output, int form, double* out) {
    if (h->s.ps.nq <= 64) {
        uint64_t x = 0, y = 0;
        for (uint32_t q = 0; q < h->s.ps.nq; ++q) { x |= (uint64_t)(input[q] ? 1 : 0) << q; y |= (uint64_t)(output[q] ? 1 : 0) << q; }
        return pq_sim_amplitudes(h, 1, &x, &y, form, out);
    }
    PQ_TRY({
        // sim.rs:171-176 for any number of qubits
        std::vector<std::pair<uint32_t, uint8_t>> asg;
        std::complex<double> v(0, 0);
        if (h->s.ps.boundary(input, output, asg)) v = eval_assignment_form(*h->s.eng, asg, form);
        out[0] = v.real(); out[1] = v.imag();
        return 0;
    })
}

int pq_sim_probability_bytes(pq_sim* h, const uint8_t* input, const uint8_t* output, int form, double* out) {
    double a[2];
    const int r = pq_sim_amplitude(h, input, output, form, a);
    if (r == 0) *out = a[0] * a[0] + a[1] * a[1];
    return r;
}

int pq_sim_probability(pq_sim* h, uint64_t x, uint64_t y, int form, double* out) {
    double a[2];
    int r = pq_sim_amplitudes(h, 1, &x, &y, form, a);
    if (r == 0) *out = a[0] * a[0] + a[1] * a[1];
    return r;
}

int pq_sim_marginal_probability(pq_sim* h, uint64_t x, const uint32_t* qubits, const uint8_t* outcomes, size_t k, int form, double* out) {
    PQ_TRY({
        if (!qubits_ok(h, x)) throw std::invalid_argument("basis index has bits beyond the qubit count");
        for (size_t i = 0; i < k; ++i) if (qubits[i] >= h->s.ps.nq) throw std::invalid_argument("measured qubit out of range");
        std::vector<uint32_t> qs(qubits, qubits + k);
        const auto Rp = h->s.marginal_engine(qs);
        const RankEng& R = *Rp;
        std::vector<std::pair<uint32_t, uint8_t>> a;
        if (!h->s.marginal_assignment(x, qubits, outcomes, k, a)) { *out = 0; return 0; }
        *out = std::max(eval_assignment_form(R, a, form).real(), 0.0);   // sim.rs:230
        return 0;
    })
}

int pq_sim_marginal_probability_bytes(pq_sim* h, const uint8_t* input, const uint32_t* qubits, const uint8_t* outcomes, size_t k, int form,
                                      double* out) {
    PQ_TRY({
        for (size_t i = 0; i < k; ++i) if (qubits[i] >= h->s.ps.nq) throw std::invalid_argument("measured qubit out of range");
        std::vector<uint32_t> qs(qubits, qubits + k);
        const auto Rp = h->s.marginal_engine(qs);
        std::vector<std::pair<uint32_t, uint8_t>> a;
        if (!h->s.marginal_assignment_bytes(input, qubits, outcomes, k, a)) { *out = 0; return 0; }
        *out = std::max(eval_assignment_form(*Rp, a, form).real(), 0.0);   // sim.rs:230
        return 0;
    })
}

namespace {
// SplitMix64 as the reference's Rng (rng.rs:9-27): state = seed + gamma,
// next adds gamma to the state and mixes the new state; next_f64 takes the
// top 53 bits.
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed + 0x9E3779B97F4A7C15ull) {}
    uint64_t next_u64() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double next_f64() { return (double)(next_u64() >> 11) / (double)(1ull << 53); }
};
}  // namespace

// Simulator::amplitude_estimate (sim.rs:187-193) -> RankEngine::estimate
// (rank.rs:149-171) -> Engine::estimate (engine.rs:612-659).  out = {re, im,
// standard error}.  One branch engine: the exact sum when 2^d <= samples,
// else the Monte-Carlo mean over `samples` uniform points of the branch
// subspace (k_estimate).  Several: exact values of uniformly drawn branch
// engines; each drawn engine is evaluated once (in one batch) and counted as
// often as it was drawn.
int pq_sim_amplitude_estimate(pq_sim* h, uint64_t x, uint64_t y, size_t samples, uint64_t seed, int form, double* out) {
    PQ_TRY({
        out[0] = out[1] = out[2] = 0;
        if (!qubits_ok(h, x) || !qubits_ok(h, y)) throw std::invalid_argument("basis index has bits beyond the qubit count");
        if (samples == 0) throw std::invalid_argument("estimate with zero samples");
        std::vector<std::pair<uint32_t, uint8_t>> asg;
        if (!h->s.assignment(x, y, asg)) return 0;                 // sim.rs:190
        const RankEng& R = *h->s.eng;
        R.ensure_device();
        const size_t nb = R.host.size();
        auto exact_values = [&](const std::vector<size_t>& which, std::vector<std::complex<double>>& val) {
            val.assign(nb, {0, 0});
            if (form == PQ_FORM_PLAIN_FP64) {
                for (size_t b : which) {
                    Prepared P;
                    if (!prepare_assignment(*R.host[b], asg, P)) continue;
                    const HostEngine& E = *R.host[b];
                    val[b] = phase_complex(P.phase) * dev_plain_sum(*R.dev[b], P, nullptr) * std::pow(2.0, E.log2mag - (double)E.nh / 2.0);
                }
                return;
            }
            const int walk = walk_of_form(form);
            std::vector<Z8Job> jobs;
            std::vector<size_t> owner;
            Prepared P;
            for (size_t b : which) {
                if (!prepare_assignment(*R.host[b], asg, P)) continue;
                Z8Job J;
                J.eng = R.dev[b].get();
                if (walk == kWalkSeg) J.sq = seg_query(*R.host[b], *R.seg[b], P);
                else J.q = z8_query(*R.host[b], *R.z8[b], P);
                J.scale = P.scale;
                J.mid_bits = 0;
                jobs.push_back(J);
                owner.push_back(b);
            }
            run_jobs(jobs, owner, nb, walk, val);
        };
        if (nb == 1) {
            const HostEngine& E = *R.host[0];
            Prepared P;
            if (!prepare_assignment(E, asg, P)) return 0;           // engine.rs:618
            const uint32_t d = E.d;
            if (d < 63 && (1ull << d) <= samples) {                  // engine.rs:622-624
                std::vector<std::complex<double>> val;
                exact_values({0}, val);
                out[0] = val[0].real(); out[1] = val[0].imag();
                return 0;
            }
            const McSums S = dev_plain_estimate(*R.dev[0], P, samples, seed);
            const std::complex<double> norm = phase_complex(P.phase) * std::pow(2.0, E.log2mag - (double)E.nh / 2.0);
            const double n = (double)S.n;
            const std::complex<double> mean(S.re / n, S.im / n);
            const double var = std::max(S.sq / n - std::norm(mean), 0.0);
            const double scale = std::pow(2.0, (double)d);
            const std::complex<double> est = norm * mean * scale;
            out[0] = est.real(); out[1] = est.imag();
            out[2] = std::abs(norm) * scale * std::sqrt(var / n);
            return 0;
        }
        // rank.rs:158-170: sample i draws its branch from Rng(seed ^ i * gamma)
        std::vector<uint64_t> count(nb, 0);
        for (size_t i = 0; i < samples; ++i) {
            SplitMix r(seed ^ ((uint64_t)i * 0x9E3779B97F4A7C15ull));
            ++count[r.next_u64() % nb];
        }
        std::vector<size_t> which;
        for (size_t b = 0; b < nb; ++b) if (count[b]) which.push_back(b);
        std::vector<std::complex<double>> val;
        exact_values(which, val);
        std::complex<double> sum(0, 0);
        double sq = 0;
        for (size_t b : which) { sum += (double)count[b] * val[b]; sq += (double)count[b] * std::norm(val[b]); }
        const double n = (double)samples;
        const std::complex<double> mean = sum / n;
        const double var = std::max(sq / n - std::norm(mean), 0.0);
        out[0] = mean.real() * (double)nb; out[1] = mean.imag() * (double)nb;
        out[2] = (double)nb * std::sqrt(var / n);
        return 0;
    })
}

// Simulator::statevector (sim.rs:201-213): all 2^n amplitudes <y|C|x>, in
// batches of at most 2^16 queries through the batched form.
int pq_sim_statevector(pq_sim* h, uint64_t x, int form, double* out) {
    PQ_TRY({
        const uint32_t n = h->s.ps.nq;
        if (n >= 48) throw std::invalid_argument("statevector of 48 or more qubits");
        if (!qubits_ok(h, x)) throw std::invalid_argument("basis index has bits beyond the qubit count");
        const uint64_t N = 1ull << n;
        constexpr uint64_t kChunk = 1ull << 16;
        std::vector<uint64_t> xs, ys;
        for (uint64_t y0 = 0; y0 < N; y0 += kChunk) {
            const uint64_t m = std::min(kChunk, N - y0);
            xs.assign(m, x); ys.resize(m);
            for (uint64_t i = 0; i < m; ++i) ys[i] = y0 + i;
            const int rc = pq_sim_amplitudes(h, m, xs.data(), ys.data(), form, out + 2 * y0);
            if (rc != 0) return rc;
        }
        return 0;
    })
}


// Simulator::sample_many (sim.rs:253-284): every shot draws its qubits in
// order by the chain rule over qubit prefixes, p0 = P(prefix, 0) from the
// cached prefix (marginal) engines, with the shot's own SplitMix64 stream
// seeded from the run's.  All shots advance one qubit at a time: each qubit
// is one batched evaluation over the distinct prefixes of the shots (a
// query's value does not depend on its batch, so shots with the same prefix
// share it).
int pq_sim_sample_many(pq_sim* h, uint64_t x, size_t shots, uint64_t seed, int form, uint64_t* out) {
    PQ_TRY({
        const uint32_t n = h->s.ps.nq;
        if (n > 64) throw std::invalid_argument("sampling more than 64 qubits");
        if (!qubits_ok(h, x)) throw std::invalid_argument("basis index has bits beyond the qubit count");
        SplitMix top(seed);
        std::vector<SplitMix> rng;
        rng.reserve(shots);
        for (size_t i = 0; i < shots; ++i) rng.emplace_back(top.next_u64());
        std::vector<uint64_t> y(shots, 0);
        std::vector<double> pprev(shots, 1.0);
        std::vector<uint32_t> qs;
        std::vector<uint8_t> outc;
        std::vector<std::pair<uint32_t, uint8_t>> a;
        for (uint32_t l = 0; l < n; ++l) {
            qs.push_back(l);
            const auto Rp = h->s.marginal_engine(qs);       // prefix of l + 1 qubits
            const RankEng& R = *Rp;
            // distinct prefixes y & (2^l - 1), with qubit l = 0
            const uint64_t mask = l == 0 ? 0 : (l >= 64 ? ~0ull : ((1ull << l) - 1));
            std::map<uint64_t, size_t> slot;
            std::vector<uint64_t> prefixes;
            std::vector<size_t> of(shots);
            for (size_t i = 0; i < shots; ++i) {
                auto it = slot.emplace(y[i] & mask, prefixes.size());
                if (it.second) prefixes.push_back(y[i] & mask);
                of[i] = it.first->second;
            }
            std::vector<double> p0(prefixes.size(), 0.0);
            std::vector<Z8Job> jobs;
            std::vector<size_t> owner;
            const int walk = form == PQ_FORM_PLAIN_FP64 ? 0 : walk_of_form(form);
            if (form != PQ_FORM_PLAIN_FP64) R.ensure_device();
            outc.assign(l + 1, 0);
            for (size_t p = 0; p < prefixes.size(); ++p) {
                for (uint32_t j = 0; j < l; ++j) outc[j] = (uint8_t)((prefixes[p] >> j) & 1);
                outc[l] = 0;
                if (!h->s.marginal_assignment(x, qs.data(), outc.data(), l + 1, a)) continue;   // contradictory: 0
                if (form == PQ_FORM_PLAIN_FP64) p0[p] = std::max(plain_eval(R, a).real(), 0.0);
                else add_query_jobs(R, a, walk, jobs, owner, p);
            }
            if (form != PQ_FORM_PLAIN_FP64) {
                std::vector<std::complex<double>> res;
                run_jobs(jobs, owner, prefixes.size(), walk, res);
                for (size_t p = 0; p < prefixes.size(); ++p) p0[p] = std::max(res[p].real(), 0.0);   // sim.rs:248
            }
            for (size_t i = 0; i < shots; ++i) {
                const double q0 = std::min(p0[of[i]], pprev[i]);
                if (rng[i].next_f64() * pprev[i] >= q0) {
                    y[i] |= 1ull << l;
                    pprev[i] = std::max(pprev[i] - q0, 0.0);
                } else {
                    pprev[i] = q0;
                }
            }
        }
        for (size_t i = 0; i < shots; ++i) out[i] = y[i];
        return 0;
    })
}

int pq_batch_amplitudes(size_t n, pq_sim* const* sims, const uint64_t* xs, const uint64_t* ys, int form, double* out) {
    PQ_TRY({
        const int walk = walk_of_form(form);
        std::vector<Z8Job> jobs; std::vector<size_t> owner;
        build_jobs(n, walk, [&](size_t i) { return (const RankEng*)sims[i]->s.eng.get(); }, [&](size_t i) {
            if (!qubits_ok(sims[i], xs[i]) || !qubits_ok(sims[i], ys[i])) throw std::invalid_argument("basis index has bits beyond the qubit count");
        }, xs, ys, jobs, owner);
        std::vector<std::complex<double>> res;
        run_jobs(jobs, owner, n, walk, res);
        for (size_t i = 0; i < n; ++i) { out[2 * i] = res[i].real(); out[2 * i + 1] = res[i].imag(); }
        return 0;
    })
}

int pq_set_walk_tuning(int form, uint32_t offset, uint32_t max_mid) {
    PQ_TRY({
        if (walk_of_form(form) == kWalkSeg) seg_set_tuning(offset, max_mid);
        else z8_set_tuning(offset, max_mid);
        return 0;
    })
}
int pq_get_walk_tuning(int form, uint32_t* o) {
    PQ_TRY({
        if (walk_of_form(form) == kWalkSeg) seg_get_tuning(&o[0], &o[1]);
        else z8_get_tuning(&o[0], &o[1]);
        return 0;
    })
}
int pq_set_walk_shape(int form, uint32_t waves) {
    PQ_TRY({
        if (walk_of_form(form) == kWalkSeg) seg_set_waves(waves);
        else if (waves != 4) throw std::invalid_argument("the branch walk runs workgroups of 4 waves");
        return 0;
    })
}
int pq_get_walk_shape(int form, uint32_t* waves) {
    PQ_TRY({
        *waves = walk_of_form(form) == kWalkSeg ? seg_get_waves() : 4u;
        return 0;
    })
}
int pq_set_segment_bits(uint32_t bits) { PQ_TRY({ seg_set_bits(bits); return 0; }) }
uint32_t pq_get_segment_bits(void) { return seg_get_bits(); }

int pq_last_timing(double* o) { for (int i = 0; i < 5; ++i) o[i] = g_timing[i]; return 0; }
int pq_set_timing(int on) { dev_set_timing(on != 0); return 0; }
int pq_get_timing(void) { return dev_get_timing() ? 1 : 0; }

int pq_graph_create(size_t n, pq_sim* const* sims, const uint64_t* xs, const uint64_t* ys, int form, pq_graph** out) {
    PQ_TRY({
        const int walk = walk_of_form(form);
        auto g = std::make_unique<pq_graph>();
        std::vector<Z8Job> jobs;
        for (size_t i = 0; i < n; ++i) g->keep.push_back(sims[i]->s.eng);
        build_jobs(n, walk, [&](size_t i) { return (const RankEng*)sims[i]->s.eng.get(); }, [&](size_t i) {
            if (!qubits_ok(sims[i], xs[i]) || !qubits_ok(sims[i], ys[i])) throw std::invalid_argument("basis index has bits beyond the qubit count");
        }, xs, ys, jobs, g->owner);
        if (jobs.empty()) throw std::invalid_argument("every query of the graph is identically zero");
        assign_mid_bits(jobs, walk);
        g->nq = n;
        g->g = dev_z8_graph(jobs, walk);
        *out = g.release();
        return 0;
    })
}

int pq_graph_run(pq_graph* g, double* out, double* total_ms) {
    PQ_TRY({
        std::vector<std::complex<double>> r;
        Z8Timing T;
        dev_z8_graph_run(*g->g, r, &T);
        std::vector<std::complex<double>> acc(g->nq, {0, 0});
        for (size_t i = 0; i < r.size(); ++i) acc[g->owner[i]] += r[i];
        for (size_t i = 0; i < g->nq; ++i) { out[2 * i] = acc[i].real(); out[2 * i + 1] = acc[i].imag(); }
        if (total_ms) *total_ms = T.total_ms;
        return 0;
    })
}

void pq_graph_destroy(pq_graph* g) { delete g; }

int pq_sim_branch_dump(pq_sim* h, uint64_t x, uint64_t y, size_t n, const uint64_t* s, uint8_t* zero, uint8_t* k8, uint8_t* ktot, uint32_t* r) {
    PQ_TRY({
        std::vector<std::pair<uint32_t, uint8_t>> asg;
        if (!h->s.assignment(x, y, asg)) return 1;
        const RankEng& R = *h->s.eng;
        const HostEngine& E = *R.host[0];
        if (E.zero) throw std::invalid_argument("path sum vanishes identically");
        diag_terms_ok(E);
        Prepared P;
        prepare_assignment(E, asg, P);   // u0 is filled even when K s = kappa is inconsistent
        if (P.u0.empty()) throw std::logic_error("prepare did not produce the boundary shift");
        if (r) *r = 2 * E.L.g + 2 * E.L.neven + E.L.nodd;
        R.ensure_device();
        dev_dump_s(*R.dev[0], P, s, n, zero, k8, ktot);
        return 0;
    })
}

int pq_sim_walk_dump(pq_sim* h, uint64_t x, uint64_t y, uint32_t mid_bits, size_t cap, uint8_t* k, float* mag, uint64_t* s, size_t* n) {
    PQ_TRY({
        std::vector<std::pair<uint32_t, uint8_t>> asg;
        *n = 0;
        if (!h->s.assignment(x, y, asg)) return 1;
        const RankEng& R = *h->s.eng;
        const HostEngine& E = *R.host[0];
        const Z8Image& Z = *R.z8[0];
        diag_terms_ok(E);
        Prepared P;
        if (!prepare_assignment(E, asg, P)) return 1;
        Z8QueryHost q = z8_query(E, Z, P);
        const uint32_t K = Z.kin, nout = Z.nslots - K, M = std::min(mid_bits, nout), H = nout - M;
        const size_t total = (size_t)1 << (K + M + H);
        if (total > cap) throw std::invalid_argument("walk dump buffer too small");
        R.ensure_device();
        dev_dump_walk(*R.dev[0], q, M, k, mag);
        for (size_t idx = 0; idx < total; ++idx) {
            const uint64_t i = idx & ((1ull << K) - 1), o = (idx >> K) & ((1ull << M) - 1), hi = idx >> (K + M);
            const uint64_t gi = i ^ (i >> 1), go = o ^ (o >> 1);
            uint64_t sv = P.s0;
            for (uint32_t p = 0; p < Z.nslots; ++p) {
                bool bit = p < K ? ((gi >> p) & 1) : (p < K + M ? ((go >> (p - K)) & 1) : ((hi >> (p - K - M)) & 1));
                if (bit && Z.slot_real[p]) sv ^= Z.slot_n[p];
            }
            s[idx] = sv;
        }
        *n = total;
        return 0;
    })
}

int pq_sim_selfcheck(pq_sim* h, uint64_t x, uint64_t y, uint32_t trials, uint64_t seed, uint64_t* bad) {
    PQ_TRY({
        std::vector<std::pair<uint32_t, uint8_t>> asg;
        *bad = 0;
        if (!h->s.assignment(x, y, asg)) return 1;
        for (size_t b = 0; b < h->s.eng->host.size(); ++b) {
            const HostEngine& E = *h->s.eng->host[b];
            diag_terms_ok(E);
            Prepared P;
            if (!prepare_assignment(E, asg, P)) continue;
            prepared_ell(E, P);
            const uint32_t W = E.L.words();
            uint64_t st = seed + 0x9E3779B97F4A7C15ull * (b + 1);
            for (uint32_t tr = 0; tr < trials; ++tr) {
                st += 0x9E3779B97F4A7C15ull;
                uint64_t z = st; z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; z ^= z >> 31;
                const uint64_t c = E.d >= 64 ? z : (z & ((1ull << E.d) - 1));
                std::vector<uint64_t> u(P.u_start);
                uint64_t sv = P.s0;
                uint32_t kq = P.K0;
                for (uint32_t j = 0; j < E.d; ++j) if ((c >> j) & 1) {
                    for (uint32_t w = 0; w < W; ++w) u[w] ^= E.kernel_u[(size_t)j * W + w];
                    sv ^= E.kernel[j];
                    kq += P.ell[j] + 4 * ((uint32_t)__builtin_popcountll(E.Q[j] & c & ((1ull << j) - 1)) & 1);
                }
                if (!survives_row(E, u.data()) || total_phase(E, u.data(), sv) != (kq & 7)) (*bad)++;
            }
        }
        return 0;
    })
}

int pq_sim_marginal_stats(pq_sim* h, const uint32_t* qubits, size_t k, uint64_t* o) {
    PQ_TRY({
        for (size_t i = 0; i < k; ++i) if (qubits[i] >= h->s.ps.nq) throw std::invalid_argument("measured qubit out of range");
        std::vector<uint32_t> qs(qubits, qubits + k);
        const auto Rp = h->s.marginal_engine(qs);
        const RankEng& R = *Rp;
        EngineStats s = R.host[0]->stats();
        o[0] = s.free_vars; o[1] = s.planes; o[2] = s.lines; o[3] = s.radical; o[4] = s.terms; o[5] = s.branch_dim; o[6] = R.pairs;
        return 0;
    })
}

int pq_sim_prologue_selfcheck(pq_sim* h, size_t n, const uint64_t* xs, const uint64_t* ys, uint64_t* bad) {
    PQ_TRY({
        *bad = 0;
        const RankEng& R = *h->s.eng;
        std::vector<std::pair<uint32_t, uint8_t>> asg;
        Prepared A, B;
        for (size_t i = 0; i < n; ++i) {
            const bool ok = h->s.assignment(xs[i], ys[i], asg);
            for (size_t b = 0; b < R.host.size(); ++b) {
                bool contra = false;
                const bool nb = prepare_bits(*R.host[b], R.bprog, xs[i], ys[i], B, &contra);
                if (!ok) { if (!contra || nb) (*bad)++; continue; }
                const bool na = prepare_assignment(*R.host[b], asg, A);
                if (contra || na != nb) { (*bad)++; continue; }
                if (!na) continue;
                if (A.phase != B.phase || A.s0 != B.s0 || A.s0_rest != B.s0_rest || A.K0 != B.K0 || A.u_start != B.u_start || A.dq != B.dq ||
                    A.scale != B.scale)
                    (*bad)++;
            }
        }
        return 0;
    })
}

int pq_sim_seg_info(pq_sim* h, uint64_t* o) {
    PQ_TRY({
        const SegImage& G = *h->s.eng->seg[0];
        uint32_t off, mx;
        seg_get_tuning(&off, &mx);
        o[0] = G.d; o[1] = G.I; o[2] = G.npin; o[3] = G.nout; o[4] = G.L6; o[5] = G.Mcap; o[6] = G.table.size() * 4;
        o[7] = seg_mid_bits(G, off, mx);
        return 0;
    })
}

// Shared set-up of the segment dumps: query data and sizes for <y|C|x>.
static bool seg_dump_setup(pq_sim* h, uint64_t x, uint64_t y, uint32_t mid_bits, size_t cap, SegQueryHost& q, uint32_t& M, size_t& total) {
    std::vector<std::pair<uint32_t, uint8_t>> asg;
    if (!h->s.assignment(x, y, asg)) return false;
    const RankEng& R = *h->s.eng;
    const HostEngine& E = *R.host[0];
    const SegImage& G = *R.seg[0];
    Prepared P;
    if (!prepare_assignment(E, asg, P)) return false;
    q = seg_query(E, G, P);
    M = std::min(mid_bits, std::min(G.Mcap, G.nout - G.L6));
    if (G.nout > 24) throw std::invalid_argument("segment dumps are limited to 2^24 outer points");
    total = (size_t)1 << G.nout;
    if (total > cap) throw std::invalid_argument("segment dump buffer too small");
    return true;
}

int pq_sim_seg_dump(pq_sim* h, uint64_t x, uint64_t y, uint32_t mid_bits, size_t cap, uint8_t* k, uint32_t* hs, uint32_t* pio, size_t* n) {
    PQ_TRY({
        SegQueryHost q; uint32_t M; size_t total;
        *n = 0;
        if (!seg_dump_setup(h, x, y, mid_bits, cap, q, M, total)) return 1;
        const RankEng& R = *h->s.eng;
        R.ensure_device();
        dev_dump_seg(*R.dev[0], q, M, k, hs, pio);
        *n = total;
        return 0;
    })
}

int pq_sim_seg_lane_ref(pq_sim* h, uint64_t x, uint64_t y, uint32_t mid_bits, size_t cap, uint8_t* k, uint32_t* hs, uint32_t* pio, size_t* n) {
    PQ_TRY({
        SegQueryHost q; uint32_t M; size_t total;
        *n = 0;
        if (!seg_dump_setup(h, x, y, mid_bits, cap, q, M, total)) return 1;
        const SegImage& G = *h->s.eng->seg[0];
        for (size_t idx = 0; idx < total; ++idx) {
            const SegLane L = seg_lane_state(G, q, M, idx >> M, (uint32_t)(idx & ((1ull << M) - 1)));
            k[idx] = (uint8_t)L.K; hs[idx] = L.hs; pio[idx] = L.pio;
        }
        *n = total;
        return 0;
    })
}

int pq_sim_seg_selfcheck(pq_sim* h, uint64_t x, uint64_t y, uint32_t trials, uint64_t seed, double* max_err) {
    PQ_TRY({
        std::vector<std::pair<uint32_t, uint8_t>> asg;
        *max_err = 0;
        if (!h->s.assignment(x, y, asg)) return 1;
        static const double Hh = 0.70710678118654752440;
        const double w8r[8] = {1, Hh, 0, -Hh, -1, -Hh, 0, Hh}, w8i[8] = {0, Hh, 1, Hh, 0, -Hh, -1, -Hh};
        for (size_t b = 0; b < h->s.eng->host.size(); ++b) {
            const HostEngine& E = *h->s.eng->host[b];
            const SegImage& G = *h->s.eng->seg[b];
            diag_terms_ok(E);
            Prepared P;
            if (!prepare_assignment(E, asg, P)) continue;
            const SegQueryHost q = seg_query(E, G, P);
            const uint32_t W = E.L.words(), M = 0;
            uint64_t st = seed + 0x9E3779B97F4A7C15ull * (b + 1);
            for (uint32_t tr = 0; tr < trials; ++tr) {
                st += 0x9E3779B97F4A7C15ull;
                uint64_t z = st; z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; z ^= z >> 31;
                const uint64_t lane = G.nout >= 64 ? z : (z & ((1ull << G.nout) - 1));
                const SegLane L = seg_lane_state(G, q, M, lane, 0);
                // segment value from the image: omega^K T[pio][hs] * lane magnitude (FP64 from hi + lo)
                const float* e = &G.table[(((size_t)L.pio << G.I) | L.hs) * 4];
                double fl = 1.0;
                for (uint32_t k = 0; k < E.t; ++k) {
                    const bool inner_or_pin = !((G.lane_terms >> k) & 1);
                    bool walk_free = false;
                    for (uint32_t m = 0; m < G.Mcap; ++m) if (E.free_col[G.outer_j[G.L6 + m]] == k) walk_free = true;
                    if (inner_or_pin && !walk_free) continue;
                    fl *= ((L.s >> k) & 1) ? E.r1[k] : E.r0[k];
                }
                const std::complex<double> seg = std::complex<double>(((double)e[0] + (double)e[1]) * fl, ((double)e[2] + (double)e[3]) * fl) *
                                                 std::complex<double>(w8r[L.K], w8i[L.K]);
                // direct sum over the 2^I inner branches with the per-branch phase of the packed row
                std::complex<double> dir(0, 0);
                double mass = 0;
                std::vector<uint64_t> u0(P.u_start);
                uint64_t s0 = P.s0;
                for (uint32_t k = 0; k < G.nout; ++k) if ((L.c >> k) & 1) {
                    const uint32_t j = G.outer_j[k];
                    for (uint32_t w = 0; w < W; ++w) u0[w] ^= E.kernel_u[(size_t)j * W + w];
                    s0 ^= E.kernel[j];
                }
                for (uint32_t ci = 0; ci < (1u << G.I); ++ci) {
                    std::vector<uint64_t> u(u0);
                    uint64_t sv = s0;
                    for (uint32_t p = 0; p < G.I; ++p) if ((ci >> p) & 1) {
                        const uint32_t j = G.inner_j[p];
                        for (uint32_t w = 0; w < W; ++w) u[w] ^= E.kernel_u[(size_t)j * W + w];
                        sv ^= E.kernel[j];
                    }
                    if (!survives_row(E, u.data())) throw std::logic_error("branch outside the surviving subspace");
                    const uint32_t kk = total_phase(E, u.data(), sv);
                    double mg = 1.0;
                    for (uint32_t k = 0; k < E.t; ++k) mg *= ((sv >> k) & 1) ? E.r1[k] : E.r0[k];
                    dir += std::complex<double>(mg * w8r[kk], mg * w8i[kk]);
                    mass += mg;
                }
                *max_err = std::max(*max_err, std::abs(dir - seg) / std::max(mass, 1e-300));
            }
        }
        return 0;
    })
}

int pq_sim_z8_info(pq_sim* h, uint32_t* o) {
    PQ_TRY({
        const Z8Image& Z = *h->s.eng->z8[0];
        o[0] = Z.d; o[1] = Z.nin; o[2] = Z.nslots; o[3] = Z.npin; o[4] = h->s.eng->host[0]->rk; o[5] = Z.imag0 ? 1 : 0; o[6] = Z.ops_per_branch();
        o[7] = Z.nx;
        o[8] = Z.low_pivot_free ? 1u : 0u;
        o[9] = Z.kin;
        return 0;
    })
}

}  // extern "C"
