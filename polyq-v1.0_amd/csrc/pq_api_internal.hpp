// Internal host structures shared by the C ABI and the serving loop.
#pragma once
#include "pq_host.hpp"
#include "pq_device.hpp"
#include <functional>
#include <map>
#include <mutex>

namespace pq {

// RankEngine (src/rank.rs:47-51): 2^p branch engines, each with its device image.
struct RankEng {
    std::vector<std::shared_ptr<HostEngine>> host;
    std::vector<std::shared_ptr<Z8Image>> z8;
    std::vector<std::shared_ptr<SegImage>> seg;
    // Device images are uploaded on first device use, so host-only callers
    // (engine statistics, self checks) never need a GPU.
    mutable std::vector<std::shared_ptr<DevEngine>> dev;
    mutable std::mutex dev_mu;
    uint32_t pairs = 0;
    size_t image_bytes = 0;        // device bytes the upload of every engine takes (set by build_rank)
    // boundary map of the amplitude engines (fixed = the circuit's io vars)
    bool has_bprog = false;
    BoundaryProg bprog;
    void ensure_device() const {
        std::lock_guard<std::mutex> lk(dev_mu);
        for (size_t i = dev.size(); i < host.size(); ++i) dev.push_back(dev_engine_upload(host[i], z8[i], seg[i]));
    }
    void drop_device() const {
        std::lock_guard<std::mutex> lk(dev_mu);
        dev.clear();
    }
    bool on_device() const {
        std::lock_guard<std::mutex> lk(dev_mu);
        return dev.size() == host.size();
    }
    size_t device_bytes() const {
        std::lock_guard<std::mutex> lk(dev_mu);
        size_t b = 0;
        for (auto& d : dev) b += dev_engine_bytes(*d);
        return b;
    }
};

std::shared_ptr<RankEng> build_rank(const PathSum& ps, const std::vector<uint32_t>& fixed, int strategy, uint32_t pairs);

// Simulator (src/sim.rs:128-135)
struct Sim {
    Circuit circ;
    PathSum ps;
    std::shared_ptr<RankEng> eng;
    int strategy = 0;
    uint32_t pairs_opt = 0;
    std::mutex mu;
    std::unique_ptr<Sandwich> sw;
    // Marginal (sandwich) engine sets per measured-qubit set, least recently
    // used dropped beyond kMargCap sets.  Callers hold the shared_ptr while
    // they use a set, so a dropped set lives until its last user is done.
    static constexpr size_t kMargCap = 32;
    struct MargSet { std::shared_ptr<RankEng> r; uint64_t last_use = 0; };
    std::map<std::vector<uint32_t>, MargSet> marg;
    uint64_t marg_clock = 0;

    bool assignment(uint64_t x, uint64_t y, std::vector<std::pair<uint32_t, uint8_t>>& asg) const;
    size_t device_bytes() {
        std::lock_guard<std::mutex> lk(mu);
        size_t b = eng ? eng->device_bytes() : 0;
        for (auto& kv : marg) b += kv.second.r->device_bytes();
        return b;
    }
    // Host bytes of the built images (the engine set and the cached marginal
    // sets), the size their device images take.
    size_t host_bytes();
    // Free every device image (host engines stay; the next use re-uploads).
    void drop_device() {
        std::lock_guard<std::mutex> lk(mu);
        if (eng) eng->drop_device();
        for (auto& kv : marg) kv.second.r->drop_device();
    }
    std::shared_ptr<const RankEng> marginal_engine(const std::vector<uint32_t>& qubits);
    bool marginal_assignment(uint64_t x, const uint32_t* qubits, const uint8_t* outcomes, size_t k, std::vector<std::pair<uint32_t, uint8_t>>& a);
    // The same with the input as one byte per qubit (any number of qubits).
    bool marginal_assignment_bytes(const uint8_t* x, const uint32_t* qubits, const uint8_t* outcomes, size_t k,
                                   std::vector<std::pair<uint32_t, uint8_t>>& a);
};

// Error of the u64 basis-index entry points on circuits above 64 qubits.
constexpr const char* kWideCircuit =
    "circuit has more than 64 qubits: basis indices do not fit u64 (use pq_sim_amplitude, pq_sim_probability_bytes or "
    "pq_sim_marginal_probability_bytes)";

void add_jobs(const RankEng& R, const std::vector<std::pair<uint32_t, uint8_t>>& asg, int walk, std::vector<Z8Job>& jobs,
              std::vector<size_t>& owner, size_t qid);
void add_query_jobs_xy(const RankEng& R, uint64_t x, uint64_t y, int walk, std::vector<Z8Job>& jobs, std::vector<size_t>& owner, size_t qid);
void assign_mid_bits(std::vector<Z8Job>& jobs, int walk);
void run_jobs(std::vector<Z8Job>& jobs, const std::vector<size_t>& owner, size_t nq, int walk, std::vector<std::complex<double>>& out);
// Walk of a batched form (PQ_FORM_Z8 -> segment walk, PQ_FORM_BRANCH_WALK ->
// per-branch walk); throws for the plain form.
int walk_of_form(int form);
std::complex<double> plain_eval(const RankEng& R, const std::vector<std::pair<uint32_t, uint8_t>>& asg);

extern thread_local std::string g_err;
extern thread_local double g_timing[5];

}  // namespace pq
