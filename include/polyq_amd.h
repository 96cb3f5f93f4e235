/* polyq_amd: MI355X (gfx950) branch evaluator for Clifford+T path sums.
 *
 * C ABI of the drop-in GPU strategy.  Every entry point takes plain pointers
 * and sizes, returns 0 on success and a negative status on error (message in
 * pq_last_error(), thread local), and never aborts or throws across the ABI.
 * An amplitude that is exactly zero because the boundary is contradictory,
 * the path sum vanishes identically, or K s = kappa is inconsistent is a
 * successful result 0 (src/pathsum.rs:430-450, src/engine.rs:570-572,
 * :595-597), not an error.
 *
 * Conventions: basis-state index bit q = qubit q (src/sim.rs:74-76,
 * src/circuit.rs:18-19); gate codes follow the Gate enum of
 * src/circuit.rs:25-42; a phase is an angle 2*pi*k/2^32 (src/phase.rs:22-24).
 *
 * Handles are immutable after creation apart from internal caches guarded by
 * locks, so concurrent calls from several host threads are safe; device work
 * is serialised on the library's stream.
 */
#ifndef POLYQ_AMD_H
#define POLYQ_AMD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Gate codes (src/circuit.rs:25-42). */
enum pq_gate {
    PQ_H = 0, PQ_X = 1, PQ_Y = 2, PQ_Z = 3, PQ_S = 4, PQ_SDG = 5, PQ_T = 6, PQ_TDG = 7,
    PQ_P = 8,       /* diag(1, e^{i phase}) */
    PQ_CZ = 9, PQ_CX = 10 /* q0 control, q1 target */, PQ_SWAP = 11,
    PQ_GPHASE = 12  /* global phase e^{i phase} */
};

/* Non-Clifford strategy (src/sim.rs:40-48): Auto picks the number of
 * decomposed T-pairs with the cost model of src/rank.rs:205-222, Cached is the
 * plain Z-branch engine (2^d branches per amplitude), Pairs decomposes exactly
 * `pairs` pairs (2^p stabiliser branch engines, src/rank.rs:88-103). */
enum pq_strategy { PQ_AUTO = 0, PQ_CACHED = 1, PQ_PAIRS = 2,
    /* Auto with the cost model refit to the GPU (no counterpart in the
     * reference): 2^p host engine builds against 2^(d - removed) branches at
     * the device's walk rate; picks p = 0 unless d is far beyond 40. */
    PQ_GPU_AUTO = 3 };

/* Device evaluation form.
 *  Z8 (default): segment walk.  The sum over 2^I inner kernel coordinates is
 *    a per-engine Walsh-Hadamard table (built once in FP64, stored as FP32
 *    hi/lo pairs); lanes walk the outer coordinates in Gray order with integer
 *    Z8 phases and add one table entry per outer point in double-float (FP32
 *    pair) arithmetic; FP64 only in the cross-workgroup reduction and scaling.
 *  BRANCH_WALK: the per-branch Z8 walk (one FMA per branch over a 256-branch
 *    inner table, FP32 two-sum compensated lane accumulators).
 *  PLAIN_FP64: a straightforward FP64 port of Engine::gray, one query per
 *    launch (src/engine.rs:669-722). */
enum pq_form { PQ_FORM_Z8 = 0, PQ_FORM_BRANCH_WALK = 1, PQ_FORM_PLAIN_FP64 = 2 };

typedef struct pq_sim pq_sim;

/* Simulator::with_options (src/sim.rs:144-148): compile the circuit
 * (src/pathsum.rs:274-297), build the branch engines on the host
 * (src/engine.rs:289-390, src/rank.rs:88-123) and upload them to HBM, where
 * they stay resident until pq_sim_destroy.  phase[] is read for PQ_P and
 * PQ_GPHASE only; q1[] for two-qubit gates only. */
int pq_sim_create(uint32_t num_qubits, size_t num_gates, const uint8_t* ops, const uint32_t* q0, const uint32_t* q1,
                  const uint32_t* phase, int strategy, uint32_t pairs, pq_sim** out);
void pq_sim_destroy(pq_sim* sim);

/* Simulator::stats + pairs (src/sim.rs:161-168; order of the Python tuple,
 * src/pybind.rs:177-180): free_vars, planes, lines, radical, terms,
 * branch_dim, pairs. */
int pq_sim_stats(const pq_sim* sim, uint64_t out7[7]);
uint32_t pq_sim_num_qubits(const pq_sim* sim);

/* Simulator::amplitude (src/sim.rs:171-176): <output|C|input>, one byte per
 * qubit (0/1), any number of qubits.  Circuits above 64 qubits are served by
 * this and the other byte-array entry points only: the u64 basis-index entry
 * points refuse them. */
int pq_sim_amplitude(pq_sim* sim, const uint8_t* input, const uint8_t* output, int form, double out[2]);
/* Simulator::probability (src/sim.rs:196-198) with byte-array basis states. */
int pq_sim_probability_bytes(pq_sim* sim, const uint8_t* input, const uint8_t* output, int form, double* out);
/* Simulator::marginal_probability (src/sim.rs:225-231) with a byte-array input. */
int pq_sim_marginal_probability_bytes(pq_sim* sim, const uint8_t* input, const uint32_t* qubits, const uint8_t* outcomes, size_t k,
                                      int form, double* out);
/* Simulator::amplitude_index (src/sim.rs:179-182). */
int pq_sim_amplitude_index(pq_sim* sim, uint64_t input, uint64_t output, int form, double out[2]);
/* Simulator::probability (src/sim.rs:196-198) = |amplitude|^2. */
int pq_sim_probability(pq_sim* sim, uint64_t input, uint64_t output, int form, double* out);
/* Simulator::marginal_probability (src/sim.rs:225-231): probability that the
 * k listed qubits measure the listed outcomes; sandwich engine
 * (src/pathsum.rs:304-314), cached per measured-qubit set (the reference
 * rebuilds it on every call). */
int pq_sim_marginal_probability(pq_sim* sim, uint64_t input, const uint32_t* qubits, const uint8_t* outcomes, size_t k,
                                int form, double* out);
/* Simulator::amplitude_estimate (src/sim.rs:187-193, pybind.rs:149-153):
 * unbiased Monte-Carlo estimate of <y|C|x> from `samples` points
 * (Engine::estimate, src/engine.rs:612-659, or random branch engines,
 * src/rank.rs:149-171); out = {re, im, standard error}.  The exact sum when
 * 2^d <= samples. */
int pq_sim_amplitude_estimate(pq_sim* sim, uint64_t input, uint64_t output, size_t samples, uint64_t seed, int form, double* out3);
/* Simulator::statevector (src/sim.rs:201-213): all 2^n amplitudes <y|C|x>,
 * out holds 2 * 2^n doubles (re, im by y); n < 48. */
int pq_sim_statevector(pq_sim* sim, uint64_t input, int form, double* out);
/* Simulator::sample_many (src/sim.rs:253-284, pybind.rs:169-172): shots
 * measurement outcomes of all qubits as basis indices, by the chain rule over
 * qubit prefixes with the reference's per-shot SplitMix64 streams. */
int pq_sim_sample_many(pq_sim* sim, uint64_t input, size_t shots, uint64_t seed, int form, uint64_t* out);
/* Many amplitudes <outputs[i]|C|inputs[i]> of one simulator in one launch
 * (the query loop of Simulator::statevector, src/sim.rs:201-213). */
int pq_sim_amplitudes(pq_sim* sim, size_t n, const uint64_t* inputs, const uint64_t* outputs, int form, double* out);

/* ---- engine-level seam: one Engine (src/engine.rs:273-284) handed over by
 * the fields it holds, so a Strategy::Gpu arm in src/rank.rs keeps its own
 * path-sum compiler and normal form and evaluates its branch engines here
 * (the call chain Strategy -> RankEngine::eval_assignment -> Engine::eval_with
 * -> sum_subspace, src/sim.rs:65-71, src/rank.rs:184-190,
 * src/engine.rs:557-561, :669).
 *
 * Rows are packed F2 vectors, LSB-first in u64 words, tail bits zero (BitVec,
 * src/bits.rs:10-19).  A shift row (UVec, src/normal_form.rs:161-201) is its
 * four parts a | b | o | e, each in ceil(len/64) words, for part lengths
 * planes, planes, lines, radical.  The caller owns every array; create copies
 * what it needs. */
typedef struct pq_engine_desc {
    uint32_t nf;               /* fixed variables (Engine::fixed order) */
    uint32_t planes, lines, radical;   /* Structure: g, #odd, #even */
    uint32_t terms;            /* rotation terms t (Branching::term_vars) */
    uint32_t branch_dim;       /* kernel vectors d (Elim::kernel) */
    uint32_t rank;             /* RREF pivots (Elim::pivots) */
    uint32_t consistency_rows; /* Elim::consistency */
    uint32_t num_h;            /* Engine::num_h */
    uint32_t global_phase;     /* Engine::global (Phase) */
    int32_t zero;              /* Engine::zero */
    int32_t sigma;             /* Diag::sigma = #o1 - #o3 */
    uint32_t free_vars;        /* Core::free.len() (statistics only) */
    const uint64_t* ucols;     /* nf shift rows (Core::ucols) */
    const uint64_t* mcols;     /* terms shift rows (Branching::mcols) */
    const uint64_t* diag;      /* one shift-shaped row: Diag wa | wb | o1 | we */
    const uint64_t* o3;        /* Diag::o3 over the lines */
    const uint8_t* fixed_v;    /* nf Z4 linear coefficients (Engine::fixed_v) */
    const uint32_t* fixed_theta;   /* nf phases (Engine::fixed_theta) */
    const uint64_t* b_fixed;   /* nf rows of ceil(nf/64) words (Core::b_fixed) */
    const uint32_t* theta;     /* terms phases (the Term angles) */
    const uint32_t* pivot_col; /* rank increasing pivot columns (Elim::pivots .0) */
    const uint64_t* pivot_row; /* rank rows over the radical (Elim::pivots .1) */
    const uint64_t* consistency;   /* consistency_rows rows over the radical */
    const uint64_t* kernel;    /* branch_dim rows of ceil(terms/64) words (Elim::kernel) */
    const uint64_t* kernel_u;  /* branch_dim shift rows (Elim::kernel_u), or NULL: then derived from mcols */
} pq_engine_desc;
typedef struct pq_engine pq_engine;

/* Replaces the Engine the Rust side built (src/engine.rs:289-390) by a GPU
 * engine built from its fields; the descriptor is checked for consistency
 * (RREF shape, kernel inside ker K, kernel_u = xor of mcols, Diag lines,
 * symmetric b_fixed, zero tails).  The device images are uploaded on first
 * evaluation and stay until pq_engine_destroy.  Any number of terms, of which
 * at most 64 may vary over the branch subspace (the others are constant and
 * folded into the prologue). */
int pq_engine_create(const pq_engine_desc* desc, pq_engine** out);
void pq_engine_destroy(pq_engine* engine);
/* Engine::stats (src/engine.rs:415-425): free_vars, planes, lines, radical,
 * terms, branch_dim. */
int pq_engine_stats(const pq_engine* engine, uint64_t out6[6]);
/* Engine::eval (src/engine.rs:551-561): values aligned with the fixed
 * variables, one byte each (0/1). */
int pq_engine_eval(pq_engine* engine, const uint8_t* values, size_t nf, int form, double out[2]);
/* Engine::eval for n (engine, values) queries in one device launch (batched
 * forms; the plain form runs them one by one).  A RankEngine's eval_assignment
 * is the sum of its 2^p branch engines' values for the same fixed values
 * (src/rank.rs:184-190).  out has 2n doubles. */
int pq_engine_eval_batch(size_t n, pq_engine* const* engines, const uint8_t* const* values, int form, double* out);
/* Engine::sum_subspace (src/engine.rs:669-693): the sum of
 * weight(s) ROOT8[phase8(u(s))] over the affine branch subspace
 * s_start + span(kernel), walked from the start shift u_start (one shift row,
 * packed as the descriptor's rows) and branch vector s_start (ceil(terms/64)
 * words, the reference's term order): what prepare leaves in sc.u and sc.s0
 * (src/engine.rs:596-606).  c_start[2] is prepare's c_start, the weight at
 * s_start; the device derives every branch's weight from the term angles, as
 * the reference does for every chunk but the first, and a c_start that is not
 * that weight (relative 1e-9) is an error.  eval_with multiplies the result
 * by the boundary phase and 2^(log2_magnitude - h/2) (src/engine.rs:557-565).
 * An engine whose path sum vanishes identically gives 0. */
int pq_engine_sum_subspace(pq_engine* engine, const uint64_t* u_start, const uint64_t* s_start, const double c_start[2], int form,
                           double out[2]);
/* Per-branch Gauss-sum outputs of the engine for explicit t-bit branch vectors
 * s, as pq_sim_branch_dump (normal_form.rs:472-493). */
int pq_engine_branch_dump(pq_engine* engine, const uint8_t* values, size_t n, const uint64_t* s, uint8_t* zero, uint8_t* k8,
                          uint8_t* ktot, uint32_t* r);

/* Many amplitudes, as pq_sim_amplitudes (batched forms only), allocating
 * nothing: no counterpart in the reference; for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The staging records, work-queue counters, item
 * partials and results live in a fixed device array of
 * pq_noalloc_scratch_bytes() (8 MiB) per GPU, held until the process ends; a
 * batch that needs more is rejected with an error.  The simulator's engines
 * must already be resident (any earlier evaluation through an allocating entry
 * point uploads them), else an error.  No device or pinned host memory is
 * allocated (host-side planning uses ordinary heap memory).  Calls on one GPU
 * share the array: the library serialises them, and each call overwrites the
 * previous call's contents.  Results are bitwise those of pq_sim_amplitudes. */
int pq_sim_amplitudes_noalloc(pq_sim* sim, size_t n, const uint64_t* inputs, const uint64_t* outputs, int form, double* out);
size_t pq_noalloc_scratch_bytes(void);

/* Heterogeneous batch: query i is <outputs[i]|sims[i]|inputs[i]>; all
 * branch-engine sums of all queries run in one persistent launch (Z8 forms
 * only).  out has 2n doubles. */
int pq_batch_amplitudes(size_t n, pq_sim* const* sims, const uint64_t* inputs, const uint64_t* outputs, int form, double* out);

/* The number of T-pairs Strategy::Auto decomposes (src/rank.rs:205-222) for
 * an engine with the given stats (free_vars, planes, lines, radical, terms,
 * branch_dim), available pairs and T-class pairs; gpu_model != 0 evaluates
 * the refit of PQ_GPU_AUTO instead.  Returns p >= 0, or -1. */
int pq_best_pairs(const uint64_t st6[6], uint32_t avail, uint32_t tclass, int gpu_model);

/* Lane walk length of a batched form (no counterpart in the reference).  A
 * lane walks 2^M outer Gray points, M = clamp(n - offset, 0, max_mid), with n
 * the outer coordinates of the query (PQ_FORM_Z8: d - I; PQ_FORM_BRANCH_WALK:
 * d minus the engine's inner slots, pq_sim_z8_info).  Set by the autotuner; a
 * fixed setting makes every result independent of the batch it runs in. */
int pq_set_walk_tuning(int form, uint32_t offset, uint32_t max_mid);
int pq_get_walk_tuning(int form, uint32_t out2[2]);
/* Workgroup shape of a batched form (no counterpart in the reference): waves
 * of 64 lanes per workgroup, and so lanes per work item.  PQ_FORM_Z8 takes 1,
 * 2 or 4 (default 4); PQ_FORM_BRANCH_WALK only 4 (its register-bound
 * occupancy is 2 workgroups of 4 waves per CU).  Global like the walk
 * length, and set by the autotuner with it. */
int pq_set_walk_shape(int form, uint32_t waves);
int pq_get_walk_shape(int form, uint32_t* waves);
/* Cap on the table bits I + npin of the segment tables of engines built
 * afterwards (default 16).  An engine of branch dimension d gets
 * min(cap, max(14, d - 14)) bits: 2^bits entries x 16 B, so 256 KB at 14 bits
 * and 1 MB at 16. */
int pq_set_segment_bits(uint32_t bits);
uint32_t pq_get_segment_bits(void);

/* Timing of the last device evaluation on this thread: kernel_ms, total_ms,
 * branches, items, grid (doubles).  kernel_ms and total_ms come from HIP
 * events and are NaN unless timing is on. */
int pq_last_timing(double out5[5]);
/* Whether eager evaluations record HIP timing events (no counterpart in the
 * reference; default off: each event is a queue barrier that adds several
 * microseconds to a short batch). */
int pq_set_timing(int on);
int pq_get_timing(void);

/* Replayable batch captured in a hipGraph: the descriptors are uploaded once
 * when the graph is made; a replay clears the queue counters, runs the walk
 * and, for large batches, the FP64 reduction and result download. */
typedef struct pq_graph pq_graph;
int pq_graph_create(size_t n, pq_sim* const* sims, const uint64_t* inputs, const uint64_t* outputs, int form, pq_graph** out);
int pq_graph_run(pq_graph* g, double* out, double* total_ms);
void pq_graph_destroy(pq_graph* g);

/* ---- request loop (no counterpart in the reference; the serving layer of
 * the GPU strategy).  Circuits are keyed by content (hash, then the gate list
 * compared), compiled and built on the host once, and stay resident in HBM
 * within hbm_budget_bytes (LRU eviction of idle entries, back-pressure while
 * requests in flight hold the rest; a request that cannot fit fails with an
 * "HBM budget" error instead of waiting).
 * Requests that arrive while batches run are launched as the next batch on
 * another stream (four batch slots).  kind: 0 amplitude <y|C|x>, 1 probability |<y|C|x>|^2,
 * 2 marginal probability of the k listed (qubit, outcome) pairs for input x
 * (src/sim.rs:171-231). */
typedef struct pq_server pq_server;
int pq_server_create(uint64_t hbm_budget_bytes, uint32_t host_workers, int form, pq_server** out);
void pq_server_destroy(pq_server* srv);
int pq_server_submit(pq_server* srv, uint32_t num_qubits, size_t num_gates, const uint8_t* ops, const uint32_t* q0,
                     const uint32_t* q1, const uint32_t* phase, int strategy, int kind, uint64_t x, uint64_t y,
                     const uint32_t* measured_qubits, const uint8_t* outcomes, size_t k, uint64_t* ticket);
/* Waits for a ticket (timeout_s < 0: forever).  Returns 0 with the result in
 * out[0..1] and the submit-to-completion latency, 1 on timeout, -1 if the
 * request failed (message in pq_last_error()). */
int pq_server_wait(pq_server* srv, uint64_t ticket, double timeout_s, double out[2], double* latency_ms);
/* completed, failed, batches, evictions, resident bytes, builds, largest
 * batch, batches launched while another was in flight, walk-kernel ms (timed
 * on every 8th batch and scaled to all batches), resident circuits. */
int pq_server_stats(pq_server* srv, double out10[10]);
/* Host bytes of built engine images kept for idle circuits (default 8 GiB):
 * past it the least recently used circuits with no request in flight are
 * dropped (device images included) and rebuilt if they come back. */
int pq_server_set_host_budget(pq_server* srv, uint64_t bytes);
/* host image bytes held, host budget, circuits dropped by the host LRU. */
int pq_server_host_stats(pq_server* srv, double out3[3]);
/* Where requests spend their time (no counterpart in the reference), over the
 * last 65536 completed requests: out[0] = their count; out[1..6] = mean ms
 * per phase over all of them; out[7..12] = the same over the slowest 1%.
 * Phases: queue (submit to a host worker), host build (compile, normal form,
 * walk images; marginal sandwich engines), upload (admission and copy to
 * HBM), prologue (boundary, prepare, query records), wait for launch, device
 * and finish (launch to result). */
int pq_server_phase_stats(pq_server* srv, double out13[13]);

/* ---- parity and diagnostics (no counterpart in the reference) ---- */
/* Per-branch Gauss-sum outputs on the first branch engine of the cached
 * strategy for <y|C|x>, for explicit t-bit branch vectors s (bit k = term k):
 * zero flag (normal_form.rs:472-474), k8 = phase8 (normal_form.rs:479-493)
 * and the total branch phase k8 + 6|s| + 4|neg \ s| mod 8.  r = 2 log2|G| is
 * constant per engine and returned in *r.  Returns 1 if the boundary is
 * contradictory. */
int pq_sim_branch_dump(pq_sim* sim, uint64_t x, uint64_t y, size_t n, const uint64_t* s, uint8_t* zero, uint8_t* k8,
                       uint8_t* ktot, uint32_t* r);
/* Every branch the Z8 walk visits for <y|C|x> with the given mid bits: k(c)
 * and prod_k r_k[s_k], plus the t-bit branch vector s of each index (host
 * reconstruction from the walk's slot layout).  n = 2^(8 + slots) entries. */
int pq_sim_walk_dump(pq_sim* sim, uint64_t x, uint64_t y, uint32_t mid_bits, size_t cap, uint8_t* k, float* mag, uint64_t* s,
                     size_t* n);
/* Host-only check of the Z8 quadratic form against the direct phase for
 * `trials` random coset points; returns the number of mismatches in *bad. */
int pq_sim_selfcheck(pq_sim* sim, uint64_t x, uint64_t y, uint32_t trials, uint64_t seed, uint64_t* bad);
/* Host-only check of the direct boundary-to-prologue path used by the batched
 * forms against the (variable, value) assignment path (pathsum.rs:430-450,
 * engine.rs:569-610): same zero/contradiction outcome, phase, s0, shift row,
 * Z8 data and scale for each query and branch engine.  *bad counts mismatches. */
int pq_sim_prologue_selfcheck(pq_sim* sim, size_t n, const uint64_t* xs, const uint64_t* ys, uint64_t* bad);
/* Host-only: stats of the marginal (sandwich) engine for a measured-qubit set,
 * in pq_sim_stats order; builds and caches it (src/sim.rs:225-231). */
int pq_sim_marginal_stats(pq_sim* sim, const uint32_t* qubits, size_t k, uint64_t out7[7]);
/* Segment image summary: d, inner directions I, inner pivots, outer
 * coordinates, lane-low slots, walkable slots, table bytes, lane walk length
 * M at the current tuning. */
int pq_sim_seg_info(pq_sim* sim, uint64_t out8[8]);
/* Integer state (K mod 8, table-basis sign mask, inner pivots) at every outer
 * point of the segment walk for <y|C|x>, indexed by (lane << M) + o: from the
 * device walk (seg_dump) and from the host restatement (seg_lane_ref).
 * n = 2^(d - I) entries. */
int pq_sim_seg_dump(pq_sim* sim, uint64_t x, uint64_t y, uint32_t mid_bits, size_t cap, uint8_t* k, uint32_t* hs, uint32_t* pio,
                    size_t* n);
int pq_sim_seg_lane_ref(pq_sim* sim, uint64_t x, uint64_t y, uint32_t mid_bits, size_t cap, uint8_t* k, uint32_t* hs,
                        uint32_t* pio, size_t* n);
/* Host-only check of the segment tables: for `trials` random outer points,
 * the table entry times the lane magnitude against the direct sum of the
 * 2^I inner branches with the packed-row phase (normal_form.rs:479-493);
 * *max_err = max |difference| / sum of branch magnitudes. */
int pq_sim_seg_selfcheck(pq_sim* sim, uint64_t x, uint64_t y, uint32_t trials, uint64_t seed, double* max_err);
/* Branch-walk image summary: d, inner directions, slots, compact inner pivots,
 * rank K, whether inner entries alternate real/imaginary, VALU lane-ops per
 * branch of the inner step, and the pivot-free outer slots that can serve as
 * mid slots with per-query magnitudes (a lane walk of M <= that many mid
 * coordinates takes the kernel's fast path), whether the four low inner
 * slots touch no pivot, and the inner slots per outer point (the walk sums
 * 2^that branches per outer point). */
int pq_sim_z8_info(pq_sim* sim, uint32_t out10[10]);

/* VALU issue-rate probe (no counterpart in the reference): one kernel per
 * instruction (int32 xor, and, add, popcount; FP32 FMA; FP64 FMA; one int32
 * xor per FP64 FMA, the branch walk's mix; FP64 FMA with an SGPR operand; FP64
 * FMA with a DPP row broadcast operand), each lane running 8 independent
 * chains of iters x 64 instructions on blocks_per_cu 256-thread workgroups
 * per CU.  out (3 + 4 x 9 doubles): CUs, wall-clock MHz, blocks per CU, then
 * per op {lane-ops/s, lanes per clock per CU at the in-kernel clock, that
 * clock in GHz, kernel ms}.  Returns the number of ops (9) or -1. */
int pq_valu_probe(double* out, size_t cap, uint32_t blocks_per_cu, uint32_t iters);
const char* pq_valu_probe_name(int op);

const char* pq_last_error(void);
/* Device name, architecture and CU count; returns -1 without a device. */
int pq_device_info(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
