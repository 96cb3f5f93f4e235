// Device-side engine images and launches (implemented in pq_kernels.hip).
// No HIP types appear here, so host C++ compiled by g++ can include it.
#pragma once
#include "pq_host.hpp"
#include <complex>
#include <memory>
#include <vector>

namespace pq {

struct DevEngine;   // engine resident in HBM: plain FP64, branch-walk and segment-table images

// Walk forms of a batch.
enum { kWalkSeg = 0, kWalkBranch = 1 };

// Upload the device images of one engine to the current device.
std::shared_ptr<DevEngine> dev_engine_upload(std::shared_ptr<const HostEngine> E, std::shared_ptr<const Z8Image> Z,
                                             std::shared_ptr<const SegImage> G);
size_t dev_engine_bytes(const DevEngine& D);
// Exact device bytes the upload of (E, Z, G) will allocate (no device needed).
size_t dev_engine_image_bytes(const HostEngine& E, const Z8Image& Z, const SegImage& G);
const HostEngine& dev_host(const DevEngine& D);
const Z8Image& dev_z8(const DevEngine& D);
const SegImage& dev_seg(const DevEngine& D);

// Plain form: a straightforward FP64 port of engine.rs:669-722 (sum_subspace +
// gray), one query per launch, Gray chunks of 2^L branches per thread.
// Returns the raw FP64 sum over the affine subspace (before the boundary phase
// and the 2^{log2mag - h/2} scale).  kernel_ms receives the walk kernel time.
std::complex<double> dev_plain_sum(const DevEngine& D, const Prepared& P, double* kernel_ms);
// Monte-Carlo sums of Engine::estimate (engine.rs:612-659) over `samples`
// uniform points of the branch subspace: sum of values (re, im) and of |x|^2.
struct McSums { double re = 0, im = 0, sq = 0; uint64_t n = 0; };
McSums dev_plain_estimate(const DevEngine& D, const Prepared& P, uint64_t samples, uint64_t seed);

// Batched Z8 walk: one persistent launch over all (query, lane-block) items.
struct Z8Job {
    const DevEngine* eng;
    Z8QueryHost q;                // branch walk query data
    SegQueryHost sq;              // segment walk query data
    std::complex<double> scale;   // applied in FP64 after the cross-item reduction
    uint32_t mid_bits;            // outer Gray bits per lane (autotuned)
};
struct Z8Timing {
    double kernel_ms = 0;         // walk kernel, HIP events on the launch stream (NaN when not timed)
    double total_ms = 0;          // copies + launches + reduction, events (NaN when not timed)
    uint64_t branches = 0;        // sum over jobs of 2^d (real coordinates)
    uint64_t items = 0;
    uint32_t grid = 0;
};
// Results are scale * (EV + omega * OD) per job, FP64.
// Eager batches record HIP timing events only while this is on (default off).
void dev_set_timing(bool on);
bool dev_get_timing();
void dev_z8_batch(const std::vector<Z8Job>& jobs, std::vector<std::complex<double>>& out, Z8Timing* timing, int walk);
// The same batch with every device buffer (staging, queue counters, item
// partials, results) in a fixed file-scope device array of
// dev_noalloc_scratch_bytes(): allocates no device or pinned host memory.
// Throws if the batch needs more.  Results are bitwise those of dev_z8_batch.
void dev_z8_batch_noalloc(const std::vector<Z8Job>& jobs, std::vector<std::complex<double>>& out, int walk);
size_t dev_noalloc_scratch_bytes();
// Asynchronous form: a slot holds one batch in flight on its own stream.
struct Z8Slot;
std::shared_ptr<Z8Slot> dev_slot_create();
void dev_z8_launch(Z8Slot& S, const std::vector<Z8Job>& jobs, int walk, bool timed);
bool dev_z8_ready(Z8Slot& S);
void dev_z8_wait(Z8Slot& S, std::vector<std::complex<double>>& out, Z8Timing* timing);
// Mid-bit heuristic: enough items to fill every CU several times, long
// enough walks to amortise the lane seed.
uint32_t z8_auto_mid_bits(uint32_t d, uint32_t kin);   // mid coordinates of a lane walk (kin inner slots)
void z8_set_tuning(uint32_t offset, uint32_t max_mid);
void z8_get_tuning(uint32_t* offset, uint32_t* max_mid);
void seg_set_tuning(uint32_t offset, uint32_t max_mid);
void seg_get_tuning(uint32_t* offset, uint32_t* max_mid);
void seg_set_waves(uint32_t waves);     // segment walk workgroup shape: 1, 2 or 4 waves
uint32_t seg_get_waves();
// Lane walk length of an engine's queries for the given form.
uint32_t dev_mid_bits(const DevEngine& D, int walk);

// Replayable batch: the whole batch (descriptor upload, counter clear, walk,
// reduction, result download) captured once into a hipGraph and relaunched.
struct Z8Graph;
std::shared_ptr<Z8Graph> dev_z8_graph(const std::vector<Z8Job>& jobs, int walk);
void dev_z8_graph_run(Z8Graph& G, std::vector<std::complex<double>>& out, Z8Timing* timing);

// Per-branch dumps for parity tests.
// (a) Gauss-sum outputs for explicit t-bit branch vectors s via the packed
//     popcount form (normal_form.rs:472-493): zero flag, k8, and the total
//     phase k8 + 6|s| + 4|neg \ s| mod 8.
void dev_dump_s(const DevEngine& D, const Prepared& P, const uint64_t* s, size_t n, uint8_t* zero, uint8_t* k8, uint8_t* ktot);
// (b) Every branch visited by the Z8 walk for one query with the given mid
//     bits: the integer phase k(c) and the magnitude prod r, indexed by
//     ((hi << M) + o) * 256 + i.  Phantom entries get k = 255.
void dev_dump_walk(const DevEngine& D, const Z8QueryHost& q, uint32_t mid_bits, uint8_t* k, float* mag);
// (c) Integer state (K mod 8, table-basis sign mask, inner pivots) at every
//     outer point of the segment walk, indexed by (lane << M) + o.
void dev_dump_seg(const DevEngine& D, const SegQueryHost& q, uint32_t mid_bits, uint8_t* k, uint32_t* hs, uint32_t* pio);

// Device / runtime helpers
int dev_count();
std::string dev_name();
void dev_sync();

}  // namespace pq
