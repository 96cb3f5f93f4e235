// VALU issue-rate probe for gfx950: long independent chains of one vector
// instruction per kernel (int32 xor / and / add / popcount, FP32 FMA, FP64 FMA,
// the branch walk's own mix of one int32 xor per FP64 FMA, and the FP64 FMA
// with one operand from an SGPR pair or broadcast across each 16-lane row by
// DPP, the two ways a uniform table entry can reach it).  Each lane runs
// 8 independent chains, so the rate measured is issue throughput, not latency.
// Two more kernels measure the FP64 matrix core for comparison: 16x16x4 FP64
// MFMAs alone, and MFMAs with VALU FP64 FMAs beside them (rates in lane-FMAs:
// one such MFMA is 16 FMAs per lane).
// The kernel's clock is read in-kernel (s_memtime against the constant-rate
// s_memrealtime) so that rates can be quoted in lanes per clock per CU.
//
// The instructions are written as inline asm so the compiler can neither fold
// nor vectorise the chains; none of them touches memory.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>
#include <stdexcept>

namespace pq {
extern thread_local std::string g_err;     // pq_last_error() (pq_api.cpp)

namespace {
enum ProbeOp { P_XOR = 0, P_AND, P_ADD, P_BCNT, P_FMA32, P_FMA64, P_XOR_FMA64, P_FMAC64_S, P_FMAC64_DPP, P_MFMA64, P_MFMA64_FMA64, P_COUNT };
const char* const kProbeNames[P_COUNT] = {"v_xor_b32", "v_and_b32", "v_add_u32", "v_bcnt_u32_b32", "v_fma_f32", "v_fma_f64",
                                          "v_xor_b32+v_fma_f64", "v_fmac_f64 (SGPR operand)", "v_fmac_f64_dpp row_newbcast",
                                          "v_mfma_f64_16x16x4f64", "v_mfma_f64_16x16x4f64+2 v_fma_f64"};
// FMA lanes per lane and iteration of the MFMA kernels: 8 MFMAs of 16 lane-FMAs,
// plus 16 VALU FMAs in the mix.
constexpr double kMfmaLaneOps[2] = {8.0 * 16.0, 8.0 * 16.0 + 16.0};
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int OP>
__device__ __forceinline__ void probe_step(uint32_t (&a)[8], double (&d)[8], uint32_t b, float fb, double db, double sdb) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if constexpr (OP == P_XOR) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
        else if constexpr (OP == P_AND) asm volatile("v_and_b32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
        else if constexpr (OP == P_ADD) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
        else if constexpr (OP == P_BCNT) asm volatile("v_bcnt_u32_b32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
        else if constexpr (OP == P_FMA32) {
            float f = __uint_as_float(a[k]);
            asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(f) : "v"(fb), "v"(0.0f));
            a[k] = __float_as_uint(f);
        } else if constexpr (OP == P_FMA64) asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(d[k]) : "v"(db), "v"(0.0));
        else if constexpr (OP == P_FMAC64_S) asm volatile("v_fmac_f64 %0, %1, %2" : "+v"(d[k]) : "s"(sdb), "v"(db));
        else if constexpr (OP == P_FMAC64_DPP)
            asm volatile("v_fmac_f64_dpp %0, %1, %2 row_newbcast:5 row_mask:0xf bank_mask:0xf" : "+v"(d[k]) : "v"(db), "v"(db));
        else {
            if (k & 1) asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(d[k]) : "v"(db), "v"(0.0));
            else asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
        }
    }
}

// iters x 8 steps x 8 chains = 64 instructions of the op per lane and iteration.
template <int OP>
__global__ __launch_bounds__(256) void k_valu(uint32_t iters, uint32_t seed, uint32_t* __restrict__ sink, uint64_t* __restrict__ clk) {
    const uint32_t tid = blockIdx.x * 256 + threadIdx.x;
    uint32_t a[8];
    double d[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        // FP chains start at small normal values (x * 1 + 0 keeps them there)
        a[k] = OP == P_FMA32 ? __float_as_uint(1.0f + k) : seed ^ (tid * 0x9E3779B9u + k);
        d[k] = 1.0 + k;
    }
    const uint32_t b = seed * 3u + 1u;
    const float fb = 1.0f;
    const double db = 1.0;
    const double sdb = __builtin_amdgcn_readfirstlane(seed) == 7u ? 1.0 : 0.5;   // uniform: an SGPR pair
    uint64_t c0 = 0, w0 = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) { c0 = clock64(); w0 = wall_clock64(); }
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < 8; ++r) probe_step<OP>(a, d, b, fb, db, sdb);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clk[0] = clock64() - c0;
        clk[1] = wall_clock64() - w0;
    }
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) x ^= a[k] ^ (uint32_t)__double2loint(d[k]);
    sink[tid] = x;
}

// iters x (8 FP64 MFMAs on independent accumulators, and with MIX 2 VALU
// FP64 FMAs after each).
template <bool MIX>
__global__ __launch_bounds__(256) void k_mfma(uint32_t iters, uint32_t seed, uint32_t* __restrict__ sink, uint64_t* __restrict__ clk) {
    const uint32_t tid = blockIdx.x * 256 + threadIdx.x;
    f64x4 acc[8];
    double d[16];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = f64x4{1.0 + k, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 16; ++k) d[k] = 1.0 + k;
    const double a = seed == 7u ? 1.0 : 0.5, b = 1.0;
    uint64_t c0 = 0, w0 = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) { c0 = clock64(); w0 = wall_clock64(); }
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[k], 0, 0, 0);
            if constexpr (MIX) {
                asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(d[2 * k]) : "v"(b), "v"(0.0));
                asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(d[2 * k + 1]) : "v"(b), "v"(0.0));
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        clk[0] = clock64() - c0;
        clk[1] = wall_clock64() - w0;
    }
    double x = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) x += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
#pragma unroll
    for (int k = 0; k < 16; ++k) x += d[k];
    sink[tid] = (uint32_t)__double2loint(x);
}

#define VCHECK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
    } while (0)

template <int OP>
void run_probe(uint32_t grid, uint32_t iters, uint32_t* sink, uint64_t* clk, hipEvent_t e0, hipEvent_t e1, hipStream_t st,
               double& ms, uint64_t (&hclk)[2]) {
    // one untimed launch to warm the clock, then the timed one
    auto kern = [] {
        if constexpr (OP == P_MFMA64) return k_mfma<false>;
        else if constexpr (OP == P_MFMA64_FMA64) return k_mfma<true>;
        else return k_valu<OP>;
    }();
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, iters / 4 + 1, 7u, sink, clk);
    VCHECK(hipGetLastError());
    VCHECK(hipEventRecord(e0, st));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, iters, 7u, sink, clk);
    VCHECK(hipGetLastError());
    VCHECK(hipEventRecord(e1, st));
    VCHECK(hipEventSynchronize(e1));
    float f = 0;
    VCHECK(hipEventElapsedTime(&f, e0, e1));
    ms = f;
    VCHECK(hipMemcpy(hclk, clk, 16, hipMemcpyDeviceToHost));
}
}  // namespace

// Runs every probe op; out (3 + 4 * P_COUNT doubles): CUs, wall-clock rate
// (MHz), blocks per CU, then per op {lane-ops/s, lanes per clock per CU at the
// in-kernel clock, in-kernel clock GHz, kernel ms}.
int valu_probe(double* out, size_t cap, uint32_t blocks_per_cu, uint32_t iters) {
    if (cap < 3 + 4 * (size_t)P_COUNT) throw std::invalid_argument("output too small");
    int dev = 0;
    VCHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    VCHECK(hipGetDeviceProperties(&prop, dev));
    int wall_khz = 0;
    VCHECK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev));
    const uint32_t cus = (uint32_t)prop.multiProcessorCount;
    const uint32_t grid = cus * blocks_per_cu;
    uint32_t* sink = nullptr;
    uint64_t* clk = nullptr;
    hipStream_t st;
    hipEvent_t e0, e1;
    VCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    VCHECK(hipEventCreate(&e0));
    VCHECK(hipEventCreate(&e1));
    VCHECK(hipMalloc(&sink, (size_t)grid * 256 * 4));
    VCHECK(hipMalloc(&clk, 16));
    out[0] = cus; out[1] = wall_khz / 1e3; out[2] = blocks_per_cu;
    // an FP64 MFMA issues for about as long as 16 VALU FP64 FMAs
    const uint32_t miters = std::max<uint32_t>(1, iters / 16);
    for (int op = 0; op < P_COUNT; ++op) {
        double ms = 0;
        uint64_t hc[2] = {0, 0};
        switch (op) {
            case P_XOR: run_probe<P_XOR>(grid, iters, sink, clk, e0, e1, st, ms, hc); break;
            case P_AND: run_probe<P_AND>(grid, iters, sink, clk, e0, e1, st, ms, hc); break;
            case P_ADD: run_probe<P_ADD>(grid, iters, sink, clk, e0, e1, st, ms, hc); break;
            case P_BCNT: run_probe<P_BCNT>(grid, iters, sink, clk, e0, e1, st, ms, hc); break;
            case P_FMA32: run_probe<P_FMA32>(grid, iters, sink, clk, e0, e1, st, ms, hc); break;
            case P_FMA64: run_probe<P_FMA64>(grid, iters, sink, clk, e0, e1, st, ms, hc); break;
            case P_FMAC64_S: run_probe<P_FMAC64_S>(grid, iters, sink, clk, e0, e1, st, ms, hc); break;
            case P_FMAC64_DPP: run_probe<P_FMAC64_DPP>(grid, iters, sink, clk, e0, e1, st, ms, hc); break;
            case P_MFMA64: run_probe<P_MFMA64>(grid, miters, sink, clk, e0, e1, st, ms, hc); break;
            case P_MFMA64_FMA64: run_probe<P_MFMA64_FMA64>(grid, miters, sink, clk, e0, e1, st, ms, hc); break;
            default: run_probe<P_XOR_FMA64>(grid, iters, sink, clk, e0, e1, st, ms, hc); break;
        }
        const bool mfma = op == P_MFMA64 || op == P_MFMA64_FMA64;
        const double lane_ops = (double)grid * 256.0 * (mfma ? (double)miters * kMfmaLaneOps[op - P_MFMA64] : (double)iters * 64.0);
        const double rate = lane_ops / (ms * 1e-3);
        const double ghz = hc[1] ? (double)hc[0] / ((double)hc[1] / (wall_khz * 1e3)) / 1e9 : 0.0;
        double* o = out + 3 + 4 * op;
        o[0] = rate;
        o[1] = ghz > 0 ? rate / (cus * ghz * 1e9) : 0.0;
        o[2] = ghz;
        o[3] = ms;
    }
    (void)hipFree(sink);
    (void)hipFree(clk);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    return P_COUNT;
}

const char* valu_probe_name(int op) { return op >= 0 && op < P_COUNT ? kProbeNames[op] : ""; }

}  // namespace pq

extern "C" {
int pq_valu_probe(double* out, size_t cap, uint32_t blocks_per_cu, uint32_t iters) {
    try {
        if (blocks_per_cu == 0 || blocks_per_cu > 8) throw std::invalid_argument("blocks_per_cu must be 1..8");
        if (iters == 0 || iters > (1u << 24)) throw std::invalid_argument("iters must be 1..2^24");
        return pq::valu_probe(out, cap, blocks_per_cu, iters);
    } catch (std::exception& e) {
        pq::g_err = e.what();
        return -1;
    }
}
const char* pq_valu_probe_name(int op) { return pq::valu_probe_name(op); }
}
