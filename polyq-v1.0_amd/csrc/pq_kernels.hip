// HIP kernels for gfx950 (MI355X): the branch evaluator of the polyq path-sum
// engine.
//
//  k_plain<NW>    straightforward FP64 port of Engine::gray / sum_subspace
//                 (src/engine.rs:669-722): one thread per Gray chunk, packed
//                 shift state (a, b, o words) in VGPRs, phase8 by popcounts
//                 (normal_form.rs:479-493), FP64 complex weight updated with
//                 the up/down multipliers and recomputed every 2048 steps.
//  k_z8_walk      persistent, work-queue kernel over (query, 256-lane block)
//                 items of many queries and engines.  A lane owns one coset of
//                 the outer coordinates, walks the mid coordinates in Gray
//                 order by single-slot updates of its integer state, and sums
//                 the 2^kin inner branches of each outer point from a
//                 uniform FP64 table held in VGPRs (DPP row broadcasts): per
//                 branch one FP64 FMA, the sign flipped once per 16 branches
//                 (an LDS pivot-factor lookup per branch when the low inner
//                 directions touch pivots).  Every branch is visited;
//                 partials accumulate in FP64, split by the parity of the
//                 lane phase.
//  k_z8_reduce    FP64 cross-workgroup reduction and global scaling.
//  k_dump_s, k_dump_walk   per-branch outputs for parity tests.
//
// Uniform data (structure rows, tables indexed by the loop counter) is read
// through the scalar cache; lane-indexed tables (lane-factor byte tables, the
// inner pivot table) are staged in LDS.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include "pq_device.hpp"
#include <mutex>
#include <chrono>
#include <thread>
#include <functional>
#include <atomic>
#include <cstring>
#include <algorithm>
#include <numeric>
#include <unordered_map>
#include <cmath>

namespace pq {

#define HIPCHECK(x)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_) + " (" #x ")"); \
    } while (0)

// ----------------------------------------------------------------- runtime --
namespace {
struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    void* get(size_t n) {
        if (n > cap) {
            if (p) HIPCHECK(hipFree(p));
            size_t c = std::max<size_t>(n, cap * 2);
            c = std::max<size_t>(c, 4096);
            HIPCHECK(hipMalloc(&p, c));
            cap = c;
        }
        return p;
    }
};
struct HBuf {
    void* p = nullptr;
    void* dev = nullptr;     // the device's address of p
    size_t cap = 0;
    void* get(size_t n) {
        if (n > cap) {
            if (p) HIPCHECK(hipHostFree(p));
            p = dev = nullptr;
            size_t c = std::max<size_t>(n, cap * 2);
            c = std::max<size_t>(c, 4096);
            HIPCHECK(hipHostMalloc(&p, c, hipHostMallocMapped));   // pinned, device-addressable
            HIPCHECK(hipHostGetDevicePointer(&dev, p, 0));
            cap = c;
        }
        return p;
    }
};
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0, e1, e2, e3;
    int cus = 0;
    std::mutex mu;
    DBuf d_engs, d_qs, d_counter, d_partial, d_out, d_misc, d_misc2, d_misc3;
    HBuf h_in, h_out;
};
// The context is created by the first call that finds a device; a call that
// fails (no device) throws and leaves the next call to try again.  A plain
// mutex rather than std::call_once: an exception leaving call_once's callable
// is not reset by every pthread_once (ThreadSanitizer's interceptor leaves the
// flag taken, and every later caller waits for ever).
Ctx& ctx() {
    static std::atomic<Ctx*> made{nullptr};
    static std::mutex init_mu;
    if (Ctx* c = made.load(std::memory_order_acquire)) return *c;
    std::lock_guard<std::mutex> lk(init_mu);
    if (Ctx* c = made.load(std::memory_order_relaxed)) return *c;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw std::runtime_error("no HIP device available");
    std::unique_ptr<Ctx> c(new Ctx());
    HIPCHECK(hipGetDevice(&c->device));
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&c->e0)); HIPCHECK(hipEventCreate(&c->e1));
    HIPCHECK(hipEventCreate(&c->e2)); HIPCHECK(hipEventCreate(&c->e3));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, c->device));
    c->cus = prop.multiProcessorCount;
    made.store(c.get(), std::memory_order_release);
    return *c.release();
}
}  // namespace

int dev_count() { int n = 0; if (hipGetDeviceCount(&n) != hipSuccess) return 0; return n; }
std::string dev_name() {
    Ctx& C = ctx();
    hipDeviceProp_t p; HIPCHECK(hipGetDeviceProperties(&p, C.device));
    return std::string(p.name) + " / " + p.gcnArchName + " / " + std::to_string(p.multiProcessorCount) + " CUs";
}
void dev_sync() { HIPCHECK(hipStreamSynchronize(ctx().stream)); }

// ------------------------------------------------------------ device images --
struct PlainDev {
    const uint32_t* ku;      // d rows x 3NW: [a | b | o]
    const uint64_t* kn;      // d kernel vectors over t terms
    const uint32_t* soff;    // d+1 support offsets
    const uint8_t* supp;     // kernel supports
    const double2* a0; const double2* a1; const double2* up; const double2* down;
    const uint32_t* diag;    // 4NW: wa | wb | o1 | o3
    const uint32_t* mrow;    // t rows x (3NW + NWe): [a | b | o | e]
    const uint32_t* we;      // NWe
    uint32_t d, t, NW, NWe;
    int32_t sigma;
    uint32_t rbits;          // r = 2g + 2#even + #odd
    uint64_t neg;
};
struct Z8Dev {
    const uint64_t* slot_q;
    const uint64_t* slot_n;
    const uint32_t* slot_pin;
    const uint32_t* slot_lam;
    const double* inner;     // 2^kin signed FP64 values W[i]
    const double* inner_acc; // the same in accumulator-walk order
    const uint32_t* inaux;   // 2^kin x pin*8, then 2^kin x phase
    const double* lt;        // nbytes x 256
    const double* gin;       // 2^npin
    const double* lt_rest;   // nbytes x 256, without the terms of the pivot-free outer slots kin .. kin + nx - 1
    const double* mid_r;     // nx x {r0, r1}: magnitudes of those slots' terms
    const uint32_t* mid_term;  // nx: their term indices
    uint32_t nslots, nbytes, npin, imag0, nx, lowfree;   // lowfree: the four low inner slots touch no pivot
    uint32_t kin, pad0;      // inner slots (kInnerWide, or kInnerPiv for the variants with pivot lookups)
};
struct Z8Q {                 // one query of a batch (64 B)
    uint64_t delta, s0, item0, nitems;
    uint32_t eng, K0, pio0, M, H, pad0, pad1, pad2;
};
// Segment-table image (see pq_seg.cpp): outer slot data, the FP32 hi/lo
// Walsh-Hadamard table and the magnitude tables.
struct SegDev {
    const uint64_t* oq;      // nout: Q rows over outer slots
    const uint32_t* ox;      // nout: cross masks over inner slots (table basis)
    const uint32_t* opin;    // nout: compact inner pivots
    const uint32_t* olam;    // nout: lambda
    const uint64_t* on;      // nout: kernel vectors over terms
    const float4* table;     // 2^(I+npin) entries, index (pio << I) | hs
    const float2* lt;        // nbytes x 256
    const float* fmid;       // walk magnitudes {hi, lo}, M = 0..Mcap at 2(2^M - 1)
    const float* rmid;       // Mcap x {r0 hi, r0 lo, r1 hi, r1 lo}
    const uint32_t* wtab;    // 2^Mcap: sign-mask change | engine phase change << 24
    const uint32_t* pts;     // walk points sorted by phase class mod 4, per M at 4(2^M - 1)
    const uint32_t* pcls;    // class boundaries, 5 per M
    const uint32_t* lowtab;  // 2^L6 x 6 words: seed terms of the lane-low slots
    uint32_t I, npin, nout, L6, Mcap, nbytes, pad0, pad1;
};
struct SegQ {                // one query of a segment batch (64 B; item0 / nitems where Z8Q has them)
    uint64_t delta, s0, item0, nitems;
    uint32_t eng, K0, hs0, pio0, M, H, pad0, pad1;
};
static_assert(sizeof(Z8Q) == 64 && sizeof(SegQ) == 64, "query records are 64 bytes");

struct DevEngine {
    std::shared_ptr<const HostEngine> host;
    std::shared_ptr<const Z8Image> z8;
    std::shared_ptr<const SegImage> seg;
    void* blob = nullptr;
    size_t bytes = 0;
    PlainDev plain{};
    Z8Dev z{};
    SegDev sg{};
    ~DevEngine() { if (blob) (void)hipFree(blob); }
};

size_t dev_engine_bytes(const DevEngine& D) { return D.bytes; }
const HostEngine& dev_host(const DevEngine& D) { return *D.host; }
const Z8Image& dev_z8(const DevEngine& D) { return *D.z8; }
const SegImage& dev_seg(const DevEngine& D) { return *D.seg; }

namespace {
// The blob is assembled in a per-thread buffer that the next blob on the
// thread reuses (a blob holds the segment tables, hundreds of KB: a fresh
// vector of that size is mmap'd and page-faulted on every build).  At most
// one builder is alive per thread at a time.
std::vector<uint8_t>& blob_scratch() {
    thread_local std::vector<uint8_t> b;
    return b;
}
struct BlobBuilder {
    std::vector<uint8_t>& bytes;
    BlobBuilder() : bytes(blob_scratch()) { bytes.clear(); }
    BlobBuilder(const BlobBuilder&) = delete;
    size_t put(const void* p, size_t n) {
        size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + std::max<size_t>(n, 16), 0);
        if (n) std::memcpy(bytes.data() + off, p, n);
        return off;
    }
};
// u64 packed words -> u32 dwords padded to nw dwords
void put_u32(std::vector<uint32_t>& dst, const uint64_t* src, uint32_t nwords64, uint32_t nbits, uint32_t nw) {
    for (uint32_t k = 0; k < nw; ++k) {
        uint32_t v = 0;
        if (k / 2 < nwords64 && 32 * k < std::max(nbits, 1u)) v = (uint32_t)(src[k / 2] >> (32 * (k & 1)));
        dst.push_back(v);
    }
}
uint32_t pick_nw(uint32_t bits) {
    uint32_t need = (bits + 31) / 32;
    for (uint32_t nw : {2u, 4u, 8u, 16u, 32u}) if (need <= nw) return nw;
    throw std::invalid_argument("normal form too large for the plain kernel (more than 1024 planes or lines)");
}
}  // namespace

namespace {
struct BlobLayout {
    size_t o_ku, o_kn, o_soff, o_supp, o_a0, o_a1, o_up, o_dn, o_diag, o_mrow, o_we, o_sq, o_sn, o_sp, o_sl, o_in, o_inacc, o_inaux, o_lt, o_gin;
    size_t o_ltr, o_mr, o_mt;
    size_t g_oq, g_ox, g_opin, g_olam, g_on, g_tab, g_lt, g_fmid, g_rmid, g_wtab, g_pts, g_pcls, g_low;
    uint32_t NW, NWe;
};
// Host image of one engine: the plain FP64 part and the Z8 walk part in one blob.
BlobLayout build_blob(const HostEngine& E, const Z8Image& Z, const SegImage& G, BlobBuilder& B) {
    const ULayout& L = E.L;
    BlobLayout O;
    // plain image: only when the normal form fits the plain kernel's packed
    // rows (at most 1024 planes or lines); NW = 0 marks an engine without it
    const bool has_plain = std::max(L.g, L.nodd) <= 1024;
    const uint32_t NW = has_plain ? pick_nw(std::max(L.g, L.nodd)) : 0;
    const uint32_t NWe = std::max<uint32_t>(2, (L.neven + 31) / 32 + ((L.neven + 31) / 32) % 2);
    const uint32_t W = L.words();
    std::vector<uint32_t> ku, diag, mrow, we;
    std::vector<uint32_t> soff{0};
    std::vector<uint8_t> supp;
    std::vector<double> a0, a1, up, dn;
    if (has_plain) {
        for (uint32_t j = 0; j < E.d; ++j) {
            const uint64_t* r = (E.kernel_u.data() + (size_t)j * W);
            put_u32(ku, r + L.off_a(), L.Wg, L.g, NW); put_u32(ku, r + L.off_b(), L.Wg, L.g, NW); put_u32(ku, r + L.off_o(), L.Wo, L.nodd, NW);
        }
        put_u32(diag, E.diag.data() + L.off_a(), L.Wg, L.g, NW); put_u32(diag, E.diag.data() + L.off_b(), L.Wg, L.g, NW);
        put_u32(diag, E.diag.data() + L.off_o(), L.Wo, L.nodd, NW); put_u32(diag, E.o3.data(), L.Wo, L.nodd, NW);
        for (uint32_t k = 0; k < E.tdev; ++k) {
            const uint64_t* r = (E.mcols.data() + (size_t)k * W);
            put_u32(mrow, r + L.off_a(), L.Wg, L.g, NW); put_u32(mrow, r + L.off_b(), L.Wg, L.g, NW);
            put_u32(mrow, r + L.off_o(), L.Wo, L.nodd, NW); put_u32(mrow, r + L.off_e(), L.We, L.neven, NWe);
        }
        put_u32(we, E.diag.data() + L.off_e(), L.We, L.neven, NWe);
        for (uint32_t j = 0; j < E.d; ++j) {
            for (uint32_t k = 0; k < E.tdev; ++k) if ((E.kernel[j] >> k) & 1) supp.push_back((uint8_t)k);
            soff.push_back((uint32_t)supp.size());
        }
        for (uint32_t k = 0; k < E.tdev; ++k) {
            a0.push_back(E.a0[k].real()); a0.push_back(E.a0[k].imag());
            a1.push_back(E.a1[k].real()); a1.push_back(E.a1[k].imag());
            up.push_back(E.up[k].real()); up.push_back(E.up[k].imag());
            dn.push_back(E.down[k].real()); dn.push_back(E.down[k].imag());
        }
    }
    O.NW = NW; O.NWe = NWe;
    O.o_ku = B.put(ku.data(), ku.size() * 4); O.o_kn = B.put(E.kernel.data(), E.kernel.size() * 8);
    O.o_soff = B.put(soff.data(), soff.size() * 4); O.o_supp = B.put(supp.data(), supp.size());
    O.o_a0 = B.put(a0.data(), a0.size() * 8); O.o_a1 = B.put(a1.data(), a1.size() * 8);
    O.o_up = B.put(up.data(), up.size() * 8); O.o_dn = B.put(dn.data(), dn.size() * 8);
    O.o_diag = B.put(diag.data(), diag.size() * 4); O.o_mrow = B.put(mrow.data(), mrow.size() * 4); O.o_we = B.put(we.data(), we.size() * 4);
    // Z8 image
    std::vector<uint32_t> lam32(Z.slot_lam.begin(), Z.slot_lam.end());
    O.o_sq = B.put(Z.slot_q.data(), Z.slot_q.size() * 8); O.o_sn = B.put(Z.slot_n.data(), Z.slot_n.size() * 8);
    O.o_sp = B.put(Z.slot_pin.data(), Z.slot_pin.size() * 4); O.o_sl = B.put(lam32.data(), lam32.size() * 4);
    O.o_in = B.put(Z.inner.data(), Z.inner.size() * 8); O.o_inacc = B.put(Z.inner_acc.data(), Z.inner_acc.size() * 8);
    O.o_inaux = B.put(Z.inner_aux.data(), Z.inner_aux.size() * 4);
    O.o_lt = B.put(Z.lt.data(), Z.lt.size() * 8); O.o_gin = B.put(Z.gin.data(), Z.gin.size() * 8);
    O.o_ltr = B.put(Z.lt_rest.data(), Z.lt_rest.size() * 8); O.o_mr = B.put(Z.mid_r.data(), Z.mid_r.size() * 8);
    O.o_mt = B.put(Z.mid_term.data(), Z.mid_term.size() * 4);
    // segment-table image
    O.g_oq = B.put(G.oq.data(), G.oq.size() * 8); O.g_ox = B.put(G.ox.data(), G.ox.size() * 4);
    O.g_opin = B.put(G.opin.data(), G.opin.size() * 4); O.g_olam = B.put(G.olam.data(), G.olam.size() * 4);
    O.g_on = B.put(G.on.data(), G.on.size() * 8); O.g_tab = B.put(G.table.data(), G.table.size() * 4);
    O.g_lt = B.put(G.lt.data(), G.lt.size() * 4); O.g_fmid = B.put(G.fmid.data(), G.fmid.size() * 4);
    O.g_rmid = B.put(G.rmid.data(), G.rmid.size() * 4);
    O.g_wtab = B.put(G.wtab.data(), G.wtab.size() * 4);
    O.g_low = B.put(G.lowtab.data(), G.lowtab.size() * 4);
    O.g_pts = B.put(G.pts.data(), G.pts.size() * 4);
    O.g_pcls = B.put(G.pcls.data(), G.pcls.size() * 4);
    return O;
}
}  // namespace

// Streams for engine uploads (default priority), each used by one thread at
// a time.  Made once, on the first upload (the context exists by then); a
// failed creation is retried by the next upload.
struct UploadStreams {
    static constexpr size_t N = 8;
    hipStream_t s[N] = {};
    std::mutex mu[N];
};
UploadStreams& upload_streams() {
    static UploadStreams* made = [] {
        std::unique_ptr<UploadStreams> u(new UploadStreams());
        for (auto& x : u->s) HIPCHECK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        return u.release();
    }();
    return *made;
}

size_t dev_engine_image_bytes(const HostEngine& E, const Z8Image& Z, const SegImage& G) {
    BlobBuilder B;
    build_blob(E, Z, G, B);
    return B.bytes.size();
}

std::shared_ptr<DevEngine> dev_engine_upload(std::shared_ptr<const HostEngine> Ep, std::shared_ptr<const Z8Image> Zp,
                                             std::shared_ptr<const SegImage> Gp) {
    (void)ctx();                 // the device context exists before the first upload
    const HostEngine& E = *Ep;
    const Z8Image& Z = *Zp;
    const SegImage& G = *Gp;
    const ULayout& L = E.L;
    auto D = std::make_shared<DevEngine>();
    D->host = Ep; D->z8 = Zp; D->seg = Gp;
    BlobBuilder B;
    const BlobLayout O = build_blob(E, Z, G, B);
    const uint32_t NW = O.NW, NWe = O.NWe;

    // the copy goes on an upload stream of its own: host workers uploading
    // different engines neither wait for each other nor for the context's
    // stream
    HIPCHECK(hipMalloc(&D->blob, B.bytes.size()));
    D->bytes = B.bytes.size();
    {
        UploadStreams& U = upload_streams();
        const size_t k = std::hash<std::thread::id>{}(std::this_thread::get_id()) % UploadStreams::N;
        std::lock_guard<std::mutex> lk(U.mu[k]);
        HIPCHECK(hipMemcpyAsync(D->blob, B.bytes.data(), B.bytes.size(), hipMemcpyHostToDevice, U.s[k]));
        HIPCHECK(hipStreamSynchronize(U.s[k]));
    }
    auto at = [&](size_t off) { return (const uint8_t*)D->blob + off; };
    PlainDev& P = D->plain;
    P.ku = (const uint32_t*)at(O.o_ku); P.kn = (const uint64_t*)at(O.o_kn); P.soff = (const uint32_t*)at(O.o_soff); P.supp = at(O.o_supp);
    P.a0 = (const double2*)at(O.o_a0); P.a1 = (const double2*)at(O.o_a1); P.up = (const double2*)at(O.o_up); P.down = (const double2*)at(O.o_dn);
    P.diag = (const uint32_t*)at(O.o_diag); P.mrow = (const uint32_t*)at(O.o_mrow); P.we = (const uint32_t*)at(O.o_we);
    P.d = E.d; P.t = E.tdev; P.NW = NW; P.NWe = NWe; P.sigma = E.sigma;
    P.rbits = 2 * L.g + 2 * L.neven + L.nodd; P.neg = E.neg_mask;
    Z8Dev& z = D->z;
    z.slot_q = (const uint64_t*)at(O.o_sq); z.slot_n = (const uint64_t*)at(O.o_sn); z.slot_pin = (const uint32_t*)at(O.o_sp); z.slot_lam = (const uint32_t*)at(O.o_sl);
    z.inner = (const double*)at(O.o_in); z.inner_acc = (const double*)at(O.o_inacc); z.inaux = (const uint32_t*)at(O.o_inaux);
    z.lt = (const double*)at(O.o_lt); z.gin = (const double*)at(O.o_gin);
    z.lt_rest = (const double*)at(O.o_ltr); z.mid_r = (const double*)at(O.o_mr); z.mid_term = (const uint32_t*)at(O.o_mt);
    z.nslots = Z.nslots; z.nbytes = Z.nbytes; z.npin = Z.npin; z.imag0 = Z.imag0 ? 1u : 0u; z.nx = Z.nx;
    z.lowfree = Z.low_pivot_free ? 1u : 0u;
    z.kin = Z.kin;
    SegDev& g = D->sg;
    g.oq = (const uint64_t*)at(O.g_oq); g.ox = (const uint32_t*)at(O.g_ox); g.opin = (const uint32_t*)at(O.g_opin);
    g.olam = (const uint32_t*)at(O.g_olam); g.on = (const uint64_t*)at(O.g_on); g.table = (const float4*)at(O.g_tab);
    g.lt = (const float2*)at(O.g_lt); g.fmid = (const float*)at(O.g_fmid); g.rmid = (const float*)at(O.g_rmid);
    g.wtab = (const uint32_t*)at(O.g_wtab);
    g.lowtab = (const uint32_t*)at(O.g_low);
    g.pts = (const uint32_t*)at(O.g_pts); g.pcls = (const uint32_t*)at(O.g_pcls);
    g.I = G.I; g.npin = G.npin; g.nout = G.nout; g.L6 = G.L6; g.Mcap = G.Mcap; g.nbytes = G.nbytes;
    if (G.nbytes > 8) throw std::invalid_argument("more than 64 terms");
    return D;
}

// ============================================================= plain FP64 ===
struct PlainQ {
    uint32_t ustart[96];     // [a | b | o] padded to NW dwords each
    uint64_t s0;
    double cre, cim;
    double fre, fim;         // product of the term multipliers of the terms beyond the device terms (1 when t <= 64)
    uint32_t L;
};

__device__ __forceinline__ void cmul(double& r, double& i, double mr, double mi) {
    const double nr = r * mr - i * mi;
    const double ni = r * mi + i * mr;
    r = nr; i = ni;
}
__device__ __forceinline__ void plain_weight(const PlainDev& E, const PlainQ& Q, uint64_t sv, double& cr, double& ci) {
    cr = Q.fre; ci = Q.fim;
    for (uint32_t k = 0; k < E.t; ++k) {
        const double2 a = ((sv >> k) & 1) ? E.a1[k] : E.a0[k];
        cmul(cr, ci, a.x, a.y);
    }
}

template <int NW>
__global__ __launch_bounds__(256) void k_plain(PlainDev E, PlainQ Q, double* __restrict__ partial) {
    __shared__ double r8[16];
    __shared__ double red[8];
    const uint32_t tid = threadIdx.x;
    if (tid < 8) {
        const double h = 0.70710678118654752440;
        const double re[8] = {1, h, 0, -h, -1, -h, 0, h}, im[8] = {0, h, 1, h, 0, -h, -1, -h};
        r8[2 * tid] = re[tid]; r8[2 * tid + 1] = im[tid];
    }
    __syncthreads();
    const uint64_t chunk = (uint64_t)blockIdx.x * 256 + tid;
    const uint32_t L = Q.L, dhi = E.d - L;
    double ar = 0.0, ai = 0.0;
    if ((chunk >> dhi) == 0) {
        uint32_t ua[NW], ub[NW], uo[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) { ua[w] = Q.ustart[w]; ub[w] = Q.ustart[NW + w]; uo[w] = Q.ustart[2 * NW + w]; }
        uint64_t sv = Q.s0;
        double cr = Q.cre, ci = Q.cim;
        if (chunk) {
            for (uint32_t j = 0; j < dhi; ++j) {
                if ((chunk >> j) & 1) {
                    const uint32_t* row = E.ku + (size_t)(L + j) * 3 * NW;
#pragma unroll
                    for (int w = 0; w < NW; ++w) { ua[w] ^= row[w]; ub[w] ^= row[NW + w]; uo[w] ^= row[2 * NW + w]; }
                    sv ^= E.kn[L + j];
                }
            }
            plain_weight(E, Q, sv, cr, ci);
        }
        const uint32_t steps = 1u << L;
        for (uint32_t i = 0; i < steps; ++i) {
            int pc1 = 0, pc3 = 0, cnt = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                pc1 += __popc(uo[w] & E.diag[2 * NW + w]);
                pc3 += __popc(uo[w] & E.diag[3 * NW + w]);
                cnt += __popc((E.diag[w] ^ ua[w]) & (E.diag[NW + w] ^ ub[w]));
            }
            const int n = E.sigma - 2 * (pc1 - pc3) + ((cnt & 1) << 2);
            const int k8 = n & 7;
            const double rr = r8[2 * k8], ri = r8[2 * k8 + 1];
            const double tr = cr * rr - ci * ri, ti = cr * ri + ci * rr;
            ar += tr; ai += ti;
            if (i + 1 < steps) {
                const uint32_t j = __builtin_ctz(i + 1);
                const uint32_t* row = E.ku + (size_t)j * 3 * NW;
#pragma unroll
                for (int w = 0; w < NW; ++w) { ua[w] ^= row[w]; ub[w] ^= row[NW + w]; uo[w] ^= row[2 * NW + w]; }
                if (((i + 1) & 2047u) == 0) {
                    sv ^= E.kn[j];
                    plain_weight(E, Q, sv, cr, ci);
                } else {
                    for (uint32_t e = E.soff[j]; e < E.soff[j + 1]; ++e) {
                        const uint32_t k = E.supp[e];
                        const double2 m = ((sv >> k) & 1) ? E.down[k] : E.up[k];
                        cmul(cr, ci, m.x, m.y);
                        sv ^= 1ull << k;
                    }
                }
            }
        }
    }
    // block reduction (FP64)
    for (int off = 32; off > 0; off >>= 1) { ar += __shfl_xor(ar, off); ai += __shfl_xor(ai, off); }
    if ((tid & 63) == 0) { red[2 * (tid >> 6)] = ar; red[2 * (tid >> 6) + 1] = ai; }
    __syncthreads();
    if (tid == 0) {
        partial[2 * blockIdx.x] = (red[0] + red[2]) + (red[4] + red[6]);
        partial[2 * blockIdx.x + 1] = (red[1] + red[3]) + (red[5] + red[7]);
    }
}

// Monte-Carlo samples of Engine::estimate (engine.rs:612-659), one per
// thread: sample g is sample g mod 1024 of chunk g / 1024, drawn from
// Rng(seed ^ chunk * gamma).  SplitMix64's state after m steps is
// start + m * gamma, so each thread computes its own random words directly.
// The point s = s0 + sum of the drawn kernel vectors, its value
// weight(s) * omega^phase8(u) in FP64 (the reference's product order), and
// per workgroup the sums of re, im and |x|^2 in a fixed order.
template <int NW>
__global__ __launch_bounds__(256) void k_estimate(PlainDev E, PlainQ Q, uint64_t samples, uint64_t seed, double* __restrict__ partial) {
    __shared__ double r8[16];
    __shared__ double red[12];
    const uint32_t tid = threadIdx.x;
    if (tid < 8) {
        const double h = 0.70710678118654752440;
        const double re[8] = {1, h, 0, -h, -1, -h, 0, h}, im[8] = {0, h, 1, h, 0, -h, -1, -h};
        r8[2 * tid] = re[tid]; r8[2 * tid + 1] = im[tid];
    }
    __syncthreads();
    constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
    const uint64_t g = (uint64_t)blockIdx.x * 256 + tid;
    double xr = 0.0, xi = 0.0, xq = 0.0;
    if (g < samples) {
        const uint64_t ci = g >> 10, k = g & 1023u;
        const uint64_t start = (seed ^ (ci * kGamma)) + kGamma;       // Rng::new
        uint32_t ua[NW], ub[NW], uo[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) { ua[w] = Q.ustart[w]; ub[w] = Q.ustart[NW + w]; uo[w] = Q.ustart[2 * NW + w]; }
        uint64_t sv = Q.s0;
        // d <= 64 (kernel vectors over at most 64 terms): one word per sample
        uint64_t z = start + (k + 1) * kGamma;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        for (uint32_t j = 0; j < E.d; ++j) {
            if ((z >> j) & 1) {
                const uint32_t* row = E.ku + (size_t)j * 3 * NW;
#pragma unroll
                for (int w = 0; w < NW; ++w) { ua[w] ^= row[w]; ub[w] ^= row[NW + w]; uo[w] ^= row[2 * NW + w]; }
                sv ^= E.kn[j];
            }
        }
        double cr, ci2;
        plain_weight(E, Q, sv, cr, ci2);
        int pc1 = 0, pc3 = 0, cnt = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            pc1 += __popc(uo[w] & E.diag[2 * NW + w]);
            pc3 += __popc(uo[w] & E.diag[3 * NW + w]);
            cnt += __popc((E.diag[w] ^ ua[w]) & (E.diag[NW + w] ^ ub[w]));
        }
        const int k8 = (E.sigma - 2 * (pc1 - pc3) + ((cnt & 1) << 2)) & 7;
        const double rr = r8[2 * k8], ri = r8[2 * k8 + 1];
        xr = cr * rr - ci2 * ri;
        xi = cr * ri + ci2 * rr;
        xq = xr * xr + xi * xi;
    }
    for (int off = 32; off > 0; off >>= 1) { xr += __shfl_xor(xr, off); xi += __shfl_xor(xi, off); xq += __shfl_xor(xq, off); }
    if ((tid & 63) == 0) { red[3 * (tid >> 6)] = xr; red[3 * (tid >> 6) + 1] = xi; red[3 * (tid >> 6) + 2] = xq; }
    __syncthreads();
    if (tid < 3) partial[3 * (uint64_t)blockIdx.x + tid] = (red[tid] + red[3 + tid]) + (red[6 + tid] + red[9 + tid]);
}

__global__ __launch_bounds__(256) void k_sum_pairs(const double* __restrict__ in, uint64_t n, double* __restrict__ out) {
    __shared__ double red[8];
    const uint32_t tid = threadIdx.x;
    double r = 0, i = 0;
    for (uint64_t k = tid; k < n; k += 256) { r += in[2 * k]; i += in[2 * k + 1]; }
    for (int off = 32; off > 0; off >>= 1) { r += __shfl_xor(r, off); i += __shfl_xor(i, off); }
    if ((tid & 63) == 0) { red[2 * (tid >> 6)] = r; red[2 * (tid >> 6) + 1] = i; }
    __syncthreads();
    if (tid == 0) { out[0] = (red[0] + red[2]) + (red[4] + red[6]); out[1] = (red[1] + red[3]) + (red[5] + red[7]); }
}

namespace {
PlainQ plain_query(const DevEngine& D, const Prepared& P) {
    const PlainDev& E = D.plain;
    const HostEngine& H = *D.host;
    PlainQ Q{};
    const ULayout& L = H.L;
    std::vector<uint32_t> us;
    put_u32(us, P.u_start.data() + L.off_a(), L.Wg, L.g, E.NW);
    put_u32(us, P.u_start.data() + L.off_b(), L.Wg, L.g, E.NW);
    put_u32(us, P.u_start.data() + L.off_o(), L.Wo, L.nodd, E.NW);
    if (us.size() > 96) throw std::invalid_argument("plain kernel state above 96 dwords");
    std::copy(us.begin(), us.end(), Q.ustart);
    const std::complex<double> cf = plain_c_rest(H, P);
    const std::complex<double> cs = plain_c_start(H, P.s0) * cf;
    Q.s0 = P.s0; Q.cre = cs.real(); Q.cim = cs.imag();
    Q.fre = cf.real(); Q.fim = cf.imag();
    return Q;
}
}  // namespace

McSums dev_plain_estimate(const DevEngine& D, const Prepared& P, uint64_t samples, uint64_t seed) {
    Ctx& C = ctx();
    const PlainDev& E = D.plain;
    if (E.NW == 0) throw std::invalid_argument("normal form too large for the plain kernel (more than 1024 planes or lines)");
    if (E.d > 64) throw std::invalid_argument("branch dimension above 64");
    McSums S;
    if (samples == 0 || !P.nonzero) return S;
    const PlainQ Q = plain_query(D, P);
    const uint64_t blocks = (samples + 255) / 256;
    if (blocks > (1ull << 31)) throw std::invalid_argument("too many samples for one launch");
    std::lock_guard<std::mutex> lk(C.mu);
    double* part = (double*)C.d_misc.get(blocks * 24);
    switch (E.NW) {
        case 2: hipLaunchKernelGGL(k_estimate<2>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, samples, seed, part); break;
        case 4: hipLaunchKernelGGL(k_estimate<4>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, samples, seed, part); break;
        case 8: hipLaunchKernelGGL(k_estimate<8>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, samples, seed, part); break;
        case 16: hipLaunchKernelGGL(k_estimate<16>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, samples, seed, part); break;
        default: hipLaunchKernelGGL(k_estimate<32>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, samples, seed, part); break;
    }
    HIPCHECK(hipGetLastError());
    std::vector<double> h(blocks * 3);
    HIPCHECK(hipMemcpyAsync(h.data(), part, blocks * 24, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipStreamSynchronize(C.stream));
    for (uint64_t b = 0; b < blocks; ++b) { S.re += h[3 * b]; S.im += h[3 * b + 1]; S.sq += h[3 * b + 2]; }
    S.n = samples;
    return S;
}

std::complex<double> dev_plain_sum(const DevEngine& D, const Prepared& P, double* kernel_ms) {
    Ctx& C = ctx();
    const PlainDev& E = D.plain;
    if (E.NW == 0) throw std::invalid_argument("normal form too large for the plain kernel (more than 1024 planes or lines)");
    if (!P.nonzero) return {0, 0};
    PlainQ Q = plain_query(D, P);
    // Gray chunk length: the reference's 2^14, shortened so that at least 2^16
    // threads exist when d allows (one query has to fill the device alone).
    const uint32_t d = E.d;
    uint32_t Lb = d > 16 ? std::min<uint32_t>(14, d - 16) : 0;
    Lb = std::max<uint32_t>(Lb, std::min<uint32_t>(d, 4));
    Q.L = Lb;
    const uint64_t chunks = 1ull << (d - Lb);
    const uint64_t blocks = (chunks + 255) / 256;
    if (blocks > (1ull << 31)) throw std::invalid_argument("branch dimension too large for the plain kernel");
    std::lock_guard<std::mutex> lk(C.mu);
    double* part = (double*)C.d_misc.get(blocks * 16);
    double* res = (double*)C.d_out.get(16);
    HIPCHECK(hipEventRecord(C.e0, C.stream));
    switch (E.NW) {
        case 2: hipLaunchKernelGGL(k_plain<2>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, part); break;
        case 4: hipLaunchKernelGGL(k_plain<4>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, part); break;
        case 8: hipLaunchKernelGGL(k_plain<8>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, part); break;
        case 16: hipLaunchKernelGGL(k_plain<16>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, part); break;
        default: hipLaunchKernelGGL(k_plain<32>, dim3((uint32_t)blocks), dim3(256), 0, C.stream, E, Q, part); break;
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventRecord(C.e1, C.stream));
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(256), 0, C.stream, part, blocks, res);
    HIPCHECK(hipGetLastError());
    double h[2];
    HIPCHECK(hipMemcpyAsync(h, res, 16, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipStreamSynchronize(C.stream));
    if (kernel_ms) { float ms; HIPCHECK(hipEventElapsedTime(&ms, C.e0, C.e1)); *kernel_ms = ms; }
    return {h[0], h[1]};
}

// ============================================================== Z8 walk ====
// Engine tables are read through the constant address space: they are never
// written while a kernel runs, and uniform loads through such pointers become
// scalar (SGPR) loads even though the persistent loop also performs atomics.
#define PQ_CONST __attribute__((address_space(4)))
template <class T> __device__ __forceinline__ const PQ_CONST T* cview(const T* p) { return (const PQ_CONST T*)p; }
// Whole-struct scalar load of a descriptor (dword by dword through the
// constant address space).
template <class T> __device__ __forceinline__ T cload(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "descriptor size");
    T out;
    const PQ_CONST uint32_t* src = cview((const uint32_t*)p);
    uint32_t* dst = (uint32_t*)&out;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) dst[k] = src[k];
    return out;
}

// x with its sign bit flipped where sw has bit 31 set (sw is 0 or 0x80000000)
__device__ __forceinline__ double flip_sign(double m, uint32_t sw) { return __hiloint2double(__double2hiint(m) ^ (int)sw, __double2loint(m)); }

// Rotate (xr, xi) by i^((K>>1)&3) and add it, in FP64, into the even or odd
// accumulator (both are updated, the other with zero, so the accumulator
// index stays static and the array stays in registers).
__device__ __forceinline__ void fold64(double xr, double xi, uint32_t K, double (&acc)[4]) {
    const uint32_t r2 = (K >> 1) & 3u;
    const double a = (r2 & 1u) ? xi : xr;
    const double b = (r2 & 1u) ? xr : xi;
    // signs of i^r2 (r2 = 1, 2 negate the real part, r2 = 2, 3 the imaginary)
    const double yr = flip_sign(a, ((r2 + 1u) & 2u) << 30);
    const double yi = flip_sign(b, (r2 & 2u) << 30);
    // even / odd by FMAs with 1 and 0 (exact: x * 0 + acc = acc)
    const double e = (K & 1u) ? 0.0 : 1.0, o = (K & 1u) ? 1.0 : 0.0;
    acc[0] = fma(e, yr, acc[0]);
    acc[1] = fma(e, yi, acc[1]);
    acc[2] = fma(o, yr, acc[2]);
    acc[3] = fma(o, yi, acc[3]);
}

// Lane state over the outer coordinates: integer phase K at c_in = 0, the
// inner sign mask hs (cross terms with the outer point, xor the query bits of
// the inner slots), inner-touched pivots pio, branch vector s at c_in = 0.
struct LaneState {
    uint64_t cout, sout;
    uint32_t K, hs, pio;
};
// Inner-table geometry of a walk with NIN inner slots: 2^NIN entries, 16
// accumulators for the four low slots, 2^(NIN - 4) steps over the high ones.
template <uint32_t NIN>
struct InnerGeom {
    static constexpr uint32_t N = 1u << NIN, mask = N - 1, high = NIN - 4, steps = 1u << (NIN - 4);
};
template <uint32_t NIN>
__device__ __forceinline__ void seed_lane(const Z8Dev& E, const Z8Q& Q, uint64_t lane, LaneState& S) {
    S.cout = 0; S.sout = Q.s0; S.K = Q.K0; S.hs = (uint32_t)Q.delta & InnerGeom<NIN>::mask; S.pio = Q.pio0;
    const uint32_t base = NIN + Q.M;
    for (uint32_t b = 0; b < Q.H; ++b) {
        const uint32_t p = base + b;
        const uint64_t qp = cview(E.slot_q)[p];
        const uint32_t ell = cview(E.slot_lam)[p] + (((uint32_t)(Q.delta >> p) & 1u) << 2);
        const uint32_t pin = cview(E.slot_pin)[p];
        const uint64_t np = cview(E.slot_n)[p];
        if ((lane >> b) & 1) {
            S.K += ell + ((__popcll(qp & S.cout) & 1u) << 2);
            S.cout |= 1ull << p;
            S.hs ^= (uint32_t)qp & InnerGeom<NIN>::mask;
            S.pio ^= pin;
            S.sout ^= np;
        }
    }
}
template <uint32_t NIN>
__device__ __forceinline__ void flip_mid(const Z8Dev& E, const Z8Q& Q, uint32_t p, LaneState& S) {
    const uint64_t qp = cview(E.slot_q)[p];
    const uint32_t ell = cview(E.slot_lam)[p] + (((uint32_t)(Q.delta >> p) & 1u) << 2);
    const uint32_t par = (__popcll(qp & S.cout) & 1u) << 2;
    const bool was = (S.cout >> p) & 1;
    S.K += (was ? (8u - ell) : ell) + par;
    S.cout ^= 1ull << p;
    S.hs ^= (uint32_t)qp & InnerGeom<NIN>::mask;
    S.pio ^= cview(E.slot_pin)[p];
    S.sout ^= cview(E.slot_n)[p];
}
__device__ __forceinline__ double lane_factor(const double* s_lt, uint32_t nbytes, uint64_t sout) {
    double F = 1.0;
    for (uint32_t b = 0; b < nbytes; ++b) F *= s_lt[(b << 8) | (uint32_t)((sout >> (8 * b)) & 0xFFu)];
    return F;
}

// The 2^NIN inner branches of one outer point.  Each table entry is real or
// imaginary; with IMAG the class is bit 0 of c_in (the builder arranges the
// inner basis so), without it every entry is real.  The sign of branch c_in is
// (-1)^|c_in & hs|, hs being the lane's inner sign mask.
//
// Without inner pivots the table lives in VGPRs and the walk keeps 16
// accumulators, one per pattern l of the four low inner slots, walking the
// NIN - 4 high slots in Gray order (inner_steps, inner_walk_mid below).
//
// With inner pivots (inner_walk_pivots) every branch also reads its pivot
// factor gin[pio ^ pin(c_in)] from LDS (the address follows the same Gray
// recurrence as the sign) and flips that factor's sign word: per branch a sign
// XOR, an address XOR, an LDS read and one FP64 FMA with the table entry (a
// 16-dword scalar load per 16 entries).

// The 2^NIN-entry inner table held in VGPRs: lane r of each 16-lane row
// keeps entries r, 16 + r, 32 + r, ... (2^(NIN - 4) register pairs), so entry
// e is entry e & 15 of row-local register e >> 4 and reaches every lane of the row
// through a DPP row broadcast (row_newbcast) folded into the FMA itself:
// v_fmac_f64_dpp issues at the full FP64 rate, and the walk then reads no
// memory at all.  Every lane of the wave must be active while the table is
// broadcast (a broadcast from a disabled lane does not write).
template <uint32_t NIN>
struct InnerRegs { double w[InnerGeom<NIN>::steps]; };
template <uint32_t NIN>
__device__ __forceinline__ void load_inner_regs(const Z8Dev& E, InnerRegs<NIN>& T) {
    const uint32_t r = threadIdx.x & 15u;
#pragma unroll
    for (int k = 0; k < (int)InnerGeom<NIN>::steps; ++k) T.w[k] = E.inner_acc[16 * k + r];
    // a DPP read of a VGPR needs two wait states after a VALU write of it
    asm volatile("s_nop 1");
}
// acc += table entry (row lane SEL of w) * m
template <int SEL>
__device__ __forceinline__ void fmac_bcast(double& acc, double w, double m) {
    asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf" : "+v"(acc) : "v"(w), "v"(m), "i"(SEL));
}
// table entry (row lane SEL of w)
template <int SEL>
__device__ __forceinline__ double bcast(double w) {
    double r;
    asm("v_mov_b64_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf" : "=v"(r) : "v"(w), "i"(SEL));
    return r;
}
// Step Q of the inner Gray walk: flip the multiplier's sign for high slot
// ctz(Q), then entries 16Q .. 16Q + 15 (register pair Q, all 16 row lanes)
// into the 16 low-slot accumulators.  Step 0 (multiplier +1) sets the
// accumulators to its entries.
template <int L>
__device__ __forceinline__ void bcast_all(double (&a)[16], double w) {
    if constexpr (L < 16) { a[L] = bcast<L>(w); bcast_all<L + 1>(a, w); }
}
template <int L>
__device__ __forceinline__ void fmac_all(double (&a)[16], double w, double m) {
    if constexpr (L < 16) { fmac_bcast<L>(a[L], w, m); fmac_all<L + 1>(a, w, m); }
}
template <uint32_t NIN, int Q>
__device__ __forceinline__ void inner_steps(double (&a)[16], const InnerRegs<NIN>& T, double& m, const uint32_t (&sm)[InnerGeom<NIN>::high]) {
    if constexpr (Q < (int)InnerGeom<NIN>::steps) {
        if constexpr (Q == 0) {
            bcast_all<0>(a, T.w[0]);
        } else {
            m = flip_sign(m, sm[__builtin_ctz(Q)]);
            asm volatile("" : "+v"(m));     // one running value: the flip is an XOR in place
            fmac_all<0>(a, T.w[Q], m);
        }
        inner_steps<NIN, Q + 1>(a, T, m, sm);
    }
}

// Signs of the four low inner slots applied to the 16 accumulators of one
// outer point: (tr, ti) of its 256-branch sum (bit 0 of l is the class).
// The same walk for an engine whose inner directions touch pivots but whose
// four low inner slots do not: the pivot factor gin[pio ^ pin(c_in)] then
// depends on the step q alone, and the signed factor is the multiplier m of
// the step's 16 FMAs.  A step whose flipped high slot touches no pivot (ps
// zero, uniform: a scalar branch) only flips m's sign, as the walk without
// pivots does; one that touches pivots reads the new factor from LDS (the
// byte offset follows the Gray recurrence over the high slots' pivots).  The
// builder puts the pivot-touching inner directions in the highest slots,
// whose flips are the rarest of the Gray walk.  The accumulators start at
// zero.
template <uint32_t NIN, int Q>
__device__ __forceinline__ void inner_steps_piv(double (&a)[16], const InnerRegs<NIN>& T, const double* s_gin, uint32_t addr, uint32_t sgn,
                                                double m, const uint32_t (&ps)[InnerGeom<NIN>::high],
                                                const uint32_t (&sm)[InnerGeom<NIN>::high]) {
    if constexpr (Q < (int)InnerGeom<NIN>::steps) {
        if constexpr (Q == 0) {
            m = flip_sign(*(const double*)((const char*)s_gin + addr), sgn);
        } else {
            constexpr int j = __builtin_ctz(Q);
            sgn ^= sm[j];
            if (ps[j]) {
                addr ^= ps[j];
                m = flip_sign(*(const double*)((const char*)s_gin + addr), sgn);
            } else {
                m = flip_sign(m, sm[j]);
            }
        }
        fmac_all<0>(a, T.w[Q], m);
        inner_steps_piv<NIN, Q + 1>(a, T, s_gin, addr, sgn, m, ps, sm);
    }
}

// Signs of the four low inner slots, s[k] = (-1)^(bit k of hs), applied to
// the 16 accumulators of one outer point: (tr, ti) of its 256-branch sum (bit
// 0 of l is the class).
template <bool IMAG>
__device__ __forceinline__ void combine_signs(const double (&s)[4], const double (&a)[16], double& tr, double& ti) {
    double v[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const double e0 = fma(s[1], a[c + 2], a[c]), e1 = fma(s[1], a[c + 6], a[c + 4]);
        const double e2 = fma(s[1], a[c + 10], a[c + 8]), e3 = fma(s[1], a[c + 14], a[c + 12]);
        v[c] = fma(s[3], fma(s[2], e3, e2), fma(s[2], e1, e0));
    }
    if constexpr (IMAG) { tr = v[0]; ti = s[0] * v[1]; }
    else { tr = fma(s[0], v[1], v[0]); ti = 0.0; }
}


template <uint32_t NIN, bool IMAG>
__device__ __forceinline__ void inner_walk_pivots(const Z8Dev& E, const uint32_t (&ps)[NIN], const double* s_gin, const LaneState& S,
                                                  double F, double (&acc)[4]) {
    double tr = 0.0, ti = 0.0;
    {
        const PQ_CONST double* Wv = cview(E.inner);
        uint32_t sm[NIN];
#pragma unroll
        for (int j = 0; j < (int)NIN; ++j) sm[j] = (S.hs << (31 - j)) & 0x80000000u;
        uint32_t sgn = 0;
        uint32_t addr = S.pio << 3;      // LDS byte offset of gin[pio ^ pin(gray(i))]
#pragma unroll
        for (int blk = 0; blk < (int)InnerGeom<NIN>::steps; ++blk) {
            double w[16];
            const PQ_CONST double* bw = Wv + 16 * blk;
            asm volatile("" : "+s"(bw) : : "memory");
#pragma unroll
            for (int k = 0; k < 16; ++k) w[k] = bw[k];
#pragma unroll
            for (int ii = 0; ii < 16; ++ii) {
                const int i = 16 * blk + ii;
                if (i) {
                    sgn ^= sm[__builtin_ctz(i)];
                    addr ^= ps[__builtin_ctz(i)];
                }
                const double g = *(const double*)((const char*)s_gin + addr);
                const double m = __hiloint2double(__double2hiint(g) ^ (int)sgn, __double2loint(g));
                if (IMAG && ((i ^ (i >> 1)) & 1)) ti = fma(m, w[ii], ti);
                else tr = fma(m, w[ii], tr);
            }
        }
    }
    fold64(tr * F, ti * F, S.K, acc);
}

// One mid point without inner pivots, with the point's sign words sm (hs bits
// 4.. at bit 31) and low-slot signs slo (hs bits 0..3 as +-1.0) carried from
// point to point, and the fold done in the lane's own frame.  The point's
// phase is K_seed + D + 4 par with D the same for every lane, so its rotation
// i^((D >> 1) & 3) and its EV / OD routing are the same for every lane too:
// they are FMAs with uniform 0 / +-1 coefficients (SGPR operands), and the
// lane's part is the sign par (bit 31 of parw) on the point's magnitude fm.
// The lane frame, omega^K_seed and the part of the lane factor common to all
// its points, is applied once at the end of the lane's walk.  With PIV the
// inner steps carry the pivot factor (inner_steps_piv, pio the point's inner
// pivots).
template <uint32_t NIN, bool IMAG, bool PIV>
__device__ __forceinline__ void inner_walk_mid(const InnerRegs<NIN>& T, const uint32_t (&sm)[InnerGeom<NIN>::high], const double (&slo)[4],
                                               uint32_t parw, double fm, uint32_t D, double (&accl)[4], const double* s_gin, uint32_t pio,
                                               const uint32_t (&ps)[InnerGeom<NIN>::high]) {
    double a[16];
    if constexpr (PIV) {
#pragma unroll
        for (int l = 0; l < 16; ++l) a[l] = 0.0;
        inner_steps_piv<NIN, 0>(a, T, s_gin, pio << 3, 0u, 0.0, ps, sm);
    } else {
        double m = 1.0;
        asm volatile("" : "+v"(m));
        inner_steps<NIN, 0>(a, T, m, sm);
    }
    double tr, ti;
    combine_signs<IMAG>(slo, a, tr, ti);
    const double fs = flip_sign(fm, parw);
    tr *= fs;
    // uniform coefficients: i^r (tr + i ti) into EV (D even) or OD (D odd)
    const uint32_t r = (D >> 1) & 3u, odd = D & 1u;
    const double c = (r == 0u) ? 1.0 : (r == 2u) ? -1.0 : 0.0;      // re i^r
    const double sn = (r == 1u) ? 1.0 : (r == 3u) ? -1.0 : 0.0;     // im i^r
    const double ce = odd ? 0.0 : c, se = odd ? 0.0 : sn, co = odd ? c : 0.0, so = odd ? sn : 0.0;
    accl[0] = fma(ce, tr, accl[0]);
    accl[1] = fma(se, tr, accl[1]);
    accl[2] = fma(co, tr, accl[2]);
    accl[3] = fma(so, tr, accl[3]);
    if constexpr (IMAG) {
        ti *= fs;
        accl[0] = fma(-se, ti, accl[0]);
        accl[1] = fma(ce, ti, accl[1]);
        accl[2] = fma(-so, ti, accl[2]);
        accl[3] = fma(co, ti, accl[3]);
    }
}

// One instantiation per inner-step variant (pivot lookups or not, imaginary
// class pattern or not); a batch launches one persistent grid per variant
// present, each over its own contiguous range of queries and items.  The
// variant with pivot lookups walks kInnerPiv inner slots, the other
// kInnerWide (the engine images are built to match).
template <bool HASG, bool IMAG>
__global__ __launch_bounds__(256) void k_z8_walk(const Z8Dev* __restrict__ engs, const Z8Q* __restrict__ qs, uint32_t q_lo, uint32_t q_hi,
                                                 uint64_t item_begin, uint64_t item_end, unsigned long long* __restrict__ counter,
                                                 double* __restrict__ partial) {
    constexpr uint32_t NIN = HASG ? kInnerPiv : kInnerWide;
    constexpr uint32_t NH = InnerGeom<NIN>::high;
    __shared__ double s_lt[8 * 256];
    __shared__ double s_gin[HASG ? 4096 : 1];
    __shared__ double s_fm[HASG ? 1 : 256];
    __shared__ unsigned long long s_item;
    __shared__ int s_key, s_fq;
    __shared__ double s_red[4][4];
    const uint32_t tid = threadIdx.x;
    if (tid == 0) { s_key = -1; s_fq = -1; }
    // the inner table stays in VGPRs from item to item of one engine
    InnerRegs<NIN> T;
    uint32_t t_eng = 0xFFFFFFFFu;
    // a workgroup's first item is its own index; later ones come from the
    // queue (counter starts at 0)
    bool first = true;
    for (;;) {
        __syncthreads();
        if (tid == 0) s_item = first ? item_begin + blockIdx.x : item_begin + gridDim.x + atomicAdd(counter, 1ull);
        first = false;
        __syncthreads();
        const uint64_t item = s_item;
        if (item >= item_end) break;
        uint32_t lo = q_lo, hi = q_hi;
        while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (cview(qs)[mid].item0 <= item) lo = mid; else hi = mid; }
        const Z8Q Q = cload(qs + lo);
        const Z8Dev E = cload(engs + Q.eng);
        // Mid slots that are all pivot-free (one free term each, touched by no
        // other slot): their magnitudes are the same for every lane of the
        // query, one table entry per mid point, and a lane's factor is
        // computed once instead of at every mid point.
        const bool fast = !HASG && Q.M <= E.nx;
        const int key = 2 * (int)Q.eng + (fast ? 1 : 0);
        if (key != s_key) {
            const double* lt = fast ? E.lt_rest : E.lt;
            for (uint32_t k = tid; k < E.nbytes * 256u; k += 256) s_lt[k] = lt[k];
            if constexpr (HASG) for (uint32_t k = tid; k < (1u << E.npin); k += 256) s_gin[k] = E.gin[k];
            __syncthreads();
            if (tid == 0) s_key = key;
        }
        if constexpr (!HASG) {
            if (fast && (int)lo != s_fq) {
                for (uint32_t o = tid; o < (1u << Q.M); o += 256) {
                    const uint32_t g = o ^ (o >> 1);
                    double f = 1.0;
                    for (uint32_t j = 0; j < Q.M; ++j) f *= E.mid_r[2 * j + (((uint32_t)(Q.s0 >> E.mid_term[j]) ^ (g >> j)) & 1u)];
                    s_fm[o] = f;
                }
                __syncthreads();
                if (tid == 0) s_fq = (int)lo;
            }
        }
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const uint64_t lane = (item - Q.item0) * 256 + tid;
        const bool valid = (lane >> Q.H) == 0;
        // the table in VGPRs: always without inner pivots, and with them when
        // the four low inner slots are pivot-free (uniform per engine)
        const bool lanewalk = !HASG || E.lowfree;
        if (lanewalk && Q.eng != t_eng) { load_inner_regs<NIN>(E, T); t_eng = Q.eng; }
        // without pivots every lane walks (the table broadcast needs the whole
        // wave); lanes past the query's 2^H repeat a valid lane's coset and
        // are zeroed below
        if (lanewalk || valid) {
            LaneState S;
            seed_lane<NIN>(E, Q, lane, S);
            uint32_t ps[NIN];
#pragma unroll
            for (int j = 0; j < (int)NIN; ++j) ps[j] = HASG ? (cview(E.slot_pin)[j] << 3) : 0u;
            const uint32_t nout = 1u << Q.M;
            if (lanewalk) {
                // The fast path (pivot-free mid slots): the lane's factor without
                // the mid slots' terms, times the magnitudes of the pivot-free
                // slots among its hi slots; each mid point multiplies in its
                // per-query magnitude.  Otherwise a mid slot can share a pivot
                // with the hi slots: each mid point takes the whole lane factor
                // of its branch vector from the byte tables.  With inner pivots
                // (the four low inner slots pivot-free) the inner steps also
                // carry the pivot factor of the point's inner pivots pio.
                double Fl = 1.0;
                if (fast) {
                    Fl = lane_factor(s_lt, E.nbytes, S.sout);
                    for (uint32_t j = Q.M; j < E.nx; ++j) {
                        const double r0 = cview(E.mid_r)[2 * j], r1 = cview(E.mid_r)[2 * j + 1];
                        Fl *= ((S.sout >> cview(E.mid_term)[j]) & 1) ? r1 : r0;
                    }
                }
                // each mid slot's cross term with the lane's hi slots
                uint32_t P = 0;
                for (uint32_t j = 0; j < Q.M; ++j) P |= (uint32_t)(__popcll(cview(E.slot_q)[NIN + j] & S.cout) & 1) << j;
                // along the mid walk the phase step D of each flip is the same
                // for every lane; the lane's cross term only flips the sign
                // (parw), and the sign-mask step is a uniform byte
                uint32_t cmid = 0, D = 0, parw = 0;
                double accl[4] = {0.0, 0.0, 0.0, 0.0};
                // the point's sign words and low-slot signs, updated in place by
                // the uniform sign-mask step of each flip
                uint32_t sm[NH];
                double slo[4];
#pragma unroll
                for (int k = 0; k < (int)NH; ++k) sm[k] = (S.hs << (27 - k)) & 0x80000000u;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    slo[k] = (S.hs >> k) & 1u ? -1.0 : 1.0;
                    asm volatile("" : "+v"(slo[k]));
                }
                for (uint32_t o = 0; o < nout; ++o) {
                    if (o) {
                        const uint32_t j = (uint32_t)__builtin_ctz(o), p = NIN + j;
                        const uint64_t qp = cview(E.slot_q)[p];
                        const uint32_t ell = cview(E.slot_lam)[p] + (((uint32_t)(Q.delta >> p) & 1u) << 2);
                        D += (((cmid >> j) & 1u) ? 8u - ell : ell) + ((uint32_t)(__popcll(qp & ((uint64_t)cmid << NIN)) & 1) << 2);
                        parw ^= (P << (31 - j)) & 0x80000000u;
                        const uint32_t x = (uint32_t)qp & InnerGeom<NIN>::mask;      // the sign-mask step, the same for every lane
#pragma unroll
                        for (int k = 0; k < (int)NH; ++k) sm[k] ^= (x << (27 - k)) & 0x80000000u;
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            slo[k] = flip_sign(slo[k], (x << (31 - k)) & 0x80000000u);
                            asm volatile("" : "+v"(slo[k]));
                        }
                        cmid ^= 1u << j;
                        if (!fast) S.sout ^= cview(E.slot_n)[p];
                        if (HASG) S.pio ^= cview(E.slot_pin)[p];
                    }
                    const double fm = fast ? s_fm[o] : lane_factor(s_lt, E.nbytes, S.sout);
                    uint32_t pshi[NH];
#pragma unroll
                    for (int k = 0; k < (int)NH; ++k) pshi[k] = ps[4 + k];
                    inner_walk_mid<NIN, IMAG, HASG>(T, sm, slo, parw, fm, D, accl, s_gin, S.pio, pshi);
                }
                // undo the lane frame: omega^K_seed (EV + omega OD) times the lane factor
                fold64(accl[0] * Fl, accl[1] * Fl, S.K, acc);
                fold64(accl[2] * Fl, accl[3] * Fl, S.K + 1u, acc);
            } else {
                for (uint32_t o = 0; o < nout; ++o) {
                    if (o) flip_mid<NIN>(E, Q, NIN + (uint32_t)__builtin_ctz(o), S);
                    const double F = lane_factor(s_lt, E.nbytes, S.sout);
                    inner_walk_pivots<NIN, IMAG>(E, ps, s_gin, S, F, acc);
                }
            }
        }
        if (!valid) acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
        // workgroup reduction in FP64, in a fixed order
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] += __shfl_xor(acc[k], off);
        }
        if ((tid & 63) == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) s_red[tid >> 6][k] = acc[k];
        }
        __syncthreads();
        if (tid < 4) partial[4 * item + tid] = (s_red[0][tid] + s_red[1][tid]) + (s_red[2][tid] + s_red[3][tid]);
    }
}

// FP64 reduction of one query's item partials in a fixed order (strided by
// thread, then a fixed shuffle and LDS tree, so the result does not depend on
// which workgroup runs it), EV + omega OD, and the FP64 scale.  All 256
// threads of the workgroup take part; thread 0 writes the result.
// host_query_reduce restates this order exactly.
__device__ __forceinline__ void query_reduce(uint64_t item0, uint64_t nitems, const double* partial, double2 sc, double* dst,
                                             double (&red)[4][4]) {
    const uint32_t tid = threadIdx.x;
    double v[4] = {0, 0, 0, 0};
    for (uint64_t k = tid; k < nitems; k += 256) {
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] += partial[4 * (item0 + k) + c];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] += __shfl_xor(v[c], off);
    }
    if ((tid & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) red[tid >> 6][c] = v[c];
    }
    __syncthreads();
    if (tid == 0) {
        double t[4];
        for (int c = 0; c < 4; ++c) t[c] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
        const double h = 0.70710678118654752440;
        // EV + omega * OD
        const double re = t[0] + h * (t[2] - t[3]);
        const double im = t[1] + h * (t[2] + t[3]);
        dst[0] = sc.x * re - sc.y * im;
        dst[1] = sc.x * im + sc.y * re;
    }
}

// ======================================================= segment walk ====
// Double-float (FP32 hi/lo pair) products and sums.  ph + pl is the product
// of (ah + al)(bh + bl) up to the dropped al*bl term; the sum keeps the
// rounding error of the hi part in the lo part.
__device__ __forceinline__ void df_mul(float ah, float al, float bh, float bl, float& ph, float& pl) {
    const float p = ah * bh;
    float e = fmaf(ah, bh, -p);
    e = fmaf(ah, bl, e);
    e = fmaf(al, bh, e);
    ph = p; pl = e;
}
__device__ __forceinline__ void df_add(float& sh, float& sl, float xh, float xl) {
    const float t = sh + xh;
    const float bp = t - sh;
    const float e = (sh - (t - bp)) + (xh - bp);
    sh = t;
    sl = sl + (xl + e);
}

// Lane state of the segment walk over the outer coordinates.
struct SegLaneDev {
    uint64_t cout, s;
    uint32_t K, hs, pio;
    float fh, fl;            // magnitude of the walkable slots taken as lane bits
};
// Lane bit b -> outer slot: the lane-low slots first, then past the M walked slots.
__device__ __forceinline__ uint32_t seg_slot_of_bit(const SegDev& E, uint32_t M, uint32_t b) { return b < E.L6 ? b : E.L6 + M + (b - E.L6); }
// Outer-slot arrays of an engine, read through pointers of the address space
// they live in: LDS copies in k_seg_walk, the device image in the dumps.
template <class P64, class P32, class PF>
struct SegSlots {
    P64 oq, on;
    P32 ox, opin, olam;
    PF rmid;
    P32 low;                 // lane-low table, 6 words per pattern
};
// The lane's outer coordinates c: lane bits below L6 on slots 0..L6-1, the
// rest on slots L6+M.. (the M walked slots stay 0 at the seed).  Every slot
// term depends on c alone, so the slots are independent: no carried chain.
//   K  = K0 + sum_{k in c} (lambda_k + 4 delta_k + 4 |Q_k & c & below(k)|)
//   hs = hs0 ^ X_k..., pio, s likewise; factor = prod of r_{c_k} over the
//   walkable slots beyond M.
template <class P64, class P32, class PF>
__device__ __forceinline__ void seg_seed(const SegDev& E, const SegSlots<P64, P32, PF>& T, const SegQ& Q, uint64_t lane, SegLaneDev& L) {
    const uint32_t M = Q.M, L6 = E.L6;
    const uint64_t c = (lane & ((1ull << L6) - 1ull)) | ((lane >> L6) << (L6 + M));
    // the lane-low slots from the engine's table (their query part is the
    // parity of delta over them), the rest slot by slot
    const uint32_t cl = (uint32_t)(c & ((1ull << L6) - 1ull)), e = 6 * cl;
    uint32_t K = Q.K0 + T.low[e] + (((uint32_t)__popcll(Q.delta & (uint64_t)cl) & 1u) << 2);
    uint32_t hs = Q.hs0 ^ T.low[e + 1], pio = Q.pio0 ^ T.low[e + 2];
    uint64_t s = Q.s0 ^ ((uint64_t)T.low[e + 4] | ((uint64_t)T.low[e + 5] << 32));
    auto slot = [&](uint32_t k) {
        const uint32_t on = (uint32_t)(c >> k) & 1u, m = 0u - on;
        const uint32_t ell = T.olam[k] + (((uint32_t)(Q.delta >> k) & 1u) << 2) +
                             (((uint32_t)__popcll(T.oq[k] & c & ((1ull << k) - 1ull)) & 1u) << 2);
        K += ell & m;
        hs ^= T.ox[k] & m;
        pio ^= T.opin[k] & m;
        s ^= T.on[k] & (0ull - (uint64_t)on);
    };
#pragma unroll 4
    for (uint32_t k = L6 + M; k < E.nout; ++k) slot(k);
    float fh = 1.f, fl = 0.f;
    for (uint32_t k = L6 + M; k < L6 + E.Mcap; ++k) {
        const uint32_t m = 4 * (k - L6) + 2 * ((uint32_t)(c >> k) & 1u);
        float ph, pl;
        df_mul(fh, fl, T.rmid[m], T.rmid[m + 1], ph, pl);
        fh = ph; fl = pl;
    }
    L.cout = c; L.K = K; L.hs = hs; L.pio = pio; L.s = s; L.fh = fh; L.fl = fl;
}
// Lane-factor of the terms outside the table and the walk: byte tables in LDS
// (8 x 256 entries, so every index is in bounds), read four at a time.
__device__ __forceinline__ void seg_lane_factor(const float2* s_lt, uint32_t nbytes, uint64_t s, float& fh, float& fl) {
#pragma unroll
    for (uint32_t b0 = 0; b0 < 8; b0 += 4) {
        if (b0 >= nbytes) break;
        float2 v[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) v[j] = s_lt[((b0 + j) << 8) | (uint32_t)((s >> (8 * (b0 + j))) & 0xFFu)];
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            if (b0 + j < nbytes) {
                float ph, pl;
                df_mul(fh, fl, v[j].x, v[j].y, ph, pl);
                fh = ph; fl = pl;
            }
        }
    }
}
// Over a lane's segment of the M walkable slots, at point o (c_w = gray(o)):
//   K(o)  = K_seed + Ue(o) + 4 (delta_w . c_w) + 4 (P . c_w)   (mod 8)
//   hs(o) = hs_seed ^ X(o)
// with Ue, X the XOR-accumulated Gray-walk data of the engine (one record per
// point), delta_w the query bits of the walk slots (uniform) and
// P_j = parity(Q_{walk j} . c_lane) the only lane-dependent part.  The walk
// slots are pivot-free, so pio and the lane factor are fixed over the segment.
template <class P64>
__device__ __forceinline__ uint32_t seg_lane_parities(const SegDev& E, P64 oq, uint32_t M, uint64_t cout) {
    uint32_t P = 0;
#pragma unroll 4
    for (uint32_t j = 0; j < M; ++j) P |= (uint32_t)(__popcll(oq[E.L6 + j] & cout) & 1) << j;
    return P;
}
typedef float f32x4 __attribute__((ext_vector_type(4)));
// Walks of at most 2^kSegShortWalk points issue all their table reads at once.
constexpr uint32_t kSegShortWalk = 4;
// Long walks issue the table reads of kSegGroup points of one class at a time
// (4, 6 and 8 give the same kernel time on the saturated 4096-amplitude launch).
constexpr uint32_t kSegGroup = 4;
// A += f * T (double-float real times double-float complex).
__device__ __forceinline__ void seg_point(float (&A)[4], float fh, float fl, const f32x4& T) {
    float ph, pl, qh, ql;
    df_mul(fh, fl, T.x, T.y, ph, pl);
    df_mul(fh, fl, T.z, T.w, qh, ql);
    df_add(A[0], A[1], ph, pl);
    df_add(A[2], A[3], qh, ql);
}
// (x + i y) * i^r in double-float, r lane-dependent (selects).
__device__ __forceinline__ void df_rot(float (&v)[4], uint32_t r) {
    const bool sw = r & 1u;
    const uint32_t sre = ((r ^ (r >> 1)) & 1u) << 31, sim = ((r >> 1) & 1u) << 31;
    const float a0 = sw ? v[2] : v[0], a1 = sw ? v[3] : v[1], b0 = sw ? v[0] : v[2], b1 = sw ? v[1] : v[3];
    v[0] = __uint_as_float(__float_as_uint(a0) ^ sre); v[1] = __uint_as_float(__float_as_uint(a1) ^ sre);
    v[2] = __uint_as_float(__float_as_uint(b0) ^ sim); v[3] = __uint_as_float(__float_as_uint(b1) ^ sim);
}
// sum_q omega^q A[q] over q = 0..3 split as EV + omega OD:
// dst = A[P] + i A[P+2] for P = 0 (EV) and 1 (OD).
template <int P>
__device__ __forceinline__ void seg_fold4(const float (&A)[4][4], float (&dst)[4]) {
    dst[0] = A[P][0]; dst[1] = A[P][1]; dst[2] = A[P][2]; dst[3] = A[P][3];
    df_add(dst[0], dst[1], -A[P + 2][2], -A[P + 2][3]);   // i (x + i y) = -y + i x
    df_add(dst[2], dst[3], A[P + 2][0], A[P + 2][1]);
}

// Persistent work-queue kernel over (query, block of 64 WAVES lanes) items.  A lane
// seeds its state from its coset index (integer phase, sign mask, pivots,
// lane factor), then covers the 2^M outer points of its segment.  The point
// records come grouped by engine phase class Ue, so each group adds into its
// own accumulator with no per-point branch.  Per point: one XOR for the table
// index, one scalar-cache read of the record, one global 16-byte read of
// T[pio][hs], the lane and query sign, and a double-float product and sum.
// The 64 lanes of a wavefront differ in the lane-low coordinates, whose sign
// masks the table basis maps to the low six index bits: one wavefront reads
// one contiguous 1 KB block per point.
typedef const __attribute__((address_space(1))) f32x4 GFloat4;
typedef __attribute__((address_space(3))) const uint64_t LdsU64;
typedef __attribute__((address_space(3))) const uint32_t LdsU32;
typedef __attribute__((address_space(3))) const float LdsF32;
// The LDS a workgroup keeps across the items it walks: the engine's byte
// tables and outer-slot arrays (so the lane seed reads LDS, not a chain of
// scalar-cache round trips) and the wave sums.
struct SegLds {
    float2* lt;              // 8 x 256
    uint64_t *oq, *on;       // 64 each
    uint32_t *ox, *opin, *olam;
    float* rmid;             // 4 kSegMaxMid
    double (*red)[4];        // 4 waves x (EV re, im, OD re, im)
    uint32_t* low;           // 6 x 64: the lane-low table
};
#define PQ_SEG_LDS(name)                                                   \
    __shared__ float2 name##_lt[8 * 256];                                  \
    __shared__ uint64_t name##_oq[64], name##_on[64];                      \
    __shared__ uint32_t name##_ox[64], name##_opin[64], name##_olam[64];   \
    __shared__ float name##_rmid[4 * kSegMaxMid];                          \
    __shared__ double name##_red[4][4];                                    \
    __shared__ uint32_t name##_low[6 * 64];                                \
    const SegLds name{name##_lt, name##_oq, name##_on, name##_ox, name##_opin, name##_olam, name##_rmid, name##_red, name##_low}

// Lane total omega^K (EVu + omega ODu) times the lane factor f (double-float).
__device__ __forceinline__ void seg_lane_total(const float (&eu)[4], const float (&ou)[4], uint32_t K, float fh, float fl,
                                               float (&ev)[4], float (&od)[4]) {
    const uint32_t Ks = K & 7u;
    float e2[4], o2[4];
    if (Ks & 1u) {   // omega^(2a+1) (EVu + omega ODu) = i^a (i ODu + omega EVu)
        e2[0] = -ou[2]; e2[1] = -ou[3]; e2[2] = ou[0]; e2[3] = ou[1];
        o2[0] = eu[0]; o2[1] = eu[1]; o2[2] = eu[2]; o2[3] = eu[3];
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) { e2[c] = eu[c]; o2[c] = ou[c]; }
    }
    df_rot(e2, Ks >> 1);
    df_rot(o2, Ks >> 1);
#pragma unroll
    for (int c = 0; c < 4; c += 2) {
        df_mul(e2[c], e2[c + 1], fh, fl, ev[c], ev[c + 1]);
        df_mul(o2[c], o2[c + 1], fh, fl, od[c], od[c + 1]);
    }
}
// A double's two halves moved by one DPP pattern (rows the mask leaves out
// read 0.0).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double x) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xF, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ void dpp_add4(double (&v)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] += dpp_f64<CTRL, ROWS>(v[i]);
}
// Workgroup reduction of the four sums in FP64: a lane's double-float pair
// hi + lo is exact in FP64; each wave adds over DPP in a fixed order (pairs
// and quads by quad_perm, 8 and 16 lanes by the half-row and row mirrors,
// then rows 0+1 and 2+3 by row_bcast:15 and all four by row_bcast:31; the
// total lands in lane 63), then the WAVES waves through LDS (a
// one-wave workgroup reads lane 63 directly: no LDS, no barrier).  The
// item's four partial sums (EV re, im, OD re, im), one per lane of lanes
// 0..3: a single 32-byte write, into pinned host memory for small batches
// (the host finishes each query as its items arrive) or device memory before
// k_z8_reduce.  Each word is one aligned 8-byte store.
template <uint32_t WAVES>
__device__ __forceinline__ void seg_publish(double (*s_red)[4], const float (&ev)[4], const float (&od)[4], uint64_t item,
                                            double* __restrict__ partial) {
    const uint32_t tid = threadIdx.x;
    double wv[4] = {(double)ev[0] + (double)ev[1], (double)ev[2] + (double)ev[3], (double)od[0] + (double)od[1], (double)od[2] + (double)od[3]};
    dpp_add4<0xB1, 0xF>(wv);     // quad_perm [1,0,3,2]
    dpp_add4<0x4E, 0xF>(wv);     // quad_perm [2,3,0,1]
    dpp_add4<0x141, 0xF>(wv);    // row_half_mirror
    dpp_add4<0x140, 0xF>(wv);    // row_mirror
    dpp_add4<0x142, 0xA>(wv);    // row_bcast:15
    dpp_add4<0x143, 0xC>(wv);    // row_bcast:31
    if constexpr (WAVES == 1) {
        double v = 0.0;
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            const uint64_t u = (uint64_t)__double_as_longlong(wv[c]);
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
            if (tid == c) v = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
        }
        if (tid < 4) partial[4 * item + tid] = v;
    } else {
        if ((tid & 63) == 63) {
#pragma unroll
            for (int c = 0; c < 4; ++c) s_red[tid >> 6][c] = wv[c];
        }
        __syncthreads();
        if (tid < 4) {
            double v = 0.0;
            for (uint32_t w = 0; w < WAVES; ++w) v += s_red[w][tid];
            partial[4 * item + tid] = v;
        }
    }
}
// XOR and sum of a value over the 64 lanes of the wave, broadcast to all
// (the DPP order of seg_publish; all 64 lanes must be active).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_src(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false); }
__device__ __forceinline__ uint32_t wave_xor(uint32_t v) {
    v ^= dpp_src<0xB1, 0xF>(v);
    v ^= dpp_src<0x4E, 0xF>(v);
    v ^= dpp_src<0x141, 0xF>(v);
    v ^= dpp_src<0x140, 0xF>(v);
    v ^= dpp_src<0x142, 0xA>(v);
    v ^= dpp_src<0x143, 0xC>(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ __forceinline__ uint32_t wave_add(uint32_t v) {
    v += dpp_src<0xB1, 0xF>(v);
    v += dpp_src<0x4E, 0xF>(v);
    v += dpp_src<0x141, 0xF>(v);
    v += dpp_src<0x140, 0xF>(v);
    v += dpp_src<0x142, 0xA>(v);
    v += dpp_src<0x143, 0xC>(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Short walks (2^MW points, MW <= kSegShortWalk), for latency-bound
// launches: every load that depends only on the engine and the query is
// issued first, the lane seed is formed in registers (no LDS tables, no
// workgroup barrier before the walk), and the table reads and the lane-factor
// byte lookups go out together.
//   Lane coordinates c = (c_low on the L6 lane-low slots, c_hi on the slots
// past the walk).  c_low is the lane's index in its wave; c_hi = lane >> L6
// is the same for the whole wave (L6 = 6 whenever a slot past the lane-low
// ones exists), so the slot terms of c_hi are summed once per wave, one slot
// per lane, and the lane-low terms come from the engine's lane-low table:
//   K  = K0 + low_K(c_low) + 4 (delta . c_low) + sum_{k in c_hi} (lambda_k
//        + 4 delta_k + 4 (Q_k . c_hi below k)) + 4 ((XOR_{k in c_hi} Q_k) . c_low)
//   hs, pio, s: the query's value XOR the lane-low entry XOR the c_hi slots.
// The results equal the slot-by-slot seed bit for bit (integers mod 8 and
// XORs), and the magnitudes multiply in the same slot order.
template <uint32_t MW, uint32_t WAVES>
__device__ __forceinline__ void seg_short(double (*s_red)[4], const SegDev& E, const SegQ& Q, uint64_t item, double* __restrict__ partial) {
    constexpr uint32_t NP = 1u << MW;
    constexpr uint32_t WLOG = WAVES == 4 ? 2 : WAVES == 2 ? 1 : 0;
    const uint32_t tid = threadIdx.x;
    const uint32_t L6 = E.L6, nout = E.nout, I = E.I, nbytes = E.nbytes, Mcap = E.Mcap;
    const uint64_t lane = (item - Q.item0) * (64 * WAVES) + tid;
    const bool live = (lane >> Q.H) == 0;
    const uint32_t lowm = (1u << L6) - 1u;
    const uint32_t cl = tid & lowm;
    const uint64_t chi = L6 == 6 ? (((item - Q.item0) << WLOG) | (uint64_t)__builtin_amdgcn_readfirstlane(tid >> 6)) : 0ull;
    const uint64_t cplace = chi << (L6 + MW);
    // ---- loads, none behind a branch (indices clamped into the arrays)
    const uint32_t* lw = E.lowtab + 6 * cl;
    uint32_t w0 = lw[0], w1 = lw[1], w2 = lw[2], w4 = lw[4], w5 = lw[5];
    const uint32_t j = tid & 63u, k = L6 + MW + j;
    const bool onj = k < nout && ((chi >> j) & 1u);        // lane j holds slot k, and c_hi has it
    const uint32_t ks = k < nout ? k : 0u;
    uint64_t vq0 = E.oq[ks], vn0 = E.on[ks];
    uint32_t vx0 = E.ox[ks], vp0 = E.opin[ks], vl0 = E.olam[ks];
    const uint32_t wj = MW + j;
    const uint32_t mj = wj < Mcap ? 4 * wj + 2 * ((uint32_t)(chi >> j) & 1u) : 0u;
    float rh0 = E.rmid[mj], rl0 = E.rmid[mj + 1];
    uint64_t oqw[MW > 0 ? MW : 1];
#pragma unroll
    for (uint32_t w = 0; w < MW; ++w) oqw[w] = cview(E.oq)[L6 + w];
    const PQ_CONST uint32_t* pt = cview(E.pts) + 4 * (NP - 1u);
    uint32_t rec[4 * NP];
#pragma unroll
    for (uint32_t i = 0; i < 4 * NP; ++i) rec[i] = pt[i];
    const PQ_CONST uint32_t* pc = cview(E.pcls) + 5 * MW;
    const uint32_t c1 = pc[1], c2 = pc[2], c3 = pc[3];
    // every load above is in flight; from here on they are used (the empty
    // asm statements keep the compiler from sinking a load into the
    // branch that first uses it, which would cost one memory latency each)
    asm volatile("" : "+v"(w0), "+v"(w1), "+v"(w2), "+v"(w4), "+v"(w5));
    asm volatile("" : "+v"(vq0), "+v"(vn0), "+v"(vx0), "+v"(vp0), "+v"(vl0), "+v"(rh0), "+v"(rl0));
    if constexpr (NP <= 8) {
#pragma unroll
        for (uint32_t i = 0; i < 4 * NP; ++i) asm volatile("" ::"s"(rec[i]));
    }
    asm volatile("" ::"s"(c1), "s"(c2), "s"(c3));
    // ---- the wave's slots past the walk (one slot per lane), summed over the wave
    const uint64_t vq = onj ? vq0 : 0ull, vn = onj ? vn0 : 0ull;
    const uint32_t tK = onj ? vl0 + (((uint32_t)(Q.delta >> ks) & 1u) << 2) + (((uint32_t)__popcll(vq0 & cplace & ((1ull << ks) - 1ull)) & 1u) << 2) : 0u;
    const uint32_t RK = wave_add(tK), Rhs = wave_xor(onj ? vx0 : 0u), Rpio = wave_xor(onj ? vp0 : 0u);
    const uint32_t Rs0 = wave_xor((uint32_t)vn), Rs1 = wave_xor((uint32_t)(vn >> 32)), Rq = wave_xor((uint32_t)vq & lowm);
    // magnitudes of the walkable slots past the walk, in slot order
    const float rh = wj < Mcap ? rh0 : 1.f, rl = wj < Mcap ? rl0 : 0.f;
    float fh = 1.f, fl = 0.f;
#pragma unroll
    for (uint32_t w = MW; w < kSegMaxMid; ++w) {
        if (w < Mcap) {
            float ph, pl;
            df_mul(fh, fl, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rh), (int)(w - MW))),
                   __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rl), (int)(w - MW))), ph, pl);
            fh = ph; fl = pl;
        }
    }
    // ---- the lane's seed
    const uint32_t K = Q.K0 + w0 + (((uint32_t)__popc((uint32_t)Q.delta & cl) & 1u) << 2) + RK + (((uint32_t)__popc(Rq & cl) & 1u) << 2);
    const uint32_t hs = Q.hs0 ^ w1 ^ Rhs, pio = Q.pio0 ^ w2 ^ Rpio;
    const uint64_t sv = Q.s0 ^ ((uint64_t)w4 | ((uint64_t)w5 << 32)) ^ ((uint64_t)Rs0 | ((uint64_t)Rs1 << 32));
    const uint64_t c = (uint64_t)cl | cplace;
    uint32_t P = 0;
#pragma unroll
    for (uint32_t w = 0; w < MW; ++w) P |= (uint32_t)(__popcll(oqw[w] & c) & 1) << w;
    const uint32_t dw = (uint32_t)(Q.delta >> L6) & ((1u << MW) - 1u);
    float ev[4] = {0.f, 0.f, 0.f, 0.f}, od[4] = {0.f, 0.f, 0.f, 0.f};
    if (live) {
        // ---- table reads and lane-factor byte lookups together
        const uint32_t base = (pio << I) | hs;
        GFloat4* tab = (GFloat4*)E.table;
        f32x4 Tk[NP];
#pragma unroll
        for (uint32_t q = 0; q < NP; ++q) Tk[q] = tab[base ^ rec[4 * q]];
        float2 lb[8];
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b)
            if (b < nbytes) lb[b] = E.lt[(b << 8) | (uint32_t)((sv >> (8 * b)) & 0xFFu)];
        // a point of class cc (uniform) adds i^(cc>>1) times its term into
        // EVu (cc even) or ODu (cc odd)
        float eu[4] = {0.f, 0.f, 0.f, 0.f}, ou[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (uint32_t q = 0; q < NP; ++q) {
            const uint32_t g = rec[4 * q + 1];
            const uint32_t sg = (uint32_t)((__popc(P & g) ^ __popc(dw & g)) & 1) << 31;
            const float wh = __uint_as_float(rec[4 * q + 2] ^ sg), wl = __uint_as_float(rec[4 * q + 3] ^ sg);
            const uint32_t cc = (q >= c1 ? 1u : 0u) + (q >= c2 ? 1u : 0u) + (q >= c3 ? 1u : 0u);
            float ph, pl, qh, ql;
            df_mul(wh, wl, Tk[q].x, Tk[q].y, ph, pl);
            df_mul(wh, wl, Tk[q].z, Tk[q].w, qh, ql);
            const bool r = cc >= 2;       // times i: (x, y) -> (-y, x)
            const float xh = r ? -qh : ph, xl = r ? -ql : pl, yh = r ? ph : qh, yl = r ? pl : ql;
            if (cc & 1u) { df_add(ou[0], ou[1], xh, xl); df_add(ou[2], ou[3], yh, yl); }
            else { df_add(eu[0], eu[1], xh, xl); df_add(eu[2], eu[3], yh, yl); }
        }
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            if (b < nbytes) {
                float ph, pl;
                df_mul(fh, fl, lb[b].x, lb[b].y, ph, pl);
                fh = ph; fl = pl;
            }
        }
        seg_lane_total(eu, ou, K, fh, fl, ev, od);
    }
    seg_publish<WAVES>(s_red, ev, od, item, partial);
}

// One item (64 WAVES lanes of one query): reload the engine tables into LDS if
// the workgroup's last item was of another engine (eng, uniform), seed every
// lane, walk its segment, and write the item's four partial sums.
template <bool SHORT, uint32_t WAVES>
__device__ __forceinline__ void seg_item(const SegLds& S, const SegDev& E, const SegQ& Q, uint64_t item, int& eng, double* __restrict__ partial) {
    if constexpr (SHORT) {
        switch (Q.M) {
        case 0: seg_short<0, WAVES>(S.red, E, Q, item, partial); break;
        case 1: seg_short<1, WAVES>(S.red, E, Q, item, partial); break;
        case 2: seg_short<2, WAVES>(S.red, E, Q, item, partial); break;
        case 3: seg_short<3, WAVES>(S.red, E, Q, item, partial); break;
        default: seg_short<kSegShortWalk, WAVES>(S.red, E, Q, item, partial); break;
        }
        return;
    }
    constexpr uint32_t NT = 64 * WAVES;
    const uint32_t tid = threadIdx.x;
    float2* s_lt = S.lt;
    uint64_t *s_oq = S.oq, *s_on = S.on;
    uint32_t *s_ox = S.ox, *s_opin = S.opin, *s_olam = S.olam;
    float* s_rmid = S.rmid;
    double (*s_red)[4] = S.red;
    uint32_t* s_low = S.low;
    const SegSlots<LdsU64*, LdsU32*, LdsF32*> T{(LdsU64*)s_oq, (LdsU64*)s_on, (LdsU32*)s_ox, (LdsU32*)s_opin, (LdsU32*)s_olam, (LdsF32*)s_rmid,
                                                 (LdsU32*)s_low};
    // once per engine per workgroup of a persistent launch
    if ((int)Q.eng != eng) {
        for (uint32_t k = tid; k < E.nbytes * 256u; k += NT) s_lt[k] = E.lt[k];
        for (uint32_t k = tid; k < E.nout; k += NT) {
            s_oq[k] = E.oq[k]; s_on[k] = E.on[k];
            s_ox[k] = E.ox[k]; s_opin[k] = E.opin[k]; s_olam[k] = E.olam[k];
        }
        if (tid < 4 * E.Mcap) s_rmid[tid] = E.rmid[tid];
        for (uint32_t k = tid; k < (6u << E.L6); k += NT) s_low[k] = E.lowtab[k];
        __syncthreads();
        eng = (int)Q.eng;
    }
    float ev[4] = {0.f, 0.f, 0.f, 0.f}, od[4] = {0.f, 0.f, 0.f, 0.f};
    const uint64_t lane = (item - Q.item0) * NT + tid;
    if ((lane >> Q.H) == 0) {
        SegLaneDev L;
        seg_seed(E, T, Q, lane, L);
        float fh = L.fh, fl = L.fl;
        seg_lane_factor(s_lt, E.nbytes, L.s, fh, fl);
        const uint32_t M = Q.M;
        const uint32_t P = seg_lane_parities(E, T.oq, M, L.cout);
        const uint32_t dw = (uint32_t)(Q.delta >> E.L6) & ((1u << M) - 1u);
        const uint32_t base = (L.pio << E.I) | L.hs;
        GFloat4* tab = (GFloat4*)E.table;
        const PQ_CONST uint32_t* pt = cview(E.pts) + 4 * ((1u << M) - 1u);
        const PQ_CONST uint32_t* cl = cview(E.pcls) + 5 * M;
        // the segment's sums EVu and ODu (double-float re, im)
        float eu[4] = {0.f, 0.f, 0.f, 0.f}, ou[4] = {0.f, 0.f, 0.f, 0.f};
        {
            float A[4][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) { A[c][0] = 0.f; A[c][1] = 0.f; A[c][2] = 0.f; A[c][3] = 0.f; }
            // points grouped by their engine phase class c: the class is
            // a loop constant, so each group adds into its own accumulator
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint32_t p1 = cl[c + 1];
                uint32_t p = cl[c];
                // kSegGroup points per step, all table loads first
                for (; p + kSegGroup <= p1; p += kSegGroup) {
                    f32x4 Tq[kSegGroup];
                    uint32_t sg[kSegGroup];
                    float fh4[kSegGroup], fl4[kSegGroup];
#pragma unroll
                    for (int k = 0; k < (int)kSegGroup; ++k) {
                        const PQ_CONST uint32_t* e = pt + 4 * (p + k);
                        const uint32_t X = e[0], g = e[1];
                        Tq[k] = tab[base ^ X];
                        // query sign (uniform) and lane sign of the point
                        const uint32_t us = (uint32_t)(__popc(dw & g) & 1) << 31;
                        sg[k] = ((uint32_t)(__popc(P & g) & 1) << 31) ^ us;
                        fh4[k] = __uint_as_float(e[2]); fl4[k] = __uint_as_float(e[3]);
                    }
#pragma unroll
                    for (int k = 0; k < (int)kSegGroup; ++k)
                        seg_point(A[c], __uint_as_float(__float_as_uint(fh4[k]) ^ sg[k]), __uint_as_float(__float_as_uint(fl4[k]) ^ sg[k]), Tq[k]);
                }
                for (; p < p1; ++p) {
                    const PQ_CONST uint32_t* e = pt + 4 * p;
                    const uint32_t X = e[0], g = e[1];
                    const f32x4 Tp = tab[base ^ X];
                    const uint32_t sg = (((uint32_t)(__popc(P & g) & 1)) ^ ((uint32_t)(__popc(dw & g) & 1))) << 31;
                    seg_point(A[c], __uint_as_float(e[2] ^ sg), __uint_as_float(e[3] ^ sg), Tp);
                }
            }
            seg_fold4<0>(A, eu);
            seg_fold4<1>(A, ou);
        }
        // lane total = omega^K_seed F (EVu + omega ODu)
        seg_lane_total(eu, ou, L.K, fh, fl, ev, od);
    }
    seg_publish<WAVES>(s_red, ev, od, item, partial);
}

// SHORT: every query of the launch walks at most 2^kSegShortWalk points
// (kept apart from the long-walk variant, whose register budget it would
// otherwise raise).  Items item_lo .. item_end - 1: a workgroup's first item
// is its own index; later ones come from the queue (counter starts at 0), so
// a batch of at most one item per workgroup takes no queue atomics at all.
// The long-walk variant is held to 7 waves per SIMD (70 VGPRs, no spills):
// its table reads are latency-bound and want every resident wave.
// WAVES (1, 2 or 4) is the workgroup shape: an item is the 64 WAVES lanes of
// one workgroup.  With fewer waves per workgroup the long walk's LDS tables
// (20 KB per workgroup) cap the occupancy at 2 WAVES waves per SIMD.
template <bool SHORT, uint32_t WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(SHORT ? 4 : (WAVES == 4 ? 7 : 2 * WAVES), 8))) void k_seg_walk(const SegDev* __restrict__ engs, const SegQ* __restrict__ qs, const uint32_t* __restrict__ item_q,
                                                  uint64_t item_lo, uint64_t item_end, unsigned long long* __restrict__ counter, double* __restrict__ partial) {
    PQ_SEG_LDS(lds);
    __shared__ unsigned long long s_item;
    uint64_t item = item_lo + blockIdx.x;
    int eng = -1;                // engine whose tables are in LDS (uniform)
    while (item < item_end) {
        const uint32_t qi = cview(item_q)[item];
        const SegQ Q = cload(qs + qi);
        const SegDev E = cload(engs + Q.eng);
        seg_item<SHORT, WAVES>(lds, E, Q, item, eng, partial);
        if (item_end - item_lo <= gridDim.x) break;    // every item was a first item
        __syncthreads();                               // LDS free for the next item
        if (threadIdx.x == 0) s_item = item_lo + gridDim.x + atomicAdd(counter, 1ull);
        __syncthreads();
        item = s_item;
    }
}

// A small batch passed whole as kernel arguments: no staging copy before the
// launch and no dependent loads of query and engine records in the kernel.
constexpr uint32_t kArgQueries = 16;
struct SegArgQ {
    SegQ q;
    SegDev e;                        // the query's engine descriptor
};
struct SegArgs {
    SegArgQ a[kArgQueries];          // the queries of one launch, longest first
};
// One workgroup per (item, query): blockIdx.y is the query and blockIdx.x
// the item within it (queries come longest first, so the grid's x extent is
// the first query's item count; a shorter query leaves the rest of its row
// idle).  The query record and its engine descriptor sit together at an
// offset given by blockIdx.y alone: the kernel's only batch of argument
// loads is issued as the wave starts.
template <bool SHORT, uint32_t WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_seg_walk_args(const SegArgs A, double* __restrict__ partial) {
    PQ_SEG_LDS(lds);
    const SegQ Q = A.a[blockIdx.y].q;
    const SegDev E = A.a[blockIdx.y].e;
    asm volatile("" ::"s"(Q.delta), "s"(Q.s0), "s"(Q.item0), "s"(Q.nitems), "s"(Q.K0), "s"(Q.hs0), "s"(Q.pio0), "s"(Q.M), "s"(Q.H));
    asm volatile("" ::"s"(E.oq), "s"(E.ox), "s"(E.opin), "s"(E.olam), "s"(E.on), "s"(E.table), "s"(E.lt), "s"(E.rmid));
    asm volatile("" ::"s"(E.pts), "s"(E.pcls), "s"(E.lowtab), "s"(E.I), "s"(E.nout), "s"(E.L6), "s"(E.Mcap), "s"(E.nbytes));
    if (blockIdx.x >= Q.nitems) return;          // uniform: the whole workgroup
    int eng = -1;
    seg_item<SHORT, WAVES>(lds, E, Q, Q.item0 + blockIdx.x, eng, partial);
}

// FP64 cross-workgroup reduction (fixed order) and global scaling per query.
// Separate reduction launch (branch walk): one workgroup per query.  Both
// query records (Z8Q, SegQ) carry item0 and nitems in words 2 and 3.
__global__ __launch_bounds__(256) void k_z8_reduce(const uint64_t* __restrict__ qwords, const double* __restrict__ partial, const double2* __restrict__ scale,
                                                   const uint32_t* __restrict__ order, double* __restrict__ out) {
    __shared__ double red[4][4];
    const uint32_t qi = blockIdx.x;
    query_reduce(qwords[8 * qi + 2], qwords[8 * qi + 3], partial, scale[qi], out + 2 * order[qi], red);
}


// Segment split: a lane walks 2^(kin+M) branches and 2^H lanes cover a query
// (H = d-kin-M).  M depends on the engine alone (M = clamp(d - offset, 0,
// max), at most d - kin), so a
// query is decomposed the same way in every batch and its partial sums,
// hence its result, are bitwise the same alone and in any mix.  The offset
// and cap are set by the autotuner (polyq_amd.autotune); the default keeps
// H = 19 (2048 items of 256 lanes per query) until M reaches 6.
static std::atomic<uint32_t> g_mid_offset{27}, g_mid_max{6};
void z8_set_tuning(uint32_t offset, uint32_t max_mid) {
    // M = clamp(d - offset, 0, max_mid) is further capped at d - kin, so any
    // offset up to the largest branch dimension is meaningful
    if (offset > 64 || max_mid > 16) throw std::invalid_argument("walk tuning out of range");
    g_mid_offset = offset; g_mid_max = max_mid;
}
void z8_get_tuning(uint32_t* offset, uint32_t* max_mid) { *offset = g_mid_offset; *max_mid = g_mid_max; }
uint32_t z8_auto_mid_bits(uint32_t d, uint32_t kin) {
    const uint32_t off = g_mid_offset, mx = g_mid_max;
    if (d <= off) return 0;
    return std::min<uint32_t>(std::min<uint32_t>(d - off, mx), d > kin ? d - kin : 0);
}
// Segment walk: a lane walks 2^M outer points (each a 2^I-branch table
// entry) with M = clamp(nout - offset, 0, max), capped by the walkable slots
// of the engine; again a function of the engine alone.
static std::atomic<uint32_t> g_seg_offset{12}, g_seg_max{6};
void seg_set_tuning(uint32_t offset, uint32_t max_mid) {
    if (max_mid > kSegMaxMid) throw std::invalid_argument("segment walk tuning out of range");
    g_seg_offset = offset; g_seg_max = max_mid;
}
void seg_get_tuning(uint32_t* offset, uint32_t* max_mid) { *offset = g_seg_offset; *max_mid = g_seg_max; }
// Workgroup shape of the segment walk: waves per workgroup (1, 2 or 4), so an
// item is 64, 128 or 256 lanes.  Like the walk length it is global, so a
// query's items and partial sums do not depend on the batch it is in.
static std::atomic<uint32_t> g_seg_waves{4};
void seg_set_waves(uint32_t waves) {
    if (waves != 1 && waves != 2 && waves != 4) throw std::invalid_argument("segment walk workgroups have 1, 2 or 4 waves");
    g_seg_waves = waves;
}
uint32_t seg_get_waves() { return g_seg_waves; }
uint32_t dev_mid_bits(const DevEngine& D, int walk) {
    if (walk == kWalkSeg) return seg_mid_bits(*D.seg, g_seg_offset, g_seg_max);
    return z8_auto_mid_bits(D.z8->d, D.z8->kin);
}

namespace {
constexpr int kVariants = 4;   // branch walk: bit 1 inner pivot lookups, bit 0 imaginary class pattern
int variant_of(const Z8Image& Z) { return (Z.kin == kInnerPiv ? 2 : 0) | (Z.imag0 ? 1 : 0); }
// What one batch uploads: engine descriptors, 64-byte query records, scales
// and the output order, plus the item ranges of the launches.
struct BatchPlan {
    int walk = kWalkSeg;
    std::vector<uint8_t> engs, qs;
    std::vector<uint32_t> item_q;      // segment walk: query of each item
    std::vector<double2> scale;
    std::vector<uint32_t> order;
    uint64_t total_items = 0, branches = 0;
    uint32_t waves = 4;                 // workgroup waves (segment walk; the branch walk's are 4)
    uint32_t q_lo[kVariants] = {0, 0, 0, 0}, q_hi[kVariants] = {0, 0, 0, 0};
    uint64_t it_lo[kVariants] = {0, 0, 0, 0}, it_hi[kVariants] = {0, 0, 0, 0};
    uint32_t nq() const { return (uint32_t)(qs.size() / 64); }
};
template <class T> void push_rec(std::vector<uint8_t>& v, const T& r) {
    const size_t o = v.size();
    v.resize(o + sizeof(T));
    std::memcpy(v.data() + o, &r, sizeof(T));
}
// The launch order of a batch's jobs: by key (launch variant, then heavier
// items first), then the jobs of one engine together, then job index.  The
// keys are computed once per job, and the order is total, so the plan does
// not depend on the sort algorithm.
struct PlanKey {
    uint64_t k;
    const DevEngine* eng;
    uint32_t ji;
};
void sort_plan(std::vector<PlanKey>& v) {
    std::sort(v.begin(), v.end(), [](const PlanKey& a, const PlanKey& b) {
        if (a.k != b.k) return a.k < b.k;
        if (a.eng != b.eng) return std::less<const DevEngine*>()(a.eng, b.eng);
        return a.ji < b.ji;
    });
}
// Index of engine `e` among the batch's descriptors, appending its
// descriptor on first sight (jobs of one engine are mostly adjacent: the
// previous job's engine is tried first).
struct EngineIds {
    std::unordered_map<const DevEngine*, uint32_t> ids;
    const DevEngine* last = nullptr;
    uint32_t last_id = 0;
    template <class Rec>
    uint32_t get(const DevEngine* e, const Rec& rec, std::vector<uint8_t>& engs) {
        if (e == last) return last_id;
        auto it = ids.find(e);
        uint32_t id;
        if (it == ids.end()) { id = (uint32_t)ids.size(); ids.emplace(e, id); push_rec(engs, rec); }
        else id = it->second;
        last = e; last_id = id;
        return id;
    }
};
BatchPlan plan_branch(const std::vector<Z8Job>& jobs) {
    BatchPlan B;
    B.walk = kWalkBranch;
    EngineIds seen;
    // grouped by kernel variant; heavy items first; jobs of one engine together
    std::vector<PlanKey> order(jobs.size());
    for (uint32_t ji = 0; ji < (uint32_t)jobs.size(); ++ji)
        order[ji] = {((uint64_t)variant_of(*jobs[ji].eng->z8) << 32) | (uint64_t)(0xFFFFFFFFu - jobs[ji].mid_bits), jobs[ji].eng, ji};
    sort_plan(order);
    B.qs.reserve(64 * jobs.size());
    int cur = -1;
    for (const PlanKey& pk : order) {
        const uint32_t ji = pk.ji;
        const Z8Job& J = jobs[ji];
        const Z8Image& Z = *J.eng->z8;
        const int v = variant_of(Z);
        if (v != cur) {
            B.q_lo[v] = B.nq(); B.it_lo[v] = B.total_items;
            cur = v;
        }
        const uint32_t e = seen.get(J.eng, J.eng->z, B.engs);
        Z8Q q{};
        q.delta = J.q.delta; q.s0 = J.q.s0; q.eng = e; q.K0 = J.q.K0; q.pio0 = J.q.pio0;
        const uint32_t nout = Z.nslots - Z.kin;
        q.M = std::min(J.mid_bits, nout);
        q.H = nout - q.M;
        if (q.H > 40) throw std::invalid_argument("branch dimension too large for one batch");
        q.item0 = B.total_items;
        q.nitems = ((1ull << q.H) + 255) / 256;
        B.total_items += q.nitems;
        B.branches += 1ull << Z.d;
        push_rec(B.qs, q);
        B.scale.push_back(make_double2(J.scale.real(), J.scale.imag()));
        B.order.push_back(ji);
        B.q_hi[v] = B.nq(); B.it_hi[v] = B.total_items;
    }
    return B;
}
BatchPlan plan_seg(const std::vector<Z8Job>& jobs) {
    BatchPlan B;
    B.walk = kWalkSeg;
    B.waves = g_seg_waves;
    const uint64_t lanes = 64ull * B.waves;
    EngineIds seen;
    auto walk_bits = [&](const Z8Job& J) {
        const SegImage& G = *J.eng->seg;
        return std::min(J.mid_bits, std::min(G.Mcap, G.nout - G.L6));
    };
    // Queries with long lane walks first (launch variant 0), short walks
    // after (variant 1); within each, heavy items first and jobs of one
    // engine together (the LDS tables are reloaded only when a workgroup
    // moves to another engine).
    std::vector<PlanKey> order(jobs.size());
    for (uint32_t ji = 0; ji < (uint32_t)jobs.size(); ++ji) {
        const uint32_t mb = walk_bits(jobs[ji]);
        order[ji] = {((uint64_t)(mb <= kSegShortWalk ? 1 : 0) << 32) | (uint64_t)(0xFFFFFFFFu - (jobs[ji].eng->seg->nout - mb)), jobs[ji].eng, ji};
    }
    sort_plan(order);
    B.qs.reserve(64 * jobs.size());
    int cur = -1;
    for (const PlanKey& pk : order) {
        const uint32_t ji = pk.ji;
        const Z8Job& J = jobs[ji];
        const SegImage& G = *J.eng->seg;
        const uint32_t e = seen.get(J.eng, J.eng->sg, B.engs);
        SegQ q{};
        q.delta = J.sq.delta; q.s0 = J.sq.s0; q.eng = e; q.K0 = J.sq.K0; q.hs0 = J.sq.hs0; q.pio0 = J.sq.pio0;
        q.M = walk_bits(J);
        q.H = G.nout - q.M;
        if (q.H > 40) throw std::invalid_argument("branch dimension too large for one batch");
        const int v = q.M <= kSegShortWalk ? 1 : 0;
        if (v != cur) { B.q_lo[v] = B.nq(); B.it_lo[v] = B.total_items; cur = v; }
        q.item0 = B.total_items;
        q.nitems = ((1ull << q.H) + lanes - 1) / lanes;
        B.total_items += q.nitems;
        B.item_q.insert(B.item_q.end(), q.nitems, B.nq());
        B.branches += 1ull << G.d;
        push_rec(B.qs, q);
        B.scale.push_back(make_double2(J.scale.real(), J.scale.imag()));
        B.order.push_back(ji);
        B.q_hi[v] = B.nq(); B.it_hi[v] = B.total_items;
    }
    return B;
}
BatchPlan plan_batch(const std::vector<Z8Job>& jobs, int walk) {
    if (walk == kWalkSeg) return plan_seg(jobs);
    if (walk == kWalkBranch) return plan_branch(jobs);
    throw std::invalid_argument("unknown walk form");
}
// Resident workgroups of `block` threads per CU of a kernel, asked of the
// runtime once per kernel.
template <class K>
int blocks_per_cu(K kernel, int block) {
    static std::mutex mu;
    static std::vector<std::pair<const void*, int>> known;
    std::lock_guard<std::mutex> lk(mu);
    for (const auto& e : known) if (e.first == (const void*)kernel) return e.second;
    int per_cu = 0;
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
    known.emplace_back((const void*)kernel, per_cu);
    return per_cu;
}
template <class K>
uint32_t grid_for(Ctx& C, K kernel, uint64_t items, int block = 256) {
    const int per_cu = blocks_per_cu(kernel, block);
    const uint64_t g = (uint64_t)std::max(per_cu, 1) * (uint64_t)C.cus;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, items));
}
template <bool HASG, bool IMAG>
uint32_t launch_v(Ctx& C, hipStream_t st, const BatchPlan& B, int v, const Z8Dev* de, const Z8Q* dq, unsigned long long* ctr, double* part) {
    const uint64_t items = B.it_hi[v] - B.it_lo[v];
    const uint32_t grid = grid_for(C, k_z8_walk<HASG, IMAG>, items);
    hipLaunchKernelGGL((k_z8_walk<HASG, IMAG>), dim3(grid), dim3(256), 0, st, de, dq, B.q_lo[v], B.q_hi[v], B.it_lo[v], B.it_hi[v], ctr + v, part);
    HIPCHECK(hipGetLastError());
    return grid;
}
// Launch the walk of a staged batch (one persistent grid per branch-walk
// variant present, or one segment-walk grid); returns the summed grid.
// The work-queue counters come first in the upload and are zero in the host
// copy, so an eager batch's descriptor copy also clears them (a graph uploads
// its descriptors once and clears the counters with k_clear_counters).
struct StageLayout {
    size_t oe, oq, os, oo, oi, tot;
};
StageLayout stage_layout(const BatchPlan& B) {
    StageLayout L;
    const size_t be = B.engs.size(), bq = B.qs.size(), bs = B.scale.size() * sizeof(double2), bo = B.order.size() * 4;
    L.oe = 256;   // 4 u64 queue counters at 0
    L.oq = L.oe + ((be + 255) & ~size_t(255));
    L.os = L.oq + ((bq + 255) & ~size_t(255));
    L.oo = L.os + ((bs + 255) & ~size_t(255));
    L.oi = L.oo + ((bo + 255) & ~size_t(255));
    L.tot = L.oi + B.item_q.size() * 4;
    return L;
}
void stage_copy(const BatchPlan& B, const StageLayout& L, uint8_t* h) {
    std::memset(h, 0, L.oe);
    std::memcpy(h + L.oe, B.engs.data(), B.engs.size());
    std::memcpy(h + L.oq, B.qs.data(), B.qs.size());
    std::memcpy(h + L.os, B.scale.data(), B.scale.size() * sizeof(double2));
    std::memcpy(h + L.oo, B.order.data(), B.order.size() * 4);
    std::memcpy(h + L.oi, B.item_q.data(), B.item_q.size() * 4);
}
// Variant v of a segment batch with WAVES-wave workgroups.
template <uint32_t WAVES>
uint32_t launch_seg(Ctx& C, hipStream_t st, const BatchPlan& B, int v, const void* de, const void* dq, const void* item_q,
                    unsigned long long* ctr, double* part, bool one_pass) {
    const uint64_t items = B.it_hi[v] - B.it_lo[v];
    auto kern = v ? k_seg_walk<true, WAVES> : k_seg_walk<false, WAVES>;
    const uint32_t grid = one_pass ? (uint32_t)items : grid_for(C, kern, items, 64 * WAVES);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * WAVES), 0, st, (const SegDev*)de, (const SegQ*)dq, (const uint32_t*)item_q,
                       B.it_lo[v], B.it_hi[v], ctr + v, part);
    HIPCHECK(hipGetLastError());
    return grid;
}
// Launches the walk kernels of a batch; item partials go to part (device
// memory, or pinned host memory seen through its device address).  With
// one_pass every item gets a workgroup of its own (the hardware dispatcher
// hands out the items in order; no queue, and no workgroup waits for its
// host write before taking another item); otherwise the grid is persistent.
uint32_t launch_walks(Ctx& C, hipStream_t st, const BatchPlan& B, uint8_t* dbase, const StageLayout& L, double* part, bool one_pass) {
    unsigned long long* ctr = (unsigned long long*)dbase;
    const void* de = dbase + L.oe;
    const void* dq = dbase + L.oq;
    if (B.walk == kWalkSeg) {
        uint32_t g = 0;
        for (int v = 0; v < 2; ++v) {
            const uint64_t items = B.it_hi[v] - B.it_lo[v];
            if (items == 0) continue;
            switch (B.waves) {
            case 1: g += launch_seg<1>(C, st, B, v, de, dq, dbase + L.oi, ctr, part, one_pass); break;
            case 2: g += launch_seg<2>(C, st, B, v, de, dq, dbase + L.oi, ctr, part, one_pass); break;
            default: g += launch_seg<4>(C, st, B, v, de, dq, dbase + L.oi, ctr, part, one_pass); break;
            }
        }
        return g;
    }
    uint32_t g = 0;
    const Z8Dev* ze = (const Z8Dev*)de;
    const Z8Q* zq = (const Z8Q*)dq;
    for (int v = 0; v < kVariants; ++v) {
        if (B.it_hi[v] == B.it_lo[v]) continue;
        const bool hasg = v & 2, imag = v & 1;
        if (hasg) g += imag ? launch_v<true, true>(C, st, B, v, ze, zq, ctr, part) : launch_v<true, false>(C, st, B, v, ze, zq, ctr, part);
        else g += imag ? launch_v<false, true>(C, st, B, v, ze, zq, ctr, part) : launch_v<false, false>(C, st, B, v, ze, zq, ctr, part);
    }
    return g;
}

// A host-finished segment batch small enough to travel as kernel arguments.
bool fits_args(const BatchPlan& B) {
    return B.walk == kWalkSeg && B.nq() <= kArgQueries;
}
// The arguments of variant v of a kernel-argument batch and its grid:
// (items of its first, longest query) x (its queries).
void make_args(const BatchPlan& B, int v, SegArgs& A, uint32_t& gx, uint32_t& gy) {
    std::memset(&A, 0, sizeof(A));
    gy = B.q_hi[v] - B.q_lo[v];
    gx = 0;
    for (uint32_t k = 0; k < gy; ++k) {
        SegArgQ& r = A.a[k];
        std::memcpy(&r.q, B.qs.data() + 64 * (size_t)(B.q_lo[v] + k), sizeof(SegQ));
        std::memcpy(&r.e, B.engs.data() + sizeof(SegDev) * (size_t)r.q.eng, sizeof(SegDev));
        gx = std::max<uint32_t>(gx, (uint32_t)r.q.nitems);
    }
}
template <uint32_t WAVES>
void launch_args_v(hipStream_t st, int v, const SegArgs& A, uint32_t gx, uint32_t gy, double* part, hipEvent_t e0, hipEvent_t e1) {
    auto kern = v ? k_seg_walk_args<true, WAVES> : k_seg_walk_args<false, WAVES>;
    if (e0 || e1)
        hipExtLaunchKernelGGL(kern, dim3(gx, gy), dim3(64 * WAVES), 0, st, e0, e1, 0, A, part);
    else
        hipLaunchKernelGGL(kern, dim3(gx, gy), dim3(64 * WAVES), 0, st, A, part);
    HIPCHECK(hipGetLastError());
}
// With t0/t1 set, the two events take their timestamps from the walk
// kernels' own dispatches (start of the first, end of the last), so the
// interval holds no marker packets and reads what a kernel trace reads.
uint32_t launch_args(hipStream_t st, const BatchPlan& B, double* part, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr) {
    SegArgs A[2];
    uint32_t gx[2], gy[2];
    int last = -1, first = -1;
    for (int v = 0; v < 2; ++v) {
        make_args(B, v, A[v], gx[v], gy[v]);
        if (gy[v] == 0) continue;
        if (first < 0) first = v;
        last = v;
    }
    uint32_t g = 0;
    for (int v = 0; v < 2; ++v) {
        if (gy[v] == 0) continue;
        hipEvent_t e0 = v == first ? t0 : nullptr, e1 = v == last ? t1 : nullptr;
        switch (B.waves) {
        case 1: launch_args_v<1>(st, v, A[v], gx[v], gy[v], part, e0, e1); break;
        case 2: launch_args_v<2>(st, v, A[v], gx[v], gy[v], part, e0, e1); break;
        default: launch_args_v<4>(st, v, A[v], gx[v], gy[v], part, e0, e1); break;
        }
        g += gx[v] * gy[v];
    }
    return g;
}

// The per-query data the host needs to finish a batch itself.
struct QueryTail {
    std::vector<uint64_t> item0, nitems;
    std::vector<double2> scale;
    std::vector<uint32_t> order;
    void set(const BatchPlan& B) {
        const uint32_t nq = B.nq();
        item0.resize(nq); nitems.resize(nq);
        for (uint32_t i = 0; i < nq; ++i) {
            std::memcpy(&item0[i], B.qs.data() + 64 * (size_t)i + 16, 8);
            std::memcpy(&nitems[i], B.qs.data() + 64 * (size_t)i + 24, 8);
        }
        scale = B.scale; order = B.order;
    }
};

// The device order of query_reduce restated on the host, bit for bit (IEEE
// double additions and products in the same association; this file is built
// without floating-point contraction): thread t of 256 adds items t, t + 256,
// ... starting from 0; each 64-lane wave folds by the xor butterfly (lane 0
// ends with lane 0 + lane 32 first, then with the lane-16 sum, ...); then
// (w0 + w1) + (w2 + w3), EV + omega OD and the scale.
typedef double HostV4 __attribute__((vector_size(32)));
void host_query_reduce(const double* part, uint64_t nitems, double2 sc, double* dst) {
    // the four components (EV re, im, OD re, im) as one 4-wide vector: the
    // same IEEE additions, component by component
    HostV4 w[4];
    const HostV4 z = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t wv = 0; wv < 4; ++wv) {
        const uint64_t t0 = 64 * (uint64_t)wv;
        if (t0 >= nitems) { w[wv] = z; continue; }
        // lanes at or beyond n hold no item: +0.0, and so does every sum of
        // them, so only lanes below n are computed (x + 0.0 is still done:
        // it turns -0.0 into +0.0 exactly as the device does)
        uint32_t n = (uint32_t)std::min<uint64_t>(64, nitems - t0);
        HostV4 L[64];
        for (uint32_t i = 0; i < n; ++i) {
            HostV4 v = z;
            for (uint64_t k = t0 + i; k < nitems; k += 256) { HostV4 x; std::memcpy(&x, part + 4 * k, 32); v = v + x; }
            L[i] = v;
        }
        for (uint32_t off = 32; off > 0; off >>= 1) {
            const uint32_t m = std::min(off, n);
            for (uint32_t i = 0; i < m; ++i) L[i] = L[i] + (i + off < n ? L[i + off] : z);
            n = m;
        }
        w[wv] = L[0];
    }
    const HostV4 t = (w[0] + w[1]) + (w[2] + w[3]);
    const double h = 0.70710678118654752440;
    // EV + omega * OD
    const double re = t[0] + h * (t[2] - t[3]);
    const double im = t[1] + h * (t[2] + t[3]);
    dst[0] = sc.x * re - sc.y * im;
    dst[1] = sc.x * im + sc.y * re;
}

// A pinned host word the device has not written yet holds this pattern (a
// signalling NaN, which GPU arithmetic never produces); each partial is one
// aligned 8-byte store, so a word holds either the pattern or the value.
constexpr uint64_t kPending = 0x7FF4C0FFEE5107EDull;

// Waits for the item partials of each query in turn (the walk may still be
// running later items) and finishes the query on the host.  A launch that
// failed leaves words pending: with an event (`done`) the event is polled now
// and then; without one the stream is asked after every 2 s of waiting (a
// batch small enough to finish here takes milliseconds at most), so the
// failure is reported instead of waited on.
void host_finish(const QueryTail& T, const double* part, hipEvent_t done, hipStream_t stream, double* res) {
    const volatile uint64_t* w = (const volatile uint64_t*)part;
    bool finished = false;   // the stream has completed: every store landed
    uint64_t spin = 0;
    auto t_ask = std::chrono::steady_clock::now();
    for (size_t q = 0; q < T.item0.size(); ++q) {
        const uint64_t lo = 4 * T.item0[q], hi = 4 * (T.item0[q] + T.nitems[q]);
        for (uint64_t i = lo; i < hi;) {
            if (w[i] != kPending) { ++i; continue; }
            if (finished) throw std::runtime_error("walk ended without writing every item partial");
            // Yield now and then: the runtime's own threads (completion
            // handling) may share this core, and a spin that never lets them
            // run gets the calling thread descheduled for long stretches
            // (every host phase of a step then runs several times slower).
            if ((spin & 63) == 63) std::this_thread::yield();
            if (done) {
                if ((++spin & 255) == 0) {
                    const hipError_t e = hipEventQuery(done);
                    if (e != hipErrorNotReady) { HIPCHECK(e); finished = true; }
                }
            } else if ((++spin & 4095) == 0 && std::chrono::steady_clock::now() - t_ask > std::chrono::seconds(2)) {
                const hipError_t e = hipStreamQuery(stream);
                if (e != hipErrorNotReady) { HIPCHECK(e); finished = true; }
                t_ask = std::chrono::steady_clock::now();
            }
        }
        std::atomic_thread_fence(std::memory_order_acquire);
        host_query_reduce(part + 4 * T.item0[q], T.nitems[q], T.scale[q], res + 2 * (size_t)T.order[q]);
    }
}
}  // namespace

// Small batches of either walk are finished on the host: the walk writes its
// item partials straight into pinned host memory and the host reduces each
// query as soon as its items have landed (no cross-workgroup atomics on the
// device, no reduction launch, no result copy).  Larger batches reduce on the
// device (k_z8_reduce) and copy the results.  The reduction order, hence
// every result bit, is the same either way.
constexpr uint64_t kHostFinishItems = 4096;
bool host_finishes(const BatchPlan& B) { return B.total_items <= kHostFinishItems; }

// One in-flight batch: its own stream, events and buffers, so a second batch
// can be queued while this one still runs (continuous batching).  Timing
// events (e0..e3) are recorded only for a timed batch: each one is a queue
// barrier that costs several microseconds of a short batch's latency.
struct Z8Slot {
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr, done = nullptr;
    DBuf d_in, d_part, d_res;
    HBuf h_in, h_res, h_part;
    QueryTail tail;
    uint32_t nq = 0, grid = 0;
    uint64_t items = 0, branches = 0;
    bool pending = false, timed = false, on_host = false;
    bool has_done = false;           // `done` was recorded after the batch
    Z8Slot() {
        HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));   // default priority
        HIPCHECK(hipEventCreate(&e0)); HIPCHECK(hipEventCreate(&e1)); HIPCHECK(hipEventCreate(&e2)); HIPCHECK(hipEventCreate(&e3));
        HIPCHECK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    }
    ~Z8Slot() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipEvent_t e : {e0, e1, e2, e3, done}) if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        for (void* p : {d_in.p, d_part.p, d_res.p}) if (p) (void)hipFree(p);
        for (void* p : {h_in.p, h_res.p, h_part.p}) if (p) (void)hipHostFree(p);
    }
};

std::shared_ptr<Z8Slot> dev_slot_create() { (void)ctx(); return std::make_shared<Z8Slot>(); }

void dev_z8_launch(Z8Slot& S, const std::vector<Z8Job>& jobs, int walk, bool timed) {
    if (S.pending) throw std::logic_error("batch slot already has a batch in flight");
    Ctx& C = ctx();
    BatchPlan B = plan_batch(jobs, walk);
    const uint32_t nq = B.nq();
    S.nq = nq; S.items = B.total_items; S.branches = B.branches; S.timed = timed; S.on_host = host_finishes(B);
    if (nq == 0) return;
    double* hres = (double*)S.h_res.get(nq * 16);
    if (S.on_host && fits_args(B)) {
        // everything in the kernel arguments: no upload at all
        uint64_t* hp = (uint64_t*)S.h_part.get(B.total_items * 32);
        for (uint64_t i = 0; i < 4 * B.total_items; ++i) hp[i] = kPending;
        S.tail.set(B);
        if (timed) HIPCHECK(hipEventRecord(S.e0, S.stream));
        S.grid = launch_args(S.stream, B, (double*)S.h_part.dev, timed ? S.e1 : nullptr, timed ? S.e2 : nullptr);
        // An untimed batch records no completion event: the host waits for
        // its partials (an event record costs about 1 us of host time per
        // batch and wakes the runtime's completion handling)
        if (timed) { HIPCHECK(hipEventRecord(S.e3, S.stream)); HIPCHECK(hipEventRecord(S.done, S.stream)); }
        S.has_done = timed;
        S.pending = true;
        return;
    }
    // one pinned staging buffer: counters | engines | queries | scales | order | item queries
    const StageLayout L = stage_layout(B);
    uint8_t* h = (uint8_t*)S.h_in.get(L.tot);
    stage_copy(B, L, h);
    uint8_t* d = (uint8_t*)S.d_in.get(L.tot);
    if (timed) HIPCHECK(hipEventRecord(S.e0, S.stream));
    HIPCHECK(hipMemcpyAsync(d, h, L.tot, hipMemcpyHostToDevice, S.stream));
    if (timed) HIPCHECK(hipEventRecord(S.e1, S.stream));
    if (S.on_host) {
        uint64_t* hp = (uint64_t*)S.h_part.get(B.total_items * 32);
        for (uint64_t i = 0; i < 4 * B.total_items; ++i) hp[i] = kPending;
        double* dp = (double*)S.h_part.dev;
        S.tail.set(B);
        S.grid = launch_walks(C, S.stream, B, d, L, dp, true);
        if (timed) HIPCHECK(hipEventRecord(S.e2, S.stream));
    } else {
        double* part = (double*)S.d_part.get(std::max<uint64_t>(B.total_items, 1) * 32);
        double* res = (double*)S.d_res.get(nq * 16);
        S.grid = launch_walks(C, S.stream, B, d, L, part, false);
        if (timed) HIPCHECK(hipEventRecord(S.e2, S.stream));
        hipLaunchKernelGGL(k_z8_reduce, dim3(nq), dim3(256), 0, S.stream, (const uint64_t*)(d + L.oq), part, (const double2*)(d + L.os), (const uint32_t*)(d + L.oo), res);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(hres, res, nq * 16, hipMemcpyDeviceToHost, S.stream));
    }
    if (timed) HIPCHECK(hipEventRecord(S.e3, S.stream));
    HIPCHECK(hipEventRecord(S.done, S.stream));
    S.has_done = true;
    S.pending = true;
}

bool dev_z8_ready(Z8Slot& S) {
    if (!S.pending) return true;
    if (S.on_host && !S.has_done) {
        // finished on the host: ready once every partial word has landed
        // (or once the stream is idle without them: the wait reports that)
        const volatile uint64_t* w = (const volatile uint64_t*)S.h_part.p;
        bool all = true;
        for (uint64_t i = 0; i < 4 * S.items && all; ++i) all = w[i] != kPending;
        if (all) return true;
    }
    hipError_t e = S.has_done ? hipEventQuery(S.done) : hipStreamQuery(S.stream);
    if (e == hipErrorNotReady) return false;
    HIPCHECK(e);
    return true;
}

void dev_z8_wait(Z8Slot& S, std::vector<std::complex<double>>& out, Z8Timing* timing) {
    out.assign(S.nq, {0, 0});
    if (!S.pending) return;
    double* hres = (double*)S.h_res.p;
    if (S.on_host) {
        try {
            host_finish(S.tail, (const double*)S.h_part.p, S.has_done ? S.done : nullptr, S.stream, hres);
        } catch (...) {
            S.pending = false;
            (void)hipStreamSynchronize(S.stream);
            throw;
        }
        if (S.has_done && S.timed) HIPCHECK(hipEventSynchronize(S.done));
    } else {
        HIPCHECK(hipEventSynchronize(S.done));
    }
    S.pending = false;
    for (uint32_t i = 0; i < S.nq; ++i) out[i] = {hres[2 * i], hres[2 * i + 1]};
    if (timing) {
        timing->branches = S.branches; timing->items = S.items; timing->grid = S.grid;
        timing->kernel_ms = timing->total_ms = std::nan("");
        if (S.timed) {
            float k, t;
            HIPCHECK(hipEventElapsedTime(&k, S.e1, S.e2)); HIPCHECK(hipEventElapsedTime(&t, S.e0, S.e3));
            timing->kernel_ms = k; timing->total_ms = t;
        }
    }
}

// Whether eager batches record timing events (pq_set_timing).
static std::atomic<bool> g_timed{false};
void dev_set_timing(bool on) { g_timed = on; }
bool dev_get_timing() { return g_timed; }

void dev_z8_batch(const std::vector<Z8Job>& jobs, std::vector<std::complex<double>>& out, Z8Timing* timing, int walk) {
    out.assign(jobs.size(), {0, 0});
    if (jobs.empty()) return;
    Ctx& C = ctx();
    static std::shared_ptr<Z8Slot> slot;
    std::lock_guard<std::mutex> lk(C.mu);
    if (!slot) slot = std::make_shared<Z8Slot>();
    dev_z8_launch(*slot, jobs, walk, g_timed);
    dev_z8_wait(*slot, out, timing);
}

// ------------------------------------------------- allocation-free batch ----
// A batch whose device working memory (staging: queue counters, engine
// descriptors, query records, scales, output order, item -> query table; the
// item partials; the query results) lives in one file-scope device array, so
// the call allocates no device or pinned host memory.  Each device holds its
// own copy of the array; the host reaches it only through the address
// hipGetSymbolAddress returns for that device (the array's name is a host-side
// stand-in, not a device address).  Calls are serialised on the context's
// lock, so two calls never share the array at the same time; each call
// overwrites what the previous one left (the staging copy rewrites the
// counters with zeros, every item writes all its partials, every query its
// result).
constexpr size_t kNoAllocScratch = 8u << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratch];

namespace {
constexpr int kMaxDevices = 64;
uint8_t* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;
uint8_t* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));   // the current device's copy
        g_scratch_ptr[device] = (uint8_t*)p;
    }
    return g_scratch_ptr[device];
}
}  // namespace

size_t dev_noalloc_scratch_bytes() { return kNoAllocScratch; }

void dev_z8_batch_noalloc(const std::vector<Z8Job>& jobs, std::vector<std::complex<double>>& out, int walk) {
    out.assign(jobs.size(), {0, 0});
    if (jobs.empty()) return;
    Ctx& C = ctx();
    const BatchPlan B = plan_batch(jobs, walk);
    const uint32_t nq = B.nq();
    if (nq == 0) return;
    const StageLayout L = stage_layout(B);
    const size_t o_part = (L.tot + 255) & ~size_t(255);
    const size_t o_res = (o_part + B.total_items * 32 + 255) & ~size_t(255);
    const size_t need = o_res + (size_t)nq * 16;
    if (need > kNoAllocScratch)
        throw std::invalid_argument("batch needs " + std::to_string(need) + " bytes of device scratch; the allocation-free entry point has " +
                                    std::to_string(kNoAllocScratch));
    // host staging: the batch's own records (ordinary heap memory, no
    // device or pinned allocation)
    thread_local std::vector<uint8_t> h;
    thread_local std::vector<double> hres;
    h.resize(L.tot);
    stage_copy(B, L, h.data());
    hres.resize(2 * (size_t)nq);
    std::lock_guard<std::mutex> lk(C.mu);
    uint8_t* d = scratch_on(C.device);
    double* part = (double*)(d + o_part);
    double* res = (double*)(d + o_res);
    HIPCHECK(hipMemcpyAsync(d, h.data(), L.tot, hipMemcpyHostToDevice, C.stream));
    launch_walks(C, C.stream, B, d, L, part, false);
    hipLaunchKernelGGL(k_z8_reduce, dim3(nq), dim3(256), 0, C.stream, (const uint64_t*)(d + L.oq), part, (const double2*)(d + L.os),
                       (const uint32_t*)(d + L.oo), res);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(hres.data(), res, (size_t)nq * 16, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipStreamSynchronize(C.stream));
    for (uint32_t i = 0; i < nq; ++i) out[i] = {hres[2 * i], hres[2 * i + 1]};
}

// ------------------------------------------------------------- hipGraph ----
// A replay's first node: zero the batch's work-queue counters (the first 256
// bytes of its device staging, 32 words) with a kernel rather than a memset
// node.
__global__ __launch_bounds__(64) void k_clear_counters(unsigned long long* __restrict__ ctr) {
    if (threadIdx.x < 32) ctr[threadIdx.x] = 0ull;
}

struct Z8Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *h_in = nullptr, *h_out = nullptr, *h_part = nullptr, *d_in = nullptr, *d_part = nullptr, *d_res = nullptr;
    QueryTail tail;
    bool on_host = false;
    uint32_t nq = 0, grid = 0;
    uint64_t items = 0, branches = 0;
    hipEvent_t k0 = nullptr, k1 = nullptr;
    ~Z8Graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : {h_in, h_out, h_part}) if (p) (void)hipHostFree(p);
        for (void* p : {d_in, d_part, d_res}) if (p) (void)hipFree(p);
        if (k0) (void)hipEventDestroy(k0);
        if (k1) (void)hipEventDestroy(k1);
    }
};

std::shared_ptr<Z8Graph> dev_z8_graph(const std::vector<Z8Job>& jobs, int walk) {
    Ctx& C = ctx();
    auto G = std::make_shared<Z8Graph>();
    BatchPlan B = plan_batch(jobs, walk);
    const uint32_t nq = B.nq();
    if (nq == 0) throw std::invalid_argument("a graph needs at least one nonzero query");
    G->nq = nq; G->items = B.total_items; G->branches = B.branches; G->on_host = host_finishes(B);
    const StageLayout L = stage_layout(B);
    std::lock_guard<std::mutex> lk(C.mu);
    HIPCHECK(hipHostMalloc(&G->h_in, L.tot, hipHostMallocDefault));
    HIPCHECK(hipHostMalloc(&G->h_out, nq * 16, hipHostMallocMapped));
    HIPCHECK(hipMalloc(&G->d_in, L.tot));
    HIPCHECK(hipEventCreate(&G->k0)); HIPCHECK(hipEventCreate(&G->k1));
    stage_copy(B, L, (uint8_t*)G->h_in);
    uint8_t* d = (uint8_t*)G->d_in;
    // every allocation before the capture: none is allowed inside it
    double* dp = nullptr;
    if (G->on_host) {
        HIPCHECK(hipHostMalloc(&G->h_part, B.total_items * 32, hipHostMallocMapped));
        HIPCHECK(hipHostGetDevicePointer((void**)&dp, G->h_part, 0));
        G->tail.set(B);
    } else {
        HIPCHECK(hipMalloc(&G->d_part, std::max<uint64_t>(B.total_items * 32, 32)));
        HIPCHECK(hipMalloc(&G->d_res, nq * 16));
        dp = (double*)G->d_part;
    }
    // The staged descriptors, query records, scales and tables do not change
    // from replay to replay: they are uploaded once, here, and a replay only
    // clears the queue counters before the walk.
    HIPCHECK(hipMemcpyAsync(d, G->h_in, L.tot, hipMemcpyHostToDevice, C.stream));
    HIPCHECK(hipStreamSynchronize(C.stream));
    HIPCHECK(hipHostFree(G->h_in));
    G->h_in = nullptr;
    HIPCHECK(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal));
    try {
        hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, C.stream, (unsigned long long*)d);
        HIPCHECK(hipGetLastError());
        G->grid = launch_walks(C, C.stream, B, d, L, dp, G->on_host);
        if (!G->on_host) {
            hipLaunchKernelGGL(k_z8_reduce, dim3(nq), dim3(256), 0, C.stream, (const uint64_t*)(d + L.oq), (const double*)G->d_part, (const double2*)(d + L.os), (const uint32_t*)(d + L.oo), (double*)G->d_res);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipMemcpyAsync(G->h_out, G->d_res, nq * 16, hipMemcpyDeviceToHost, C.stream));
        }
    } catch (...) {
        // leave the stream usable: end the capture and drop what it holds
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(C.stream, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
    }
    HIPCHECK(hipStreamEndCapture(C.stream, &G->graph));
    HIPCHECK(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
    return G;
}

void dev_z8_graph_run(Z8Graph& G, std::vector<std::complex<double>>& out, Z8Timing* timing) {
    Ctx& C = ctx();
    std::lock_guard<std::mutex> lk(C.mu);
    HIPCHECK(hipEventRecord(G.k0, C.stream));
    HIPCHECK(hipGraphLaunch(G.exec, C.stream));
    HIPCHECK(hipEventRecord(G.k1, C.stream));
    HIPCHECK(hipStreamSynchronize(C.stream));
    double* r = (double*)G.h_out;
    if (G.on_host)   // the stream is done: every partial is in place
        for (size_t q = 0; q < G.nq; ++q)
            host_query_reduce((const double*)G.h_part + 4 * G.tail.item0[q], G.tail.nitems[q], G.tail.scale[q], r + 2 * (size_t)G.tail.order[q]);
    out.resize(G.nq);
    for (uint32_t i = 0; i < G.nq; ++i) out[i] = {r[2 * i], r[2 * i + 1]};
    if (timing) {
        float ms; HIPCHECK(hipEventElapsedTime(&ms, G.k0, G.k1));
        timing->total_ms = ms; timing->kernel_ms = 0; timing->branches = G.branches; timing->items = G.items; timing->grid = G.grid;
    }
}

// ================================================================= dumps ====
__global__ __launch_bounds__(256) void k_dump_s(PlainDev E, const uint32_t* __restrict__ u0, const uint64_t* __restrict__ s, uint64_t n,
                                                uint8_t* __restrict__ zero, uint8_t* __restrict__ k8, uint8_t* __restrict__ ktot) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t NW = E.NW, NWe = E.NWe, RW = 3 * NW + NWe;
    const uint64_t sv = s[i];
    int pc1 = 0, pc3 = 0, cnt = 0;
    bool z = false;
    for (uint32_t w = 0; w < NW; ++w) {
        uint32_t a = u0[w], b = u0[NW + w], o = u0[2 * NW + w];
        for (uint32_t k = 0; k < E.t; ++k) if ((sv >> k) & 1) { const uint32_t* r = E.mrow + (size_t)k * RW; a ^= r[w]; b ^= r[NW + w]; o ^= r[2 * NW + w]; }
        pc1 += __popc(o & E.diag[2 * NW + w]); pc3 += __popc(o & E.diag[3 * NW + w]);
        cnt += __popc((E.diag[w] ^ a) & (E.diag[NW + w] ^ b));
    }
    for (uint32_t w = 0; w < NWe; ++w) {
        uint32_t e = u0[3 * NW + w];
        for (uint32_t k = 0; k < E.t; ++k) if ((sv >> k) & 1) e ^= E.mrow[(size_t)k * RW + 3 * NW + w];
        if (e != E.we[w]) z = true;
    }
    const int nn = E.sigma - 2 * (pc1 - pc3) + ((cnt & 1) << 2);
    const uint32_t g = (uint32_t)(nn & 7);
    zero[i] = z ? 1 : 0;
    k8[i] = z ? 0 : (uint8_t)g;
    ktot[i] = z ? 0 : (uint8_t)((g + 6u * (uint32_t)__popcll(sv) + 4u * (uint32_t)__popcll(E.neg & ~sv)) & 7u);
}

void dev_dump_s(const DevEngine& D, const Prepared& P, const uint64_t* s, size_t n, uint8_t* zero, uint8_t* k8, uint8_t* ktot) {
    Ctx& C = ctx();
    if (D.plain.NW == 0) throw std::invalid_argument("normal form too large for the per-branch dump (more than 1024 planes or lines)");
    const HostEngine& H = *D.host;
    const ULayout& L = H.L;
    std::vector<uint32_t> u0;
    put_u32(u0, P.u0.data() + L.off_a(), L.Wg, L.g, D.plain.NW); put_u32(u0, P.u0.data() + L.off_b(), L.Wg, L.g, D.plain.NW);
    put_u32(u0, P.u0.data() + L.off_o(), L.Wo, L.nodd, D.plain.NW); put_u32(u0, P.u0.data() + L.off_e(), L.We, L.neven, D.plain.NWe);
    std::lock_guard<std::mutex> lk(C.mu);
    uint32_t* du = (uint32_t*)C.d_misc.get(u0.size() * 4 + 256);
    uint64_t* ds = (uint64_t*)C.d_misc2.get(n * 8 + 256);
    uint8_t* dout = (uint8_t*)C.d_misc3.get(3 * n + 256);
    HIPCHECK(hipMemcpyAsync(du, u0.data(), u0.size() * 4, hipMemcpyHostToDevice, C.stream));
    HIPCHECK(hipMemcpyAsync(ds, s, n * 8, hipMemcpyHostToDevice, C.stream));
    if (n) hipLaunchKernelGGL(k_dump_s, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, C.stream, D.plain, du, ds, (uint64_t)n, dout, dout + n, dout + 2 * n);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(zero, dout, n, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipMemcpyAsync(k8, dout + n, n, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipMemcpyAsync(ktot, dout + 2 * n, n, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipStreamSynchronize(C.stream));
}

// Every branch of the walk for one query, with the walk's own lane state.
template <uint32_t NIN>
__global__ __launch_bounds__(256) void k_dump_walk(Z8Dev E, Z8Q Q, uint8_t* __restrict__ kout, float* __restrict__ mout) {
    __shared__ double s_lt[8 * 256];
    __shared__ double s_gin[4096];
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k < E.nbytes * 256u; k += 256) s_lt[k] = E.lt[k];
    if (E.npin) for (uint32_t k = tid; k < (1u << E.npin); k += 256) s_gin[k] = E.gin[k];
    __syncthreads();
    const uint64_t lane = (uint64_t)blockIdx.x * 256 + tid;
    if ((lane >> Q.H) != 0) return;
    LaneState S;
    seed_lane<NIN>(E, Q, lane, S);
    const uint32_t nout = 1u << Q.M;
    for (uint32_t o = 0; o < nout; ++o) {
        if (o) flip_mid<NIN>(E, Q, NIN + (uint32_t)__builtin_ctz(o), S);
        const double F = lane_factor(s_lt, E.nbytes, S.sout);
        for (uint32_t i = 0; i < InnerGeom<NIN>::N; ++i) {
            const uint32_t g = i ^ (i >> 1);
            const double wv = E.inner[i];
            const uint32_t pin8 = E.inaux[i], kin = E.inaux[InnerGeom<NIN>::N + i];
            const uint64_t idx = (((lane << Q.M) + o) << NIN) + i;
            if (wv == 0.0) { kout[idx] = 255; mout[idx] = 0.f; continue; }
            const uint32_t par = (uint32_t)__popc(g & S.hs) & 1u;
            const uint32_t k = (S.K + kin + 4u * par) & 7u;
            const double gi = E.npin ? s_gin[(S.pio ^ (pin8 >> 3))] : 1.0;
            kout[idx] = (uint8_t)k;
            mout[idx] = (float)(F * gi * fabs(wv));
        }
    }
}

// Integer state of every outer point of the segment walk for one query:
// K mod 8, table-basis sign mask and inner pivots, indexed by (lane << M) + o,
// computed by the same seed and flip code as k_seg_walk.
// Integer state (K mod 8, sign mask, inner pivots) of every outer point of
// one query, computed from the same seed and class-sorted point records as
// k_seg_walk, written at (lane << M) + o.
__global__ __launch_bounds__(256) void k_dump_seg(SegDev E, SegQ Q, uint8_t* __restrict__ kout, uint32_t* __restrict__ hsout, uint32_t* __restrict__ pout) {
    const uint64_t lane = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if ((lane >> Q.H) != 0) return;
    SegLaneDev L;
    const SegSlots<const uint64_t*, const uint32_t*, const float*> T{E.oq, E.on, E.ox, E.opin, E.olam, E.rmid, E.lowtab};
    seg_seed(E, T, Q, lane, L);
    const uint32_t M = Q.M;
    const uint32_t P = seg_lane_parities(E, T.oq, M, L.cout);
    const uint32_t dw = (uint32_t)(Q.delta >> E.L6) & ((1u << M) - 1u);
    const uint32_t* pt = E.pts + 4 * ((1u << M) - 1u);
    const uint32_t* cl = E.pcls + 5 * M;
    for (uint32_t c = 0; c < 4; ++c) {
        for (uint32_t p = cl[c]; p < cl[c + 1]; ++p) {
            const uint32_t X = pt[4 * p], g = pt[4 * p + 1] & 0x7FFFFFFFu, h = pt[4 * p + 1] >> 31;
            uint32_t o = g;
            for (uint32_t sh = 1; sh < 32; sh <<= 1) o ^= o >> sh;   // inverse Gray code
            const uint32_t lp = (uint32_t)((__popc(P & g) ^ __popc(dw & g)) & 1);
            const uint64_t idx = (lane << M) + o;
            kout[idx] = (uint8_t)((L.K + c + 4u * (lp ^ h)) & 7u);
            hsout[idx] = L.hs ^ X;
            pout[idx] = L.pio;
        }
    }
}

void dev_dump_seg(const DevEngine& D, const SegQueryHost& qh, uint32_t mid_bits, uint8_t* k, uint32_t* hs, uint32_t* pio) {
    Ctx& C = ctx();
    const SegImage& G = *D.seg;
    SegQ q{};
    q.delta = qh.delta; q.s0 = qh.s0; q.K0 = qh.K0; q.hs0 = qh.hs0; q.pio0 = qh.pio0;
    q.M = std::min(mid_bits, std::min(G.Mcap, G.nout - G.L6));
    q.H = G.nout - q.M;
    if (G.nout > 24) throw std::invalid_argument("dump_seg is limited to 2^24 outer points");
    const uint64_t n = 1ull << G.nout;
    std::lock_guard<std::mutex> lk(C.mu);
    uint8_t* dk = (uint8_t*)C.d_misc.get(n + 256);
    uint32_t* dh = (uint32_t*)C.d_misc2.get(n * 4 + 256);
    uint32_t* dp = (uint32_t*)C.d_misc3.get(n * 4 + 256);
    const uint64_t lanes = 1ull << q.H;
    hipLaunchKernelGGL(k_dump_seg, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, C.stream, D.sg, q, dk, dh, dp);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(k, dk, n, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipMemcpyAsync(hs, dh, n * 4, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipMemcpyAsync(pio, dp, n * 4, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipStreamSynchronize(C.stream));
}

void dev_dump_walk(const DevEngine& D, const Z8QueryHost& qh, uint32_t mid_bits, uint8_t* k, float* mag) {
    Ctx& C = ctx();
    const Z8Image& Z = *D.z8;
    Z8Q q{};
    q.delta = qh.delta; q.s0 = qh.s0; q.K0 = qh.K0; q.pio0 = qh.pio0;
    const uint32_t nout = Z.nslots - Z.kin;
    q.M = std::min(mid_bits, nout); q.H = nout - q.M;
    if (Z.kin + q.M + q.H > 26) throw std::invalid_argument("dump_walk is limited to 2^26 branches");
    const uint64_t n = 1ull << (Z.kin + q.M + q.H);
    std::lock_guard<std::mutex> lk(C.mu);
    uint8_t* dk = (uint8_t*)C.d_misc.get(n + 256);
    float* dm = (float*)C.d_misc2.get(n * 4 + 256);
    const uint64_t lanes = 1ull << q.H;
    if (Z.kin == kInnerPiv) hipLaunchKernelGGL(k_dump_walk<kInnerPiv>, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, C.stream, D.z, q, dk, dm);
    else if (Z.kin == kInnerWide) hipLaunchKernelGGL(k_dump_walk<kInnerWide>, dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, C.stream, D.z, q, dk, dm);
    else throw std::logic_error("walk image with an unknown inner slot count");
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(k, dk, n, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipMemcpyAsync(mag, dm, n * 4, hipMemcpyDeviceToHost, C.stream));
    HIPCHECK(hipStreamSynchronize(C.stream));
}

}  // namespace pq
