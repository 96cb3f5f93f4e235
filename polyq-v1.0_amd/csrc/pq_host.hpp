// Host side of the MI355X branch evaluator: circuit model, path-sum compiler,
// symmetric F2 normal form, branch-subspace elimination and the packed engine
// image that the HIP kernels read.  One engine is built per (path sum, fixed
// variable set), once, on the host; every amplitude query afterwards is a
// prologue of word XORs plus a device launch.
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>
#include <string>
#include <memory>
#include <complex>
#include <stdexcept>

namespace pq {

// ------------------------------------------------------------------ F2 rows
// A packed F2 vector: bit i in word i/64 at position i%64, tail bits zero
// (same byte layout as the reference BitVec, src/bits.rs:10-19, so a u32 view
// (i/32, i%32) is the same bytes).
struct Bits {
    std::vector<uint64_t> w;
    uint32_t n = 0;
    Bits() {}
    explicit Bits(uint32_t len) : w((len + 63) / 64, 0), n(len) {}
    inline bool get(uint32_t i) const { return (w[i >> 6] >> (i & 63)) & 1; }
    inline void set(uint32_t i) { w[i >> 6] |= 1ull << (i & 63); }
    inline void put(uint32_t i, bool v) { if (v) set(i); else w[i >> 6] &= ~(1ull << (i & 63)); }
    inline void flip(uint32_t i) { w[i >> 6] ^= 1ull << (i & 63); }
    inline void operator^=(const Bits& o) { for (size_t k = 0; k < w.size(); ++k) w[k] ^= o.w[k]; }
    inline void operator&=(const Bits& o) { for (size_t k = 0; k < w.size(); ++k) w[k] &= o.w[k]; }
    inline uint32_t popc() const { uint32_t c = 0; for (uint64_t x : w) c += (uint32_t)__builtin_popcountll(x); return c; }
    inline uint32_t and_popc(const Bits& o) const {
        uint32_t c = 0; for (size_t k = 0; k < w.size(); ++k) c += (uint32_t)__builtin_popcountll(w[k] & o.w[k]); return c;
    }
    inline bool dot(const Bits& o) const {
        uint64_t a = 0; for (size_t k = 0; k < w.size(); ++k) a ^= w[k] & o.w[k]; return __builtin_popcountll(a) & 1;
    }
    inline bool any() const { for (uint64_t x : w) if (x) return true; return false; }
    inline void clear() { for (auto& x : w) x = 0; }
    template <class F> inline void each(F f) const {
        for (size_t k = 0; k < w.size(); ++k) { uint64_t x = w[k]; while (x) { f((uint32_t)(k * 64 + __builtin_ctzll(x))); x &= x - 1; } }
    }
    inline long lowest() const {
        for (size_t k = 0; k < w.size(); ++k) if (w[k]) return (long)(k * 64 + __builtin_ctzll(w[k]));
        return -1;
    }
    bool operator==(const Bits& o) const { return n == o.n && w == o.w; }
};

// ------------------------------------------------------------ phase ring
// Angle 2*pi*k/2^32 with wrapping addition (src/phase.rs:22-36).
typedef uint32_t Phase;
constexpr Phase kPi = 1u << 31, kHalfPi = 1u << 30, kQuarterPi = 1u << 29;
inline Phase z4(uint32_t k) { return (k & 3u) << 30; }                     // phase.rs:112-114
inline int clifford_level(Phase p) { return (p & ((1u << 30) - 1)) ? -1 : (int)(p >> 30); }  // phase.rs:70-76
std::complex<double> phase_complex(Phase p);                               // phase.rs:60-66
Phase phase_from_radians(double r);                                        // phase.rs:39-44

// ------------------------------------------------------------- circuits
// Gate codes shared by the C ABI (the Gate enum of src/circuit.rs:25-42).
enum GateOp : uint8_t { G_H = 0, G_X, G_Y, G_Z, G_S, G_SDG, G_T, G_TDG, G_P, G_CZ, G_CX, G_SWAP, G_GPHASE, G_COUNT };
struct Gate { uint8_t op; uint32_t a, b; Phase ph; };
struct Circuit {
    uint32_t n = 0;
    std::vector<Gate> gates;
    void add(uint8_t op, uint32_t a, uint32_t b = 0, Phase ph = 0);
    uint32_t hadamards() const;     // circuit.rs:206-215 (CX counts two)
    Circuit inverse() const;        // circuit.rs:46-56, 189-191
};

// ------------------------------------------------------------- path sum
// <y|C|x> = 2^{-h/2} e^{i phi} sum_z i^{Q(z)} e^{i sum theta_k z_k}  (src/pathsum.rs:1-30)
struct PathSum {
    uint32_t nq = 0, nv = 0, nh = 0;
    std::vector<Bits> adj;          // symmetric, zero diagonal
    std::vector<uint8_t> v;         // Z4 linear part
    std::vector<Phase> theta;       // non-Clifford rotations
    Phase global = 0;
    std::vector<uint8_t> dead;
    std::vector<uint32_t> out_var;
    std::vector<uint8_t> out_neg;
    bool zero = false;

    std::vector<uint32_t> io_vars() const;                                  // pathsum.rs:328-336
    void substitute(uint32_t a, uint32_t b, bool c);                        // pathsum.rs:360-394
    // Boundary (var, value) pairs for <y|C|x>, one byte per qubit, any number
    // of qubits, with k extra (var, value) pairs; false if contradictory
    // (pathsum.rs:430-450).
    bool boundary(const uint8_t* x, const uint8_t* y, std::vector<std::pair<uint32_t, uint8_t>>& out, size_t k = 0,
                  const uint32_t* extra_var = nullptr, const uint8_t* extra_val = nullptr) const;
    // Same from basis indices (bit q = qubit q), with k extra (var, value) pairs.
    bool boundary_bits(uint64_t x, uint64_t y, std::vector<std::pair<uint32_t, uint8_t>>& out, size_t k = 0,
                       const uint32_t* extra_var = nullptr, const uint8_t* extra_val = nullptr) const;
};
PathSum compile(const Circuit& c);                                          // pathsum.rs:274-297
struct Sandwich { PathSum ps; std::vector<uint32_t> junction; std::vector<uint8_t> jneg; };
Sandwich compile_sandwich(const Circuit& c);                                // pathsum.rs:304-314

// ----------------------------------------------------------- normal form
// Block-diagonalising basis of the F2 Gram matrix (normal_form.rs:146-157):
// hyperbolic planes (a_j, b_j), non-isotropic lines, radical vectors.
struct NormalForm {
    uint32_t m = 0;
    std::vector<Bits> pa, pb, odd, even;
    // with the Z4 linear values given: q(p) mod 4 of each basis vector, the
    // values normal_form.rs:339-391 reads the Diag from
    std::vector<uint8_t> zpa, zpb, zodd, zeven;
};
// normal_form.rs:240-336 (MinWeight).  z4 (optional): the linear coefficient
// mod 4 of each variable (its parity is `parity`); the Z4 value of every basis
// vector is then carried through the row operations.
NormalForm normal_form(const std::vector<Bits>& b, const Bits& parity, const std::vector<uint8_t>* z4 = nullptr);

// Shift coordinates packed in one row: [a: Wg words | b: Wg | o: Wo | e: We]
// (u64 words).  Along kernel directions the e part is zero.
struct ULayout {
    uint32_t g = 0, nodd = 0, neven = 0;
    uint32_t Wg = 0, Wo = 0, We = 0;
    uint32_t words() const { return 2 * Wg + Wo + We; }
    uint32_t off_a() const { return 0; }
    uint32_t off_b() const { return Wg; }
    uint32_t off_o() const { return 2 * Wg; }
    uint32_t off_e() const { return 2 * Wg + Wo; }
};

struct EngineStats { uint64_t free_vars, planes, lines, radical, terms, branch_dim; };  // engine.rs:64-79

// Everything the device needs about one engine, host copy.  Built once per
// (path sum, fixed set): engine.rs:289-390 restated for a packed layout.
struct HostEngine {
    // sizes
    uint32_t m = 0, t = 0, d = 0, nf = 0, rk = 0;   // rk = rank K
    // Device terms: the terms a kernel vector can flip, placed first (term
    // indices 0..tdev-1), at most 64, so a branch vector over them is one u64.
    // With t <= 64 every term is a device term and the order is the
    // reference's; above 64, the terms no kernel vector touches come after
    // them and are constant over the walk (folded into each query's prologue).
    // rkdev pivots (the first ones) sit on device terms.
    uint32_t tdev = 0, rkdev = 0;
    uint32_t nh = 0;
    Phase global = 0;
    bool zero = false;
    double log2mag = 0;
    ULayout L;
    // fixed variables
    std::vector<uint32_t> fixed;
    std::vector<int32_t> fixed_pos;
    std::vector<uint8_t> fixed_v;
    std::vector<Phase> fixed_theta;
    std::vector<Bits> b_fixed;
    std::vector<uint64_t> b_fixed_flat;      // the same rows, nfw words each
    uint32_t nfw = 0;
    // packed shift rows (L.words() u64 each)
    std::vector<uint64_t> ucols;    // nf rows
    std::vector<uint64_t> mcols;    // t rows
    // Diag (normal_form.rs:205-218), packed like a shift row: wa in the a slot,
    // wb in the b slot, o1 in the o slot, we in the e slot.
    std::vector<uint64_t> diag;
    std::vector<uint64_t> o3;       // complement of o1 among the lines (Wo words)
    int32_t sigma = 0;
    // branch subspace: RREF of K (engine.rs:106-201)
    std::vector<uint32_t> pivot_col;         // rk
    std::vector<Bits> pivot_row;             // rk rows over neven
    std::vector<Bits> consistency;           // rows over neven
    std::vector<uint32_t> term_pos;          // t: position of the reference's term k in the order above
    std::vector<uint64_t> kernel;            // d kernel vectors over the device terms
    std::vector<uint32_t> free_col;          // d: the non-pivot column of kernel j
    std::vector<uint64_t> kernel_u;          // d packed shift rows
    // rotation terms
    std::vector<Phase> theta;                // t
    uint64_t neg_mask = 0;                   // theta_k > pi, device terms
    std::vector<uint8_t> neg_rest;           // theta_k > pi for terms tdev..t-1
    // factorised weights: weight(s) = E * B * omega^{6|s| + 4|neg\s|} * prod_k r_k[s_k]
    std::vector<double> r0, r1;              // per term, max(r0, r1) = 1
    double Bmag = 1.0;                       // prod_k max(|cos|, sin)
    std::complex<double> Ephase{1, 0};       // e^{i sum theta_k / 2}
    // quadratic Z8 form of the per-branch phase over kernel coordinates
    std::vector<uint64_t> Q;                 // d rows (bit k of row j = Q_jk), symmetric, zero diagonal
    std::vector<uint8_t> lam;                // query-independent part of the linear coefficients
    std::vector<uint64_t> dual;              // d packed rows: (b_j | a_j | y_j | 0) for the query-dependent parity
    // The query-dependent parities of the d linear coefficients, as d-bit
    // masks: bit j of ucol_dq[f] = parity(ucols[f] . dual_j); bit j of
    // mcol_dq[k] = parity(mcols[k] . dual_j) ^ bit k of kernel_j.  A query's
    // mask is the XOR over its ones and the terms of s0 (parity is linear).
    std::vector<uint64_t> ucol_dq;           // nf
    std::vector<uint64_t> mcol_dq;           // t
    double mag = 1.0;                        // 2^(log2mag - nh/2) * Bmag (engine.rs:563-565)
    // reference-style FP64 term multipliers for the plain kernel (engine.rs:221-238)
    std::vector<std::complex<double>> a0, a1, up, down;
    EngineStats stats() const { return EngineStats{m, L.g, L.nodd, L.neven, t, d}; }
};

// Build an engine for `ps` with the given fixed variables (engine.rs:289-370).
std::shared_ptr<HostEngine> build_engine(const PathSum& ps, const std::vector<uint32_t>& fixed);
// Place the terms (device terms first, see HostEngine::tdev) given the RREF
// kernel vectors over all t terms and their non-pivot columns: permutes mcols,
// theta and the pivots, and sets kernel, free_col, kernel_u, tdev and rkdev.
// Throws TermLimit when more than 64 terms vary over the branch subspace.
void place_terms(HostEngine& E, const std::vector<Bits>& kernel_full, const std::vector<uint32_t>& free_full);
// An engine whose branch vectors vary in more than 64 terms: the device walks
// keep them in one 64-bit word.  It carries the engine's stats, which are
// complete by then, so Strategy::Auto can still run its cost model on it.
struct TermLimit : std::invalid_argument {
    EngineStats st;
    explicit TermLimit(const EngineStats& s)
        : std::invalid_argument("more than 64 non-Clifford terms vary over the branch subspace"), st(s) {}
};
// Derive the evaluator's data from an engine's core fields (fixed values,
// fixed edges b_fixed, layout L, ucols, mcols, diag, o3, sigma, the RREF
// pivot_col / pivot_row / consistency, kernel, free_col, kernel_u, theta, nh,
// global, zero); build_engine ends with it, and so does an engine built from
// the fields of an existing one (pq_engine_create).
void finish_engine(HostEngine& E);

// Per-query prologue (engine.rs:569-610) extended with the Z8 form data.
struct Prepared {
    bool nonzero = false;            // false: the amplitude is exactly 0
    Phase phase = 0;                 // boundary phase (integer exact)
    uint64_t s0 = 0;                 // particular solution over the device terms
    std::vector<uint32_t> s0_rest;   // its terms beyond the device terms (t > 64 only)
    std::vector<uint64_t> u_start;   // packed shift row at c = 0
    std::vector<uint64_t> u0;        // shift row of the boundary alone (s = 0)
    uint32_t K0 = 0;                 // total Z8 phase at c = 0
    std::vector<uint8_t> ell;        // d linear Z8 coefficients (filled by prepared_ell only)
    uint64_t dq = 0;                 // bit j: ell[j] - lam[j] = 4 (the query-dependent part)
    std::complex<double> scale;      // e^{i phase} 2^{log2mag - h/2} E B
};
bool prepare(const HostEngine& E, const uint8_t* values, Prepared& P);
// The prologue of Engine::sum_subspace's own arguments (engine.rs:669): the
// start shift u (a packed shift row) and branch vector s_ref (t bits in the
// reference's term order) of the affine branch subspace, which prepare leaves
// in sc.u and sc.s0 (engine.rs:596-606), instead of fixed values.  P.phase is
// 0 and P.scale is E B (the sum is that of weight(s) ROOT8[phase8(u(s))]
// without eval_with's boundary phase and magnitude, engine.rs:557-565).
// False (the sum is 0) for an engine whose path sum vanishes identically.
bool prepare_start(const HostEngine& E, const uint64_t* u, const uint64_t* s_ref, Prepared& P);
// weight(s) (engine.rs:661-667) of a branch vector in the reference's term
// order, multiplied in that order.
std::complex<double> weight_ref(const HostEngine& E, const uint64_t* s_ref);
// Reference-style FP64 weight at s0 (engine.rs:229-238: the product of the
// term multipliers a1[k] for k in s0, a0[k] otherwise, in term order) over the
// device terms, and the constant product over the other terms; only the plain
// FP64 form uses them.
std::complex<double> plain_c_start(const HostEngine& E, uint64_t s0);
std::complex<double> plain_c_rest(const HostEngine& E, const Prepared& P);
// The d linear Z8 coefficients of a prepared query, ell_j = lam_j + 4 dq_j
// (the prologue itself keeps only the mask dq).
void prepared_ell(const HostEngine& E, Prepared& P);
// Boundary of <y|C|x> mapped straight to the engine's fixed positions
// (pathsum.rs:430-450): input qubit q fixes position in_pos[q] to x_q; output
// qubit q fixes out_pos[q] to y_q ^ out_neg[q], or, when its variable already
// has a value (out_chk[q] = q' for input q', 64 + q' for output q'), must
// agree with it.
struct BoundaryProg {
    uint32_t nq = 0;
    std::vector<uint32_t> in_pos;
    std::vector<int32_t> out_pos, out_chk;
    std::vector<uint8_t> out_neg;
    // The same map as runs: bits q0 .. q0+len-1 of x (src 0) or of y ^ yneg
    // (src 1) land on bits shift .. of fixed-row word `word`.  Inputs and
    // outputs usually map to a few such runs.
    struct Run { uint8_t src, q0, len, word, shift; };
    std::vector<Run> runs;
    uint64_t yneg = 0;                                  // out_neg as a mask
    std::vector<std::pair<uint32_t, int32_t>> checks;  // (q, out_chk[q]) for out_chk[q] >= 0
};
BoundaryProg boundary_prog(const PathSum& ps, const HostEngine& E);
// prepare() for <y|C|x> from basis indices; *contradictory tells a
// contradictory boundary from a vanishing or inconsistent engine.
bool prepare_bits(const HostEngine& E, const BoundaryProg& B, uint64_t x, uint64_t y, Prepared& P, bool* contradictory = nullptr);
// Same, with a (var, value) assignment covering every fixed variable.
bool prepare_assignment(const HostEngine& E, const std::vector<std::pair<uint32_t, uint8_t>>& asg, Prepared& P);

// Host evaluation of the Gauss-sum phase (normal_form.rs:479-493) for a packed
// shift row, and of the total per-branch phase of the factorised weight.
uint32_t phase8_row(const HostEngine& E, const uint64_t* u);
// Z8 form data of an arbitrary kernel direction (N over terms, U packed row).
uint32_t z8_lambda(const HostEngine& E, const uint64_t* U, uint64_t N);
bool z8_cross(const HostEngine& E, const uint64_t* U1, uint64_t N1, const uint64_t* U2, uint64_t N2);
void z8_dual(const HostEngine& E, const uint64_t* U, uint64_t* out);
bool survives_row(const HostEngine& E, const uint64_t* u);
uint32_t total_phase(const HostEngine& E, const uint64_t* u, uint64_t s);

// ------------------------------------------------------- Z8 walk image
// Coordinates of the branch subspace are laid out in slots: kin inner slots
// (walked from a uniform 2^kin-entry table; slots without a real coordinate
// are phantoms whose table entries are zero), then the outer slots, whose
// first M are walked in Gray order per lane and the rest index the lane's
// coset.  The walk keeps 16 accumulators for the four low inner slots and
// steps over the kin - 4 high ones in Gray order (2^(kin - 4) steps of 16
// branches per outer point).  The table lives in VGPRs (2^(kin - 3) of them),
// so its size is set by the registers the kernel variant has left: 10 inner
// slots when no inner direction touches a pivot, 9 when the walk also carries
// pivot state.
constexpr uint32_t kInnerWide = 10;
constexpr uint32_t kInnerPiv = 9;
constexpr uint32_t kMidFree = 8;     // pivot-free outer slots whose magnitudes a lane takes from a per-query table
struct Z8Image {
    uint32_t d = 0, nslots = 0, nin = 0, t = 0, npin = 0, nbytes = 0;
    uint32_t kin = 0;                  // inner slots: kInnerWide, or kInnerPiv (the walk variant with pivot lookups)
    bool imag0 = false;                // inner entries alternate real / imaginary with slot 0
    std::vector<uint64_t> slot_u;      // packed shift row of each slot's direction (host only)
    std::vector<uint64_t> slot_dual;   // query dual row of each slot (host only)
    std::vector<uint64_t> slot_q;      // Q rows over slots
    std::vector<uint64_t> slot_n;      // kernel vector N over the t terms (0 for a phantom)
    std::vector<uint32_t> slot_pin;    // inner-touched pivot bits of N, compacted
    std::vector<uint8_t> slot_lam;     // query-independent linear coefficient
    std::vector<uint8_t> slot_real;    // slot holds a direction (not a phantom)
    std::vector<double> inner;         // 2^kin signed FP64 values W[i] (entry i covers c_in = gray(i))
    // The same values in the order of the accumulator walk (no inner
    // pivots): entry 16 q + l covers c_in = (gray(q) << 4) | l.
    std::vector<double> inner_acc;
    std::vector<uint32_t> inner_aux;   // 2^kin x pin*8 (LDS byte offset of the inner pivots), then 2^kin x phase
    std::vector<double> lt;            // nbytes x 256 lane-factor byte tables (FP64)
    // Pivot-free outer slots come first after the inner ones: slots
    // kin .. kin + nx - 1 are single free terms that no other slot touches
    // (nx <= kMidFree), so as mid slots their magnitudes are the same for
    // every lane of a query.  mid_term[j] / mid_r[2 j + b] are the term of
    // slot kin + j and its magnitude r_b; lt_rest are the lane-factor
    // byte tables without those nx terms.
    uint32_t nx = 0;
    // The four low inner slots touch no pivot, so with inner pivots the pivot
    // factor of a branch depends on the high inner slots alone.
    bool low_pivot_free = false;
    std::vector<uint32_t> mid_term;
    std::vector<double> mid_r;
    std::vector<double> lt_rest;
    std::vector<double> gin;           // 2^npin inner pivot factor table (FP64)
    std::vector<uint32_t> pin_term;    // term index of each compact inner pivot bit
    uint64_t lane_terms = 0;           // terms whose ratio is carried by the lane factor
    // VALU lane-ops per branch of the inner step: one FP64 FMA (the sign, and
    // with inner pivots the pivot factor, is applied once per 16 branches),
    // except with inner pivots touched by the low inner slots: then a sign
    // XOR, an address XOR, an LDS read and one FP64 FMA.
    uint32_t ops_per_branch() const { return npin && !low_pivot_free ? 4u : 1u; }
};
Z8Image build_z8_image(const HostEngine& E);
// Query part of the Z8 walk: K0, delta bits over slots, s0, compact pivots.
struct Z8QueryHost { uint32_t K0; uint64_t delta; uint64_t s0; uint32_t pio0; };
Z8QueryHost z8_query(const HostEngine& E, const Z8Image& Z, const Prepared& P);

// ------------------------------------------------- segment-table image
// (pq_seg.cpp) Inner sums over 2^I kernel coordinates precomputed per engine
// as a Walsh-Hadamard table; the device walks only the outer coordinates.
constexpr uint32_t kSegMaxMid = 10;
struct SegImage {
    uint32_t d = 0, t = 0, I = 0, npin = 0, nout = 0, L6 = 0, Mcap = 0, nbytes = 0;
    std::vector<uint32_t> inner_j, outer_j;   // kernel index of each inner / outer slot
    std::vector<uint64_t> oq;                 // per outer slot: Q row over outer slots
    std::vector<uint32_t> ox;                 // per outer slot: cross mask over inner slots, table basis
    std::vector<uint32_t> opin;               // per outer slot: compact inner pivots touched
    std::vector<uint32_t> olam;               // per outer slot: query-independent linear coefficient
    std::vector<uint64_t> on;                 // per outer slot: kernel vector over the t terms
    std::vector<uint32_t> hsmap;              // inner sign bit p -> table-basis mask
    std::vector<uint32_t> pin_term;           // term index of each compact inner pivot bit
    uint64_t lane_terms = 0;
    std::vector<float> table;                 // 2^(I+npin) x {re hi, re lo, im hi, im lo}, index (pio << I) | h
    std::vector<float> lt;                    // nbytes x 256 x {hi, lo}: lane-factor byte tables
    std::vector<float> fmid;                  // M = 0..Mcap: 2^M x {hi, lo} walk magnitudes at 2(2^M - 1)
    std::vector<float> rmid;                  // Mcap x {r0 hi, r0 lo, r1 hi, r1 lo}
    // Walk table, o < 2^Mcap, c = gray(o) over the walkable slots: the sign-mask
    // change X(o) (table basis, bits 0..23) and the engine part of the phase
    // change Ue(o) = sum c_j lambda_j + 4 sum_{j<j'} c_j c_j' Q_jj' mod 8 (bits 24..26).
    std::vector<uint32_t> wtab;
    // Lane-low table: for each pattern c of the L6 lane-low slots, the seed
    // terms those slots contribute (query part delta excluded):
    // {sum lambda_k + 4 |Q_k & c & below(k)| mod 8, XOR ox, XOR opin, 0,
    //  XOR on (lo, hi)}, 6 words per entry, 2^L6 entries.
    std::vector<uint32_t> lowtab;
    // The walk's points sorted by Ue mod 4: for M = 0..Mcap, at entry offset
    // 2^M - 1, 2^M records {X(o), gray(o) | h << 31, (-1)^h Fmid_M(o) hi, lo}
    // with h = Ue >> 2 (omega^4 = -1), and class boundaries pcls[5 M + c]
    // (c = 0..4, relative to the M block).
    std::vector<uint32_t> pts;
    std::vector<uint32_t> pcls;
};
SegImage build_seg_image(const HostEngine& E);
void seg_set_bits(uint32_t bits);             // table bits I + npin of engines built afterwards
uint32_t seg_get_bits();
struct SegQueryHost { uint32_t K0, hs0, pio0; uint64_t delta, s0; };
SegQueryHost seg_query(const HostEngine& E, const SegImage& S, const Prepared& P);
uint32_t seg_mid_bits(const SegImage& S, uint32_t offset, uint32_t max_mid);
// Integer lane state at walk point o of a lane (host restatement of the
// device walk, for parity checks): K mod 8, table-basis sign mask, inner
// pivots, branch vector and the outer coordinates.
struct SegLane { uint32_t K, hs, pio; uint64_t s, c; };
SegLane seg_lane_state(const SegImage& S, const SegQueryHost& q, uint32_t M, uint64_t lane, uint32_t o);

// Rank-2 decomposition of rotation pairs (src/rank.rs:44-222).
constexpr uint32_t kMaxPairs = 14;
std::vector<std::pair<uint32_t, uint32_t>> pair_candidates(const PathSum& ps, const std::vector<uint32_t>& fixed, uint32_t& n_tclass);
uint32_t best_pairs(const EngineStats& st, uint32_t avail, uint32_t tclass);
// GPU refit of the cost model: host build per branch engine vs. device walk.
uint32_t best_pairs_gpu(const EngineStats& st, uint32_t avail, uint32_t tclass, double branch_s);
constexpr double kGpuBranchSeconds = 1.0 / 1.5e13;   // measured Z8 walk rate on MI355X (config 2)

}  // namespace pq
