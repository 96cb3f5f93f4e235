// Segment-table image of an engine: the data of the default device form.
//
// Over kernel coordinates c the per-branch term is
//     omega^{k(c)} prod_k r_k[s(c)_k],   k(c) = K0 + sum_j c_j l_j + 4 sum_{j<k} c_j c_k Q_jk
// (the Z8 quadratic form of pq_engine.cpp; s(c) = s0 + sum_j c_j N_j,
// engine.rs:28-38).  Split the coordinates into I inner ones (c_in) and the
// outer ones (c_out).  With c_out fixed, the sum over the 2^I inner branches is
//     omega^{K(c_out)} F(c_out) * T[pio(c_out)][hs(c_out)],
//     T[pio][h] = sum_{c_in} W(c_in) g(pio + pin(c_in)) (-1)^{c_in . h}
// where W holds the inner-only phase and magnitudes, g the magnitudes of the
// pivot terms the inner directions touch, and h the inner sign mask (query
// bits of the inner coordinates plus their cross terms with c_out).  T is a
// Walsh-Hadamard transform, the same for every query of the engine: it is
// computed here once, in FP64, and stored as FP32 hi/lo pairs (16 bytes per
// entry).  The device then walks only the outer coordinates in Gray order and
// adds one table entry per outer point, rotated by omega^K: a segment of 2^I
// branches per lookup.
//
// Outer slot order: the (up to) six lane-low slots first (lane bits 0..5 of a
// wavefront), then pivot-free slots that a lane may walk (their magnitude
// is a uniform table over the walk index), then the rest.  The inner sign
// basis is changed so that the six lane-low cross masks span the low bits of
// h: the 64 lanes of a wavefront then read one contiguous block of the table.
#include "pq_host.hpp"
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

namespace pq {

// Cap on the table bits I + npin.  An engine of branch dimension d gets
// min(cap, max(14, d - 14)) bits: its table is no larger than the outer walk
// of one query, so engines queried a few times (served, batched) keep cheap
// 2^14-entry tables while large resident engines walk 4x fewer points.
static std::atomic<uint32_t> g_seg_bits{16};
uint32_t seg_bits_for(uint32_t d) {
    const uint32_t cap = g_seg_bits;
    return std::min(cap, std::max<uint32_t>(14, d > 14 ? d - 14 : 0));
}
void seg_set_bits(uint32_t b) {
    if (b > 20) throw std::invalid_argument("segment table bits above 20");
    g_seg_bits = b;
}
uint32_t seg_get_bits() { return g_seg_bits; }

namespace {
inline void split_df(double v, float& hi, float& lo) {
    hi = (float)v;
    lo = (float)(v - (double)hi);
}
// Inverse of an n x n matrix over F2 given by columns (n <= 32).
std::vector<uint32_t> f2_inverse_cols(const std::vector<uint32_t>& cols, uint32_t n) {
    // rows of [A | I]
    std::vector<uint64_t> rows(n, 0);
    for (uint32_t c = 0; c < n; ++c)
        for (uint32_t r = 0; r < n; ++r)
            if ((cols[c] >> r) & 1) rows[r] |= 1ull << c;
    for (uint32_t r = 0; r < n; ++r) rows[r] |= 1ull << (32 + r);
    for (uint32_t c = 0; c < n; ++c) {
        uint32_t p = c;
        while (p < n && !((rows[p] >> c) & 1)) ++p;
        if (p == n) throw std::logic_error("singular lane basis");
        std::swap(rows[p], rows[c]);
        for (uint32_t r = 0; r < n; ++r) if (r != c && ((rows[r] >> c) & 1)) rows[r] ^= rows[c];
    }
    std::vector<uint32_t> inv(n, 0);   // columns of the inverse
    for (uint32_t r = 0; r < n; ++r)
        for (uint32_t c = 0; c < n; ++c)
            if ((rows[r] >> (32 + c)) & 1) inv[c] |= 1u << r;
    return inv;
}
}  // namespace

namespace {
// In-place Walsh-Hadamard transform of n = 2^k doubles, two radix-2 stages
// per pass (the same additions in the same order as stage-by-stage radix 2,
// in half the passes over memory).
struct SegScratch {
    std::vector<double> fin, re, im;
    std::vector<uint8_t> kin;
    std::vector<uint32_t> pinc, hof;
};
SegScratch& seg_scratch() {
    thread_local SegScratch s;
    return s;
}

void fwht_scalar(double* v, uint32_t n) {
    uint32_t len = 1;
    for (; 4 * len <= n; len *= 4)
        for (uint32_t base = 0; base < n; base += 4 * len)
            for (uint32_t x = base; x < base + len; ++x) {
                const double a0 = v[x], a1 = v[x + len], a2 = v[x + 2 * len], a3 = v[x + 3 * len];
                const double s01 = a0 + a1, d01 = a0 - a1, s23 = a2 + a3, d23 = a2 - a3;
                v[x] = s01 + s23; v[x + len] = d01 + d23;
                v[x + 2 * len] = s01 - s23; v[x + 3 * len] = d01 - d23;
            }
    if (len < n)
        for (uint32_t x = 0; x < len; ++x) {
            const double a = v[x], b = v[x + len];
            v[x] = a + b; v[x + len] = a - b;
        }
}
typedef double V4 __attribute__((vector_size(32)));
typedef long long M4 __attribute__((vector_size(32)));
// Two-source 4-lane shuffle: GCC's builtin takes an index vector, clang's a
// list of constants (the library is built with g++; the ThreadSanitizer build
// of the host code uses clang).
#if defined(__clang__)
#define PQ_SHUF4(x, y, i0, i1, i2, i3) __builtin_shufflevector(x, y, i0, i1, i2, i3)
#else
#define PQ_SHUF4(x, y, i0, i1, i2, i3) __builtin_shuffle(x, y, (M4){i0, i1, i2, i3})
#endif
inline V4 ld4(const double* p) { V4 r; std::memcpy(&r, p, 32); return r; }
inline void st4(double* p, V4 r) { std::memcpy(p, &r, 32); }
inline void bfly4(V4& a0, V4& a1, V4& a2, V4& a3) {
    const V4 s01 = a0 + a1, d01 = a0 - a1, s23 = a2 + a3, d23 = a2 - a3;
    a0 = s01 + s23; a1 = d01 + d23; a2 = s01 - s23; a3 = d01 - d23;
}
inline void transpose4(V4& a, V4& b, V4& c, V4& d) {
    const V4 t0 = PQ_SHUF4(a, b, 0, 4, 2, 6), t1 = PQ_SHUF4(a, b, 1, 5, 3, 7);
    const V4 t2 = PQ_SHUF4(c, d, 0, 4, 2, 6), t3 = PQ_SHUF4(c, d, 1, 5, 3, 7);
    a = PQ_SHUF4(t0, t2, 0, 1, 4, 5); c = PQ_SHUF4(t0, t2, 2, 3, 6, 7);
    b = PQ_SHUF4(t1, t3, 0, 1, 4, 5); d = PQ_SHUF4(t1, t3, 2, 3, 6, 7);
}
// Walsh-Hadamard transform in place: radix-4 passes over the index bit
// pairs (0,1), (2,3), ..., then one radix-2 pass if log2 n is odd.  Four
// lanes at a time: the passes of bits (0,1) and (2,3) on 16-entry blocks
// (a 4x4 transpose puts the four entries of a butterfly in four vectors),
// the later ones on four consecutive butterflies.  Every entry sees the same
// additions in the same order as the one-at-a-time loop, so the result is
// the same bit for bit.
void fwht(double* v, uint32_t n) {
    if (n < 16) { fwht_scalar(v, n); return; }
    for (uint32_t b = 0; b < n; b += 16) {
        V4 r0 = ld4(v + b), r1 = ld4(v + b + 4), r2 = ld4(v + b + 8), r3 = ld4(v + b + 12);
        transpose4(r0, r1, r2, r3);
        bfly4(r0, r1, r2, r3);          // bits (0,1)
        transpose4(r0, r1, r2, r3);
        bfly4(r0, r1, r2, r3);          // bits (2,3)
        st4(v + b, r0); st4(v + b + 4, r1); st4(v + b + 8, r2); st4(v + b + 12, r3);
    }
    uint32_t len = 16;
    for (; 4 * len <= n; len *= 4)
        for (uint32_t base = 0; base < n; base += 4 * len)
            for (uint32_t x = base; x < base + len; x += 4) {
                V4 a0 = ld4(v + x), a1 = ld4(v + x + len), a2 = ld4(v + x + 2 * len), a3 = ld4(v + x + 3 * len);
                bfly4(a0, a1, a2, a3);
                st4(v + x, a0); st4(v + x + len, a1); st4(v + x + 2 * len, a2); st4(v + x + 3 * len, a3);
            }
    if (len < n)
        for (uint32_t x = 0; x < len; x += 4) {
            const V4 a = ld4(v + x), c = ld4(v + x + len);
            st4(v + x, a + c); st4(v + x + len, a - c);
        }
}
}  // namespace

SegImage build_seg_image(const HostEngine& E) {
    SegImage S;
    const uint32_t d = E.d, t = E.tdev;   // device terms (see HostEngine::tdev)
    if (d > 56) throw std::invalid_argument("branch dimension above 56 is not supported");
    S.d = d; S.t = t;
    std::vector<int32_t> piv_index(t, -1);
    for (uint32_t i = 0; i < E.rkdev; ++i) piv_index[E.pivot_col[i]] = (int32_t)i;
    auto pivots_of = [&](uint64_t N) { uint64_t pm = 0; for (uint32_t k = 0; k < t; ++k) if (((N >> k) & 1) && piv_index[k] >= 0) pm |= 1ull << piv_index[k]; return pm; };

    // Inner directions within a table of 2^(I + npin) entries: pivot-free ones
    // first, then those adding the fewest new pivots.
    std::vector<uint8_t> used;
    uint64_t pin_set = 0;
    auto choose_inner = [&](uint32_t B) {
        S.inner_j.clear();
        used.assign(d, 0);
        pin_set = 0;
        for (uint32_t j = 0; j < d && S.inner_j.size() < B; ++j)
            if (pivots_of(E.kernel[j]) == 0) { S.inner_j.push_back(j); used[j] = 1; }
        for (;;) {
            int best = -1; uint32_t bestc = 99;
            for (uint32_t j = 0; j < d; ++j) {
                if (used[j]) continue;
                const uint64_t pj = pivots_of(E.kernel[j]);
                const uint32_t tot = (uint32_t)S.inner_j.size() + 1 + (uint32_t)__builtin_popcountll(pin_set | pj);
                const uint32_t c = (uint32_t)__builtin_popcountll(pj & ~pin_set);
                if (tot <= B && c < bestc) { bestc = c; best = (int)j; }
            }
            if (best < 0) break;
            S.inner_j.push_back((uint32_t)best); used[best] = 1; pin_set |= pivots_of(E.kernel[best]);
        }
    };
    choose_inner(seg_bits_for(d));
    // When the inner directions touch many pivots, the table's bits go to
    // pivot patterns and few are left for inner directions: the walk of one
    // query can then pass 2^24 outer points (a 4-qubit marginal's sandwich
    // engine of d = 39 kept 4 inner directions and walked 2^35 points, 25 ms).
    // Such an engine gets up to 2^kSegBigBits table entries: a host build of
    // a few milliseconds against a walk several times longer for every query
    // (that engine: 5 ms to build, 6 ms to walk).
    constexpr uint32_t kSegBigBits = 20, kSegBigOuter = 24;
    if (d - (uint32_t)S.inner_j.size() > kSegBigOuter && seg_bits_for(d) < kSegBigBits) {
        const std::vector<uint32_t> keep = S.inner_j;
        const std::vector<uint8_t> keep_used = used;
        const uint64_t keep_pin = pin_set;
        choose_inner(kSegBigBits);
        if (S.inner_j.size() <= keep.size()) { S.inner_j = keep; used = keep_used; pin_set = keep_pin; }
    }
    S.I = (uint32_t)S.inner_j.size();
    std::vector<int32_t> cpos(E.rkdev, -1);
    for (uint32_t i = 0; i < E.rkdev; ++i) if ((pin_set >> i) & 1) { cpos[i] = (int32_t)S.pin_term.size(); S.pin_term.push_back(E.pivot_col[i]); }
    S.npin = (uint32_t)S.pin_term.size();
    auto pin_of = [&](uint64_t N) { uint32_t pm = 0; const uint64_t pv = pivots_of(N); for (uint32_t i = 0; i < E.rkdev; ++i) if (((pv >> i) & 1) && cpos[i] >= 0) pm |= 1u << cpos[i]; return pm; };

    // Outer order: lane-low (slots touching pivots first), walkable pivot-free, rest.
    std::vector<uint32_t> rest_p, rest_f;
    for (uint32_t j = 0; j < d; ++j) if (!used[j]) (pivots_of(E.kernel[j]) ? rest_p : rest_f).push_back(j);
    const uint32_t nout = d - S.I;
    S.nout = nout;
    S.L6 = std::min<uint32_t>(6, nout);
    std::vector<uint32_t> order;
    size_t ip = 0, jf = 0;
    while (order.size() < S.L6) order.push_back(ip < rest_p.size() ? rest_p[ip++] : rest_f[jf++]);
    S.Mcap = std::min<uint32_t>((uint32_t)(rest_f.size() - jf), kSegMaxMid);
    for (uint32_t m = 0; m < S.Mcap; ++m) order.push_back(rest_f[jf++]);
    while (ip < rest_p.size()) order.push_back(rest_p[ip++]);
    while (jf < rest_f.size()) order.push_back(rest_f[jf++]);
    S.outer_j = order;

    // Lane basis of the inner sign mask: the lane-low cross masks go to unit vectors.
    const uint32_t I = S.I;
    auto cross_orig = [&](uint32_t j) { uint32_t x = 0; for (uint32_t p = 0; p < I; ++p) if ((E.Q[S.inner_j[p]] >> j) & 1) x |= 1u << p; return x; };
    std::vector<uint32_t> bcols;     // columns of B = (A^T)^{-1}
    {
        std::vector<uint32_t> basis;  // reduced vectors for independence tests
        auto reduce = [&](uint32_t v) { for (uint32_t b : basis) v = std::min(v, v ^ b); return v; };
        auto add = [&](uint32_t v) {
            const uint32_t r = reduce(v);
            if (!r) return false;
            basis.push_back(r);
            std::sort(basis.begin(), basis.end(), std::greater<uint32_t>());
            return true;
        };
        for (uint32_t b = 0; b < S.L6 && I; ++b) { const uint32_t x = cross_orig(S.outer_j[b]); if (add(x)) bcols.push_back(x); }
        for (uint32_t p = 0; p < I; ++p) if (add(1u << p)) bcols.push_back(1u << p);
    }
    std::vector<uint32_t> at = I ? f2_inverse_cols(bcols, I) : std::vector<uint32_t>{};   // columns of A^T
    S.hsmap.assign(I, 0);
    for (uint32_t p = 0; p < I; ++p) S.hsmap[p] = at[p];
    auto to_table_basis = [&](uint32_t x) { uint32_t y = 0; for (uint32_t p = 0; p < I; ++p) if ((x >> p) & 1) y ^= S.hsmap[p]; return y; };

    // Outer slot data
    // (at least one entry each, so a clamped index is always in bounds)
    const uint32_t nslot = std::max<uint32_t>(nout, 1);
    S.oq.assign(nslot, 0); S.ox.assign(nslot, 0); S.opin.assign(nslot, 0); S.olam.assign(nslot, 0); S.on.assign(nslot, 0);
    for (uint32_t k = 0; k < nout; ++k) {
        const uint32_t j = S.outer_j[k];
        for (uint32_t l = 0; l < nout; ++l) if (l != k && ((E.Q[j] >> S.outer_j[l]) & 1)) S.oq[k] |= 1ull << l;
        S.ox[k] = to_table_basis(cross_orig(j));
        S.opin[k] = pin_of(E.kernel[j]);
        S.olam[k] = E.lam[j];
        S.on[k] = E.kernel[j];
    }

    // Terms of the lane factor: all but the inner free terms, the inner pivots
    // and the free terms of the walkable slots.
    const uint64_t all = t == 64 ? ~0ull : ((1ull << t) - 1);
    uint64_t excl = 0;
    for (uint32_t j : S.inner_j) excl |= 1ull << E.free_col[j];
    for (uint32_t term : S.pin_term) excl |= 1ull << term;
    for (uint32_t m = 0; m < S.Mcap; ++m) excl |= 1ull << E.free_col[S.outer_j[S.L6 + m]];
    S.lane_terms = all & ~excl;
    S.nbytes = (t + 7) / 8;
    S.lt.assign((size_t)S.nbytes * 256 * 2, 0.f);
    for (uint32_t b = 0; b < S.nbytes; ++b)
        for (uint32_t v = 0; v < 256; ++v) {
            double f = 1.0;
            for (uint32_t q = 0; q < 8; ++q) {
                const uint32_t k = 8 * b + q;
                if (k >= t || !((S.lane_terms >> k) & 1)) continue;
                f *= ((v >> q) & 1) ? E.r1[k] : E.r0[k];
            }
            split_df(f, S.lt[2 * ((size_t)b * 256 + v)], S.lt[2 * ((size_t)b * 256 + v) + 1]);
        }
    // Walk magnitudes: for M = 0..Mcap, entry o = prod_{m < M} r_{f_m}[gray(o)_m]
    S.fmid.assign((size_t)2 * ((2u << S.Mcap) - 1), 0.f);
    for (uint32_t M = 0; M <= S.Mcap; ++M)
        for (uint32_t o = 0; o < (1u << M); ++o) {
            const uint32_t g = o ^ (o >> 1);
            double f = 1.0;
            for (uint32_t m = 0; m < M; ++m) { const uint32_t k = E.free_col[S.outer_j[S.L6 + m]]; f *= ((g >> m) & 1) ? E.r1[k] : E.r0[k]; }
            const size_t at2 = (size_t)2 * ((1u << M) - 1 + o);
            split_df(f, S.fmid[at2], S.fmid[at2 + 1]);
        }
    S.lowtab.assign((size_t)6 << S.L6, 0);
    for (uint32_t c = 0; c < (1u << S.L6); ++c) {
        uint32_t K = 0, hs = 0, pio = 0;
        uint64_t sm = 0;
        for (uint32_t k = 0; k < S.L6; ++k) {
            if (!((c >> k) & 1)) continue;
            K += S.olam[k] + 4 * ((uint32_t)__builtin_popcountll(S.oq[k] & (uint64_t)c & ((1ull << k) - 1)) & 1);
            hs ^= S.ox[k]; pio ^= S.opin[k]; sm ^= S.on[k];
        }
        uint32_t* e = &S.lowtab[(size_t)6 * c];
        e[0] = K & 7u; e[1] = hs; e[2] = pio; e[3] = 0; e[4] = (uint32_t)sm; e[5] = (uint32_t)(sm >> 32);
    }
    S.wtab.assign((size_t)1 << S.Mcap, 0);
    for (uint32_t o = 0; o < (1u << S.Mcap); ++o) {
        const uint32_t g = o ^ (o >> 1);
        uint32_t X = 0, U = 0;
        for (uint32_t m = 0; m < S.Mcap; ++m) {
            if (!((g >> m) & 1)) continue;
            const uint32_t k = S.L6 + m;
            X ^= S.ox[k];
            U += S.olam[k] + 4 * ((uint32_t)__builtin_popcountll(S.oq[k] & ((uint64_t)(g & ((1u << m) - 1)) << S.L6)) & 1);
        }
        S.wtab[o] = X | ((U & 7u) << 24);
    }
    // Point records per walk length M, grouped by class Ue mod 4: omega^4 =
    // -1, so a point of class c + 4 is a point of class c with its magnitude
    // negated, and four accumulators cover the eight phases.  Record
    // {X, gray | h << 31, f hi, f lo} with h = Ue >> 2 and f = (-1)^h fmid.
    // (plus 16 zero records, so a short walk may read 16 records from any block)
    S.pts.assign((size_t)4 * ((2u << S.Mcap) - 1 + 16), 0);
    S.pcls.assign((size_t)5 * (S.Mcap + 1), 0);
    for (uint32_t M = 0; M <= S.Mcap; ++M) {
        std::vector<uint32_t> idx(1u << M);
        for (uint32_t o = 0; o < (1u << M); ++o) idx[o] = o;
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return ((S.wtab[a] >> 24) & 3u) < ((S.wtab[b] >> 24) & 3u); });
        const size_t base = (size_t)4 * ((1u << M) - 1);
        for (uint32_t p = 0; p < (1u << M); ++p) {
            const uint32_t o = idx[p];
            const uint32_t U = S.wtab[o] >> 24, h = U >> 2;
            uint32_t* e = &S.pts[base + 4 * (size_t)p];
            e[0] = S.wtab[o] & 0xFFFFFFu;
            e[1] = (o ^ (o >> 1)) | (h << 31);
            float fm[2] = {S.fmid[2 * ((1u << M) - 1 + o)], S.fmid[2 * ((1u << M) - 1 + o) + 1]};
            if (h) { fm[0] = -fm[0]; fm[1] = -fm[1]; }
            std::memcpy(&e[2], fm, 8);
            S.pcls[5 * M + 1 + (U & 3u)]++;
        }
        for (uint32_t c = 1; c <= 4; ++c) S.pcls[5 * M + c] += S.pcls[5 * M + c - 1];
    }
    S.rmid.assign((size_t)4 * std::max<uint32_t>(S.Mcap, 1), 0.f);
    for (uint32_t m = 0; m < S.Mcap; ++m) {
        const uint32_t k = E.free_col[S.outer_j[S.L6 + m]];
        split_df(E.r0[k], S.rmid[4 * m], S.rmid[4 * m + 1]);
        split_df(E.r1[k], S.rmid[4 * m + 2], S.rmid[4 * m + 3]);
    }

    // Inner weights W(c) g(pio + pin(c)) and their Walsh-Hadamard transform.
    const uint32_t NI = 1u << I, NP = 1u << S.npin;
    // Work arrays of up to 2^16 entries (about 2 MB) live in per-thread
    // buffers that later builds on the thread reuse: fresh blocks this large
    // come from mmap, and touching their pages cost about as much as the
    // transforms themselves.
    SegScratch& W = seg_scratch();
    W.fin.assign(NI, 1.0);
    W.kin.assign(NI, 0);
    W.pinc.assign(NI, 0);
    std::vector<double>& fin = W.fin;
    std::vector<uint8_t>& kin = W.kin;
    std::vector<uint32_t>& pinc = W.pinc;
    for (uint32_t p = 0; p < I; ++p) {
        const uint32_t j = S.inner_j[p], k = E.free_col[j];
        uint32_t qlow = 0;
        for (uint32_t q = 0; q < p; ++q) if ((E.Q[j] >> S.inner_j[q]) & 1) qlow |= 1u << q;
        const uint32_t pm = pin_of(E.kernel[j]);
        for (uint32_t c = 0; c < (1u << p); ++c) {
            const uint32_t c1 = c | (1u << p);
            fin[c1] = fin[c] * E.r1[k];
            fin[c] *= E.r0[k];
            kin[c1] = (uint8_t)((kin[c] + E.lam[j] + 4 * (__builtin_popcount(qlow & c) & 1)) & 7);
            pinc[c1] = pinc[c] ^ pm;
        }
    }
    std::vector<double> gpin(NP, 1.0);
    for (uint32_t pi = 0; pi < NP; ++pi)
        for (uint32_t b = 0; b < S.npin; ++b) gpin[pi] *= ((pi >> b) & 1) ? E.r1[S.pin_term[b]] : E.r0[S.pin_term[b]];
    static const double H = 0.70710678118654752440;
    const double w8r[8] = {1, H, 0, -H, -1, -H, 0, H}, w8i[8] = {0, H, 1, H, 0, -H, -1, -H};
    S.table.assign((size_t)NP * NI * 4, 0.f);
    // table index hp -> transform index h = B hp (B: the columns bcols), one
    // XOR per entry from the entry with the lowest set bit cleared
    W.hof.assign(NI, 0);
    std::vector<uint32_t>& hof = W.hof;
    for (uint32_t hp = 1; hp < NI; ++hp) hof[hp] = hof[hp & (hp - 1)] ^ bcols[__builtin_ctz(hp)];
    W.re.resize(NI); W.im.resize(NI);
    std::vector<double>& re = W.re;
    std::vector<double>& im = W.im;
    for (uint32_t pio = 0; pio < NP; ++pio) {
        for (uint32_t c = 0; c < NI; ++c) {
            const double w = fin[c] * gpin[pio ^ pinc[c]];
            re[c] = w * w8r[kin[c]];
            im[c] = w * w8i[kin[c]];
        }
        fwht(re.data(), NI);
        fwht(im.data(), NI);
        for (uint32_t hp = 0; hp < NI; ++hp) {
            const uint32_t h = hof[hp];
            float* dst = &S.table[((size_t)pio * NI + hp) * 4];
            split_df(re[h], dst[0], dst[1]);
            split_df(im[h], dst[2], dst[3]);
        }
    }
    return S;
}

SegQueryHost seg_query(const HostEngine& E, const SegImage& S, const Prepared& P) {
    SegQueryHost q;
    q.K0 = P.K0;
    q.s0 = P.s0;
    q.delta = 0;
    for (uint32_t k = 0; k < S.nout; ++k) q.delta |= ((P.dq >> S.outer_j[k]) & 1) << k;
    q.hs0 = 0;
    for (uint32_t p = 0; p < S.I; ++p) q.hs0 ^= S.hsmap[p] & (0u - (uint32_t)((P.dq >> S.inner_j[p]) & 1));
    q.pio0 = 0;
    for (uint32_t b = 0; b < S.npin; ++b) if ((P.s0 >> S.pin_term[b]) & 1) q.pio0 |= 1u << b;
    return q;
}

uint32_t seg_mid_bits(const SegImage& S, uint32_t offset, uint32_t max_mid) {
    const uint32_t room = S.nout - S.L6;
    uint32_t M = S.nout > offset ? S.nout - offset : 0;
    return std::min(std::min(M, max_mid), std::min(room, S.Mcap));
}

SegLane seg_lane_state(const SegImage& S, const SegQueryHost& q, uint32_t M, uint64_t lane, uint32_t o) {
    // the lane's hi coordinates and the Gray point o of its walk
    uint64_t c = 0;
    const uint32_t H = S.nout - M;
    for (uint32_t b = 0; b < H; ++b) if ((lane >> b) & 1) c |= 1ull << (b < S.L6 ? b : S.L6 + M + (b - S.L6));
    const uint32_t g = o ^ (o >> 1);
    for (uint32_t m = 0; m < M; ++m) if ((g >> m) & 1) c |= 1ull << (S.L6 + m);
    SegLane L{q.K0, q.hs0, q.pio0, q.s0, c};
    for (uint32_t k = 0; k < S.nout; ++k) {
        if (!((c >> k) & 1)) continue;
        L.K += S.olam[k] + 4 * (uint32_t)((q.delta >> k) & 1) + 4 * ((uint32_t)__builtin_popcountll(S.oq[k] & c & ((1ull << k) - 1)) & 1);
        L.hs ^= S.ox[k];
        L.pio ^= S.opin[k];
        L.s ^= S.on[k];
    }
    L.K &= 7;
    return L;
}

}  // namespace pq
