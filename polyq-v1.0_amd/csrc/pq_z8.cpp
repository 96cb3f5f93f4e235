// Z8 walk image: the engine data in the form the optimized HIP kernel reads.
//
// For kernel coordinates c the branch vector is s(c) = s0 + sum c_j N_j
// (engine.rs:28-38, :158-171) and the per-branch term is
//     E B omega^{k(c)} prod_k r_k[s(c)_k]
// with k(c) the Z8 quadratic form built in pq_engine.cpp.  The product over
// terms splits into three parts, each handled where it varies:
//   inner free terms      -> folded into the uniform 2^kin-entry table W[i]
//   pivots touched by an  -> per-branch LDS lookup gin[pio ^ pin_i]
//   inner direction
//   every other term      -> lane factor, 8-bit byte tables, once per outer step
#include "pq_host.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace pq {

// The image with K inner slots.
static Z8Image build_z8_slots(const HostEngine& E, const uint32_t K) {
    Z8Image Z;
    const uint32_t d = E.d, t = E.tdev, W = E.L.words();   // device terms (see HostEngine::tdev)
    if (d > 56) throw std::invalid_argument("branch dimension above 56 is not supported by the Z8 walk");
    Z.d = d; Z.t = t; Z.kin = K;
    const uint32_t NI = 1u << K;
    std::vector<int32_t> piv_index(t, -1);
    for (uint32_t i = 0; i < E.rkdev; ++i) piv_index[E.pivot_col[i]] = (int32_t)i;
    auto pivots_of = [&](uint64_t N) { uint64_t pm = 0; for (uint32_t k = 0; k < t; ++k) if (((N >> k) & 1) && piv_index[k] >= 0) pm |= 1ull << piv_index[k]; return pm; };

    // Choose up to K inner directions: pivot-free kernel vectors first, then
    // the ones adding the fewest new pivots while at most 12 pivots are
    // touched (the FP64 gin table stays within 32 KB of LDS).
    std::vector<uint8_t> used(d, 0);
    std::vector<uint32_t> inner;
    uint64_t pin_set = 0;
    for (uint32_t j = 0; j < d && inner.size() < K; ++j) if (pivots_of(E.kernel[j]) == 0) { inner.push_back(j); used[j] = 1; }
    while (inner.size() < K) {
        int best = -1; uint32_t bestc = 99;
        for (uint32_t j = 0; j < d; ++j) {
            if (used[j]) continue;
            const uint64_t pj = pivots_of(E.kernel[j]);
            const uint32_t c = (uint32_t)__builtin_popcountll(pj & ~pin_set);
            if ((uint32_t)__builtin_popcountll(pin_set | pj) <= 12 && c < bestc) { bestc = c; best = (int)j; }
        }
        if (best < 0) break;
        inner.push_back((uint32_t)best); used[best] = 1; pin_set |= pivots_of(E.kernel[best]);
    }
    Z.nin = (uint32_t)inner.size();

    // Slot directions.  Inner ones may be recombined (any basis of the inner
    // subspace gives the same sum): every linear coefficient of the Z8 form is
    // even, so an inner entry is real or imaginary according to sum_p c_p
    // (lam_p / 2) mod 2; folding every "imaginary" direction but one into the
    // others leaves that one in slot 0, and the class of entry i becomes
    // bit 0 of gray(i), known at compile time in the kernel.
    std::vector<uint64_t> N;
    std::vector<std::vector<uint64_t>> U;
    for (uint32_t j : inner) { N.push_back(E.kernel[j]); U.emplace_back((E.kernel_u.data() + (size_t)j * W), (E.kernel_u.data() + (size_t)j * W) + W); }
    int p1 = -1;
    for (uint32_t p = 0; p < Z.nin; ++p) if ((z8_lambda(E, U[p].data(), N[p]) >> 1) & 1) { p1 = (int)p; break; }
    if (p1 >= 0) {
        for (uint32_t p = 0; p < Z.nin; ++p) {
            if ((int)p == p1 || !((z8_lambda(E, U[p].data(), N[p]) >> 1) & 1)) continue;
            N[p] ^= N[p1];
            for (uint32_t w = 0; w < W; ++w) U[p][w] ^= U[p1][w];
        }
        std::swap(N[0], N[p1]); std::swap(U[0], U[p1]);
        Z.imag0 = true;
    }
    while (N.size() < K) { N.push_back(0); U.emplace_back(W, 0ull); }
    // outer slots: pivot-free kernel vectors first (each is one free term)
    for (int pass = 0; pass < 2; ++pass)
        for (uint32_t j = 0; j < d; ++j) {
            if (used[j] || (pivots_of(E.kernel[j]) == 0) != (pass == 0)) continue;
            N.push_back(E.kernel[j]); U.emplace_back((E.kernel_u.data() + (size_t)j * W), (E.kernel_u.data() + (size_t)j * W) + W);
            if (pass == 0 && Z.nx < kMidFree) { Z.mid_term.push_back(E.free_col[j]); ++Z.nx; }
        }
    Z.nslots = (uint32_t)N.size();
    if (Z.nslots > 64) throw std::invalid_argument("too many slots");
    const uint32_t S = Z.nslots;
    Z.slot_real.assign(S, 0);
    for (uint32_t p = 0; p < S; ++p) Z.slot_real[p] = (p < Z.nin || p >= K) ? 1 : 0;

    // compact numbering of the pivots touched by inner directions
    uint64_t touched = 0;
    for (uint32_t p = 0; p < Z.nin; ++p) touched |= pivots_of(N[p]);
    std::vector<int32_t> cpos(E.rkdev, -1);
    for (uint32_t i = 0; i < E.rkdev; ++i) if ((touched >> i) & 1) { cpos[i] = (int32_t)Z.pin_term.size(); Z.pin_term.push_back(E.pivot_col[i]); }
    Z.npin = (uint32_t)Z.pin_term.size();

    Z.slot_q.assign(S, 0); Z.slot_n.assign(S, 0); Z.slot_pin.assign(S, 0); Z.slot_lam.assign(S, 0);
    Z.slot_u.assign((size_t)S * W, 0); Z.slot_dual.assign((size_t)S * W, 0);
    for (uint32_t p = 0; p < S; ++p) {
        if (!Z.slot_real[p]) continue;
        Z.slot_n[p] = N[p];
        std::copy(U[p].begin(), U[p].end(), (Z.slot_u.data() + (size_t)p * W));
        Z.slot_lam[p] = (uint8_t)z8_lambda(E, U[p].data(), N[p]);
        z8_dual(E, U[p].data(), (Z.slot_dual.data() + (size_t)p * W));
        const uint64_t pm = pivots_of(N[p]);
        for (uint32_t i = 0; i < E.rkdev; ++i) if (((pm >> i) & 1) && cpos[i] >= 0) Z.slot_pin[p] |= 1u << cpos[i];
        for (uint32_t q = 0; q < S; ++q)
            if (q != p && Z.slot_real[q] && z8_cross(E, U[p].data(), N[p], U[q].data(), N[q])) Z.slot_q[p] |= 1ull << q;
    }
    // Pivot-free inner directions come first and the imaginary-class carrier
    // moved into slot 0 is the first of its class, so the low four slots are
    // pivot-free whenever at least four inner directions are and the carrier
    // (if any) is one of them.
    Z.low_pivot_free = (Z.slot_pin[0] | Z.slot_pin[1] | Z.slot_pin[2] | Z.slot_pin[3]) == 0;

    // Terms carried by the lane factor: all but the free terms of the inner
    // directions and the inner-touched pivots.
    const uint64_t all = t == 64 ? ~0ull : ((1ull << t) - 1);
    uint64_t inner_free = 0;
    for (uint32_t j : inner) inner_free |= 1ull << E.free_col[j];
    uint64_t excl = inner_free;
    for (uint32_t term : Z.pin_term) excl |= 1ull << term;
    Z.lane_terms = all & ~excl;

    // Inner table: entry i covers c_in = gray(i); phantom bits give zero.
    Z.inner.assign(NI, 0.0);
    Z.inner_aux.assign(2 * NI, 0u);
    for (uint32_t i = 0; i < NI; ++i) {
        const uint32_t c = i ^ (i >> 1);
        uint32_t pin = 0, kin = 0;
        uint64_t sin = 0;
        const bool phantom = (c >> Z.nin) != 0;
        for (uint32_t p = 0; p < Z.nin; ++p) {
            if (!((c >> p) & 1)) continue;
            kin += Z.slot_lam[p] + 4 * ((uint32_t)__builtin_popcountll(Z.slot_q[p] & (uint64_t)(c & ((1u << p) - 1))) & 1);
            pin ^= Z.slot_pin[p];
            sin ^= N[p];
        }
        kin &= 7;
        if (!phantom) {
            if (kin & 1) throw std::logic_error("odd inner phase: the Z8 form has odd linear coefficients");
            const uint32_t cls = (kin >> 1) & 1;
            if (cls != (Z.imag0 ? (c & 1u) : 0u)) throw std::logic_error("inner real/imaginary classes do not follow slot 0");
            double fin = 1.0;
            for (uint64_t rest = inner_free; rest; rest &= rest - 1) {
                const uint32_t k = (uint32_t)__builtin_ctzll(rest);
                fin *= ((sin >> k) & 1) ? E.r1[k] : E.r0[k];
            }
            Z.inner[i] = (kin & 4) ? -fin : fin;
        }
        Z.inner_aux[i] = pin << 3;
        Z.inner_aux[NI + i] = kin;
    }
    // Accumulator-walk order: c_in = (gray(q) << 4) | l at 16 q + l; entry i
    // of the Gray order covers gray(i), so c_in sits at i = gray^-1(c_in).
    Z.inner_acc.assign(NI, 0.0);
    for (uint32_t q = 0; q < (NI >> 4); ++q)
        for (uint32_t l = 0; l < 16; ++l) {
            uint32_t c = ((q ^ (q >> 1)) << 4) | l, i = c;
            for (uint32_t sh = 1; sh < 32; sh <<= 1) i ^= i >> sh;
            Z.inner_acc[16 * q + l] = Z.inner[i];
        }
    // Lane-factor byte tables, with every lane term and without the free
    // terms of the pivot-free outer slots K .. K + nx - 1
    uint64_t mid_terms = 0;
    for (uint32_t j = 0; j < Z.nx; ++j) {
        const uint32_t k = Z.mid_term[j];
        if (Z.slot_n[K + j] != 1ull << k) throw std::logic_error("pivot-free outer slot is not its free term alone");
        mid_terms |= 1ull << k;
        Z.mid_r.push_back(E.r0[k]);
        Z.mid_r.push_back(E.r1[k]);
    }
    Z.nbytes = (t + 7) / 8;
    Z.lt.assign((size_t)Z.nbytes * 256, 1.0);
    Z.lt_rest.assign((size_t)Z.nbytes * 256, 1.0);
    for (uint32_t b = 0; b < Z.nbytes; ++b)
        for (uint32_t v = 0; v < 256; ++v) {
            double f = 1.0, fr = 1.0;
            for (uint32_t q = 0; q < 8; ++q) {
                const uint32_t k = 8 * b + q;
                if (k >= t || !((Z.lane_terms >> k) & 1)) continue;
                const double r = ((v >> q) & 1) ? E.r1[k] : E.r0[k];
                f *= r;
                if (!((mid_terms >> k) & 1)) fr *= r;
            }
            Z.lt[(size_t)b * 256 + v] = f;
            Z.lt_rest[(size_t)b * 256 + v] = fr;
        }
    // inner pivot table
    Z.gin.assign((size_t)1 << Z.npin, 1.0);
    for (uint32_t pi = 0; pi < (1u << Z.npin); ++pi) {
        double f = 1.0;
        for (uint32_t b = 0; b < Z.npin; ++b) f *= ((pi >> b) & 1) ? E.r1[Z.pin_term[b]] : E.r0[Z.pin_term[b]];
        Z.gin[pi] = f;
    }
    return Z;
}

// kInnerWide inner slots when that many pivot-free kernel vectors exist (or
// every vector is pivot-free): the inner directions, chosen pivot-free first,
// then touch no pivot and the plain walk variant runs the image.  Otherwise
// kInnerPiv, walked by the variant with pivot lookups, whose pivot state
// leaves no registers for the larger table (that variant also walks a
// kInnerPiv image that turns out pivot-free).
Z8Image build_z8_image(const HostEngine& E) {
    std::vector<int32_t> piv_index(E.tdev, -1);
    for (uint32_t i = 0; i < E.rkdev; ++i) piv_index[E.pivot_col[i]] = (int32_t)i;
    uint32_t free_vecs = 0;
    for (uint32_t j = 0; j < E.d; ++j) {
        bool pf = true;
        for (uint64_t b = E.kernel[j]; b && pf; b &= b - 1) pf = piv_index[__builtin_ctzll(b)] < 0;
        free_vecs += pf;
    }
    return build_z8_slots(E, free_vecs >= std::min(E.d, kInnerWide) ? kInnerWide : kInnerPiv);
}

Z8QueryHost z8_query(const HostEngine& E, const Z8Image& Z, const Prepared& P) {
    const uint32_t W = E.L.words();
    Z8QueryHost q;
    q.K0 = P.K0;
    q.s0 = P.s0;
    q.delta = 0;
    for (uint32_t p = 0; p < Z.nslots; ++p) {
        if (!Z.slot_real[p]) continue;
        uint64_t a = 0;
        const uint64_t* dr = (Z.slot_dual.data() + (size_t)p * W);
        for (uint32_t w = 0; w < W; ++w) a ^= P.u_start[w] & dr[w];
        const bool bit = ((__builtin_popcountll(a) + __builtin_popcountll(P.s0 & Z.slot_n[p])) & 1) != 0;
        if (bit) q.delta |= 1ull << p;
    }
    q.pio0 = 0;
    for (uint32_t b = 0; b < Z.npin; ++b) if ((P.s0 >> Z.pin_term[b]) & 1) q.pio0 |= 1u << b;
    return q;
}

}  // namespace pq
