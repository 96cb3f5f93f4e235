// Z8 walk image: the engine data in the form the optimized HIP kernel reads.
//
// For kernel coordinates c the branch vector is s(c) = s0 + sum c_j N_j
// (engine.rs:28-38, :158-171) and the per-branch term is
//     E B omega^{k(c)} prod_k r_k[s(c)_k]
// with k(c) the Z8 quadratic form built in pq_engine.cpp.  The product over
// terms splits into three parts, each handled where it varies:
//   inner free terms      -> folded into the uniform 256-entry table W[i]
//   pivots touched by an  -> per-branch LDS lookup gin[pio ^ pin_i]
//   inner direction
//   every other term      -> lane factor, 8-bit byte tables, once per outer step
#include "pq_host.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace pq {

Z8Image build_z8_image(const HostEngine& E) {
    Z8Image Z;
    const uint32_t d = E.d, t = E.t;
    if (d > 56) throw std::invalid_argument("branch dimension above 56 is not supported by the Z8 walk");
    Z.d = d; Z.t = t;
    // Pivot index of each pivot term.
    std::vector<int32_t> piv_index(t, -1);
    for (uint32_t i = 0; i < E.rk; ++i) piv_index[E.pivot_col[i]] = (int32_t)i;
    auto pivots_of = [&](uint32_t j) { uint64_t pm = 0; for (uint32_t k = 0; k < t; ++k) if (((E.kernel[j] >> k) & 1) && piv_index[k] >= 0) pm |= 1ull << piv_index[k]; return pm; };

    // Choose up to 8 inner coordinates: pivot-free directions first, then the
    // directions that add the fewest new pivots while at most 12 pivots are
    // touched (the gin table stays within 16 KB of LDS).
    std::vector<uint8_t> used(d, 0);
    std::vector<uint32_t> inner;
    uint64_t pin_set = 0;
    for (uint32_t j = 0; j < d && inner.size() < kInner; ++j) if (pivots_of(j) == 0) { inner.push_back(j); used[j] = 1; }
    while (inner.size() < kInner) {
        int best = -1; uint32_t bestc = 99;
        for (uint32_t j = 0; j < d; ++j) {
            if (used[j]) continue;
            uint32_t c = (uint32_t)__builtin_popcountll(pivots_of(j) & ~pin_set);
            if ((uint32_t)__builtin_popcountll(pin_set | pivots_of(j)) <= 12 && c < bestc) { bestc = c; best = (int)j; }
        }
        if (best < 0) break;
        inner.push_back((uint32_t)best); used[best] = 1; pin_set |= pivots_of((uint32_t)best);
    }
    Z.nin = (uint32_t)inner.size();
    Z.slot_coord.assign(kInner, -1);
    for (uint32_t p = 0; p < Z.nin; ++p) Z.slot_coord[p] = (int32_t)inner[p];
    for (uint32_t j = 0; j < d; ++j) if (!used[j]) Z.slot_coord.push_back((int32_t)j);
    Z.nslots = (uint32_t)Z.slot_coord.size();
    if (Z.nslots > 64) throw std::invalid_argument("too many slots");

    // compact pivot numbering for the inner-touched pivots
    std::vector<int32_t> cpos(E.rk, -1);
    for (uint32_t i = 0; i < E.rk; ++i) if ((pin_set >> i) & 1) { cpos[i] = (int32_t)Z.pin_term.size(); Z.pin_term.push_back(E.pivot_col[i]); }
    Z.npin = (uint32_t)Z.pin_term.size();

    const uint32_t S = Z.nslots;
    Z.slot_q.assign(S, 0); Z.slot_n.assign(S, 0); Z.slot_pin.assign(S, 0); Z.slot_lam.assign(S, 0);
    for (uint32_t p = 0; p < S; ++p) {
        int32_t j = Z.slot_coord[p];
        if (j < 0) continue;
        Z.slot_n[p] = E.kernel[j];
        Z.slot_lam[p] = E.lam[j];
        uint64_t pm = pivots_of((uint32_t)j);
        for (uint32_t i = 0; i < E.rk; ++i) if (((pm >> i) & 1) && cpos[i] >= 0) Z.slot_pin[p] |= 1u << cpos[i];
        for (uint32_t p2 = 0; p2 < S; ++p2) {
            int32_t j2 = Z.slot_coord[p2];
            if (j2 >= 0 && ((E.Q[j] >> j2) & 1)) Z.slot_q[p] |= 1ull << p2;
        }
    }

    // Terms carried by the lane factor: all but the inner free terms and the
    // inner-touched pivots.
    uint64_t all = t == 64 ? ~0ull : ((1ull << t) - 1);
    uint64_t excl = 0;
    for (uint32_t p = 0; p < Z.nin; ++p) excl |= 1ull << E.free_col[Z.slot_coord[p]];
    for (uint32_t term : Z.pin_term) excl |= 1ull << term;
    Z.lane_terms = all & ~excl;

    // Inner table.  Entry i covers c_in = gray(i); phantom bits give zero.
    static const double S2 = 0.70710678118654752440;
    static const double C8[8] = {1, S2, 0, -S2, -1, -S2, 0, S2}, S8[8] = {0, S2, 1, S2, 0, -S2, -1, -S2};
    Z.inner.assign(256 * 4, 0.f);
    for (uint32_t i = 0; i < 256; ++i) {
        uint32_t c = i ^ (i >> 1);
        uint32_t pin = 0, kin = 0;
        bool phantom = (c >> Z.nin) != 0;
        double fin = 1.0;
        for (uint32_t p = 0; p < Z.nin; ++p) {
            uint32_t term = E.free_col[Z.slot_coord[p]];
            bool b = (c >> p) & 1;
            fin *= b ? E.r1[term] : E.r0[term];
            if (b) {
                kin += Z.slot_lam[p];
                kin += 4 * ((uint32_t)__builtin_popcountll(Z.slot_q[p] & (uint64_t)(c & ((1u << p) - 1))) & 1);
                pin ^= Z.slot_pin[p];
            }
        }
        kin &= 7;
        if (!phantom) { Z.inner[2 * i] = (float)(fin * C8[kin]); Z.inner[2 * i + 1] = (float)(fin * S8[kin]); }
        uint32_t pin4 = pin << 2;
        std::memcpy(&Z.inner[512 + i], &pin4, 4);
        std::memcpy(&Z.inner[768 + i], &kin, 4);
    }
    // Lane-factor byte tables
    Z.nbytes = (t + 7) / 8;
    Z.lt.assign((size_t)Z.nbytes * 256, 1.f);
    for (uint32_t b = 0; b < Z.nbytes; ++b)
        for (uint32_t v = 0; v < 256; ++v) {
            double f = 1.0;
            for (uint32_t q = 0; q < 8; ++q) {
                uint32_t k = 8 * b + q;
                if (k >= t || !((Z.lane_terms >> k) & 1)) continue;
                f *= ((v >> q) & 1) ? E.r1[k] : E.r0[k];
            }
            Z.lt[(size_t)b * 256 + v] = (float)f;
        }
    // inner pivot table
    Z.gin.assign((size_t)1 << Z.npin, 1.f);
    for (uint32_t pi = 0; pi < (1u << Z.npin); ++pi) {
        double f = 1.0;
        for (uint32_t b = 0; b < Z.npin; ++b) f *= ((pi >> b) & 1) ? E.r1[Z.pin_term[b]] : E.r0[Z.pin_term[b]];
        Z.gin[pi] = (float)f;
    }
    return Z;
}

Z8QueryHost z8_query(const HostEngine& E, const Z8Image& Z, const Prepared& P) {
    Z8QueryHost q;
    q.K0 = P.K0;
    q.s0 = P.s0;
    q.delta = 0;
    for (uint32_t p = 0; p < Z.nslots; ++p) {
        int32_t j = Z.slot_coord[p];
        if (j >= 0 && ((P.ell[j] ^ E.lam[j]) & 4)) q.delta |= 1ull << p;
    }
    q.pio0 = 0;
    for (uint32_t b = 0; b < Z.npin; ++b) if ((P.s0 >> Z.pin_term[b]) & 1) q.pio0 |= 1u << b;
    return q;
}

}  // namespace pq
