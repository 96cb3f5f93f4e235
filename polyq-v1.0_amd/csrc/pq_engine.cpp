// Engine construction on the host (once per circuit and fixed set) and the
// per-query prologue.
//
// Reference: src/engine.rs:289-390 (Engine::new, build_branching, bind_values),
// src/normal_form.rs:56-391 (Structure::new with the MinWeight rule, diagonal),
// src/engine.rs:106-201 (Elim), src/engine.rs:569-610 (prepare).
//
// Beyond the reference, the builder derives the per-branch phase as a Z8
// quadratic form over the kernel coordinates c of the surviving affine
// subspace:
//     k(c) = K0 + sum_j c_j l_j + 4 sum_{j<k} c_j c_k Q_jk   (mod 8)
// where k(c) = phase8(u(c)) + 6|s(c)| + 4|neg \ s(c)| is the Gauss-sum phase
// (normal_form.rs:479-493) plus the integer phase of the factorised branch
// weight (alpha = e^{i th/2} cos(th/2), beta = -i e^{i th/2} sin(th/2)).
// Q is query independent; l_j = lam_j + 4 delta_j with one query-dependent bit
// delta_j per coordinate.  The device walks c and never touches the packed
// shift rows.
#include "pq_host.hpp"
#include <algorithm>
#include <cstring>
#include <cmath>
#include <queue>
#include <map>

namespace pq {

// --------------------------------------------------------- normal form ----
// Symmetric LDL^T over F2 with 1x1 and 2x2 pivots.  Pivot choice follows the
// reference's MinWeight rule (normal_form.rs:269-330): first the lightest
// alive row with a diagonal bit, else the lightest alive row with a neighbour,
// ties by row index.  Bucket bitsets per (class, weight) replace the two
// ordered sets: the lowest row of the lightest non-empty bucket is the
// smallest (weight, index) entry, so the pivot sequence is the same.
// With z4 given, z[r] = q(P_r) mod 4 follows each basis row: q is a Z4
// quadratic refinement of the Gram form beta (q(x) = sum_i v_i x_i + 2 sum_{i<j}
// b_ij x_i x_j, beta = B + diag(v mod 2)), so q(x ^ y) = q(x) + q(y) + 2
// beta(x, y) mod 4, and beta(P_r, P_s) is the current Gram entry (r, s).
NormalForm normal_form(const std::vector<Bits>& b, const Bits& parity, const std::vector<uint8_t>* z4) {
    const uint32_t m = (uint32_t)b.size();
    const bool track = z4 != nullptr;
    std::vector<uint8_t> z;
    if (track) {
        if (z4->size() != m) throw std::logic_error("normal form: Z4 values do not match the Gram rows");
        z = *z4;
        for (uint32_t i = 0; i < m; ++i) if ((z[i] & 1) != (uint8_t)parity.get(i)) throw std::logic_error("normal form: Z4 value and parity disagree");
    }
    const size_t W = (m + 63) / 64;
    // Gram and basis rows in two flat arrays (row r at r * W), scratch rows
    // allocated once: a pivot step allocates nothing.
    std::vector<uint64_t> gram(std::max<size_t>(m * W, 1), 0), basis(std::max<size_t>(m * W, 1), 0);
    for (uint32_t i = 0; i < m; ++i) {
        std::memcpy(&gram[i * W], b[i].w.data(), W * 8);
        uint64_t& g = gram[i * W + (i >> 6)];
        g = (g & ~(1ull << (i & 63))) | ((uint64_t)parity.get(i) << (i & 63));
        basis[i * W + (i >> 6)] |= 1ull << (i & 63);
    }
    auto G = [&](uint32_t r) { return &gram[(size_t)r * W]; };
    auto P = [&](uint32_t r) { return &basis[(size_t)r * W]; };
    auto bit = [](const uint64_t* x, uint32_t i) { return (x[i >> 6] >> (i & 63)) & 1; };
    auto clr = [](uint64_t* x, uint32_t i) { x[i >> 6] &= ~(1ull << (i & 63)); };
    auto each = [&](const uint64_t* x, auto f) {
        for (size_t k = 0; k < W; ++k) for (uint64_t v = x[k]; v; v &= v - 1) f((uint32_t)(k * 64 + __builtin_ctzll(v)));
    };
    auto lowest = [&](const uint64_t* x) -> long {
        for (size_t k = 0; k < W; ++k) if (x[k]) return (long)(k * 64 + __builtin_ctzll(x[k]));
        return -1;
    };
    auto to_bits = [&](const uint64_t* x) { Bits r(m); std::memcpy(r.w.data(), x, W * 8); return r; };
    std::vector<uint64_t> alive(std::max<size_t>(W, 1), 0);
    for (uint32_t i = 0; i < m; ++i) alive[i >> 6] |= 1ull << (i & 63);

    // Rows by class (1: diagonal entry set, 2: off-diagonal only) and weight
    // (alive neighbours), as bucket bitsets: the next pivot of a class is the
    // lowest row of its lightest non-empty bucket, the (weight, row) order of
    // a min-heap, with O(1) updates and no stale entries.
    struct Buckets {
        std::vector<uint64_t> rows;      // per weight, W words each
        std::vector<uint32_t> cnt;
        size_t W;
        uint32_t lo;                     // no non-empty bucket below lo
        Buckets(uint32_t m, size_t W) : rows((m + 1) * std::max<size_t>(W, 1), 0), cnt(m + 1, 0), W(W), lo(m + 1) {}
        void add(uint32_t w, uint32_t a) { rows[w * W + (a >> 6)] |= 1ull << (a & 63); cnt[w]++; if (w < lo) lo = w; }
        void del(uint32_t w, uint32_t a) { rows[w * W + (a >> 6)] &= ~(1ull << (a & 63)); cnt[w]--; }
        long top() {
            for (; lo < cnt.size(); ++lo)
                if (cnt[lo])
                    for (size_t k = 0; k < W; ++k) if (rows[lo * W + k]) return (long)(k * 64 + __builtin_ctzll(rows[lo * W + k]));
            return -1;
        }
    };
    Buckets bd(m, W), bo(m, W);        // diagonal / off-diagonal classes
    std::vector<uint8_t> cls(m, 0);
    std::vector<uint32_t> wt(m, 0);
    auto reclass = [&](uint32_t a) {
        if (cls[a] == 1) bd.del(wt[a], a);
        else if (cls[a] == 2) bo.del(wt[a], a);
        if (!bit(alive.data(), a)) { cls[a] = 0; return; }
        const uint64_t* g = G(a);
        uint32_t w = 0;
        for (size_t k = 0; k < W; ++k) w += (uint32_t)__builtin_popcountll(g[k] & alive[k]);
        if (bit(g, a)) { cls[a] = 1; wt[a] = w; bd.add(w, a); }
        else if (w > 0) { cls[a] = 2; wt[a] = w; bo.add(w, a); }
        else cls[a] = 0;
    };
    auto unclass = [&](uint32_t a) {
        if (cls[a] == 1) bd.del(wt[a], a);
        else if (cls[a] == 2) bo.del(wt[a], a);
        cls[a] = 0;
    };
    for (uint32_t a = 0; a < m; ++a) reclass(a);

    std::vector<uint64_t> scratch(6 * std::max<size_t>(W, 1));
    uint64_t* s = &scratch[0];
    uint64_t* sa = &scratch[W];
    uint64_t* sb = &scratch[2 * W];
    uint64_t* pa = &scratch[3 * W];
    uint64_t* pb = &scratch[4 * W];
    uint64_t* touched = &scratch[5 * W];
    NormalForm nf; nf.m = m;
    for (;;) {
        long a = bd.top();
        if (a >= 0) {
            // 1x1 pivot: A'_bc = A_bc + A_ba A_ac, p_b += p_a for neighbours b
            const uint32_t ua = (uint32_t)a;
            for (size_t k = 0; k < W; ++k) { s[k] = G(ua)[k] & alive[k]; pa[k] = P(ua)[k]; }
            clr(s, ua);
            each(s, [&](uint32_t r) {
                uint64_t* g = G(r); uint64_t* p = P(r);
                for (size_t k = 0; k < W; ++k) { g[k] ^= s[k]; p[k] ^= pa[k]; }
                if (track) z[r] = (uint8_t)((z[r] + z[ua] + 2) & 3);     // beta(P_r, P_a) = 1
            });
            unclass(ua);
            clr(alive.data(), ua);
            nf.odd.push_back(to_bits(pa));
            if (track) nf.zodd.push_back(z[ua]);
            each(s, [&](uint32_t r) { reclass(r); });
            continue;
        }
        a = bo.top();
        if (a < 0) break;
        // 2x2 (hyperbolic) pivot on a and its first alive neighbour
        const uint32_t ua = (uint32_t)a;
        for (size_t k = 0; k < W; ++k) sa[k] = G(ua)[k] & alive[k];
        const uint32_t bb = (uint32_t)lowest(sa);
        clr(sa, bb);
        for (size_t k = 0; k < W; ++k) { sb[k] = G(bb)[k] & alive[k]; pa[k] = P(ua)[k]; pb[k] = P(bb)[k]; }
        clr(sb, ua);
        for (size_t k = 0; k < W; ++k) touched[k] = sa[k] | sb[k];
        each(touched, [&](uint32_t r) {
            uint64_t* g = G(r); uint64_t* p = P(r);
            const bool ra = bit(sa, r), rb = bit(sb, r);
            if (ra) for (size_t k = 0; k < W; ++k) { g[k] ^= sb[k]; p[k] ^= pb[k]; }
            if (rb) for (size_t k = 0; k < W; ++k) { g[k] ^= sa[k]; p[k] ^= pa[k]; }
            if (track) {
                // P_r ^= P_b with beta(P_r, P_b) = [r in sb]; then P_r ^= P_a
                // with beta = beta(P_r, P_a) (+ beta(P_b, P_a) = 1 if r in sa),
                // which is 0 either way
                if (ra) z[r] = (uint8_t)((z[r] + z[bb] + 2 * (rb ? 1 : 0)) & 3);
                if (rb) z[r] = (uint8_t)((z[r] + z[ua]) & 3);
            }
        });
        unclass(ua); unclass(bb);
        clr(alive.data(), ua); clr(alive.data(), bb);
        nf.pa.push_back(to_bits(pa)); nf.pb.push_back(to_bits(pb));
        if (track) { nf.zpa.push_back(z[ua]); nf.zpb.push_back(z[bb]); }
        each(touched, [&](uint32_t r) { reclass(r); });
    }
    each(alive.data(), [&](uint32_t r) {
        nf.even.push_back(to_bits(P(r)));
        if (track) nf.zeven.push_back(z[r]);
    });
    return nf;
}

static inline bool row_parity(const uint64_t* x, const uint64_t* y, uint32_t nw) {
    uint64_t a = 0; for (uint32_t k = 0; k < nw; ++k) a ^= x[k] & y[k]; return __builtin_popcountll(a) & 1;
}
static inline uint32_t row_popc(const uint64_t* x, uint32_t nw) {
    uint32_t c = 0; for (uint32_t k = 0; k < nw; ++k) c += (uint32_t)__builtin_popcountll(x[k]); return c;
}
static inline uint32_t row_and_popc(const uint64_t* x, const uint64_t* y, uint32_t nw) {
    uint32_t c = 0; for (uint32_t k = 0; k < nw; ++k) c += (uint32_t)__builtin_popcountll(x[k] & y[k]); return c;
}

// Gauss-sum phase of a shift row (normal_form.rs:479-493), written as
// sigma + 2|u_o| + 4(|u_o & o1| + |(wa^ua)&(wb^ub)|) mod 8, which is the same
// integer because o3 is the complement of o1 among the lines.
uint32_t phase8_row(const HostEngine& E, const uint64_t* u) {
    const ULayout& L = E.L;
    const uint64_t* D = E.diag.data();
    uint32_t uo = row_popc(u + L.off_o(), L.Wo);
    uint32_t uo1 = row_and_popc(u + L.off_o(), D + L.off_o(), L.Wo);
    uint32_t pl = 0;
    for (uint32_t k = 0; k < L.Wg; ++k)
        pl += (uint32_t)__builtin_popcountll((D[L.off_a() + k] ^ u[L.off_a() + k]) & (D[L.off_b() + k] ^ u[L.off_b() + k]));
    int v = E.sigma + 2 * (int)uo + 4 * (int)((uo1 + pl) & 1);
    return (uint32_t)(((v % 8) + 8) % 8);
}
bool survives_row(const HostEngine& E, const uint64_t* u) {
    const ULayout& L = E.L;
    for (uint32_t k = 0; k < L.We; ++k) if (u[L.off_e() + k] != E.diag[L.off_e() + k]) return false;
    return true;
}
uint32_t total_phase(const HostEngine& E, const uint64_t* u, uint64_t s) {
    uint32_t ns = (uint32_t)__builtin_popcountll(s);
    uint32_t nn = (uint32_t)__builtin_popcountll(E.neg_mask & ~s);
    return (phase8_row(E, u) + 6 * ns + 4 * nn) & 7;
}

// Z8 form data of a kernel direction (N over the terms, U its packed shift
// row, e part zero).  Derivation: expand phase8(u0 + sum c_j U_j) + 6|s| +
// 4|neg \ s| with |x + y| = |x| + |y| - 2|x & y| mod 4; every cross term is a
// multiple of 4 and every linear coefficient is even.
uint32_t z8_lambda(const HostEngine& E, const uint64_t* U, uint64_t N) {
    const ULayout& L = E.L;
    const uint64_t* D = E.diag.data();
    const uint64_t* A = U + L.off_a(); const uint64_t* B = U + L.off_b(); const uint64_t* Y = U + L.off_o();
    uint32_t l = 2 * row_popc(Y, L.Wo) + 4 * (row_and_popc(Y, D + L.off_o(), L.Wo) & 1)
               + 4 * row_parity(A, B, L.Wg) + 4 * row_parity(A, D + L.off_b(), L.Wg) + 4 * row_parity(D + L.off_a(), B, L.Wg)
               + 6 * (uint32_t)__builtin_popcountll(N) + 4 * ((uint32_t)__builtin_popcountll(E.neg_mask & N) & 1);
    return l & 7;
}
bool z8_cross(const HostEngine& E, const uint64_t* U1, uint64_t N1, const uint64_t* U2, uint64_t N2) {
    const ULayout& L = E.L;
    return row_parity(U1 + L.off_o(), U2 + L.off_o(), L.Wo) ^ row_parity(U1 + L.off_a(), U2 + L.off_b(), L.Wg)
         ^ row_parity(U2 + L.off_a(), U1 + L.off_b(), L.Wg) ^ ((__builtin_popcountll(N1 & N2) & 1) != 0);
}
// Row whose parity with the query's start shift gives the query bit of a
// direction: <ua, b> + <ub, a> + <uo, y>.
void z8_dual(const HostEngine& E, const uint64_t* U, uint64_t* out) {
    const ULayout& L = E.L;
    std::fill(out, out + L.words(), 0ull);
    std::copy_n(U + L.off_b(), L.Wg, out + L.off_a());
    std::copy_n(U + L.off_a(), L.Wg, out + L.off_b());
    std::copy_n(U + L.off_o(), L.Wo, out + L.off_o());
}

// ---------------------------------------------------------------- engine --
std::shared_ptr<HostEngine> build_engine(const PathSum& ps, const std::vector<uint32_t>& fixed) {
    auto Ep = std::make_shared<HostEngine>();
    HostEngine& E = *Ep;
    const uint32_t nv = ps.nv;
    E.fixed = fixed;
    E.nf = (uint32_t)fixed.size();
    E.fixed_pos.assign(nv, -1);
    for (uint32_t i = 0; i < E.nf; ++i) {
        uint32_t f = fixed[i];
        if (f >= nv || ps.dead[f]) throw std::invalid_argument("fixed variable is not live");
        if (E.fixed_pos[f] >= 0) throw std::invalid_argument("duplicate fixed variable");
        E.fixed_pos[f] = (int32_t)i;
    }
    std::vector<uint32_t> freev;
    std::vector<int32_t> fidx(nv, -1);
    for (uint32_t i = 0; i < nv; ++i) if (!ps.dead[i] && E.fixed_pos[i] < 0) { fidx[i] = (int32_t)freev.size(); freev.push_back(i); }
    const uint32_t m = E.m = (uint32_t)freev.size();
    std::vector<Bits> bff(m, Bits(m));
    Bits parity(m);
    for (uint32_t i = 0; i < m; ++i) {
        ps.adj[freev[i]].each([&](uint32_t j) { if (fidx[j] >= 0) bff[i].set((uint32_t)fidx[j]); });
        parity.put(i, ps.v[freev[i]] & 1);
    }
    E.b_fixed.assign(E.nf, Bits(E.nf));
    for (uint32_t i = 0; i < E.nf; ++i) ps.adj[fixed[i]].each([&](uint32_t j) { if (E.fixed_pos[j] >= 0) E.b_fixed[i].set((uint32_t)E.fixed_pos[j]); });
    std::vector<uint8_t> z4(m);
    for (uint32_t i = 0; i < m; ++i) z4[i] = (uint8_t)(ps.v[freev[i]] & 3);
    NormalForm nf = normal_form(bff, parity, &z4);
    ULayout& L = E.L;
    L.g = (uint32_t)nf.pa.size(); L.nodd = (uint32_t)nf.odd.size(); L.neven = (uint32_t)nf.even.size();
    L.Wg = (L.g + 63) / 64; L.Wo = (L.nodd + 63) / 64; L.We = (L.neven + 63) / 64;
    const uint32_t W = L.words();

    // Column rows: for each needed free variable j, the packed coordinates
    // (p_l[j])_l (normal_form.rs:413-433).  Needed = neighbours of fixed
    // variables and rotated variables.
    std::vector<uint8_t> needed(m, 0);
    for (uint32_t f : fixed) ps.adj[f].each([&](uint32_t j) { if (fidx[j] >= 0) needed[fidx[j]] = 1; });
    std::vector<uint32_t> term_vars;
    for (uint32_t i = 0; i < m; ++i) if (ps.theta[freev[i]] != 0) { needed[i] = 1; term_vars.push_back(i); }
    std::vector<int32_t> crow(m, -1);
    uint32_t ncr = 0;
    for (uint32_t i = 0; i < m; ++i) if (needed[i]) crow[i] = (int32_t)ncr++;
    std::vector<uint64_t> cols((size_t)ncr * W, 0);
    auto scatter = [&](const std::vector<Bits>& vecs, uint32_t off) {
        for (uint32_t l = 0; l < vecs.size(); ++l)
            vecs[l].each([&](uint32_t j) { if (crow[j] >= 0) cols[(size_t)crow[j] * W + off + (l >> 6)] |= 1ull << (l & 63); });
    };
    scatter(nf.pa, L.off_a()); scatter(nf.pb, L.off_b()); scatter(nf.odd, L.off_o()); scatter(nf.even, L.off_e());
    E.ucols.assign((size_t)E.nf * W, 0);
    for (uint32_t i = 0; i < E.nf; ++i)
        ps.adj[fixed[i]].each([&](uint32_t j) {
            if (fidx[j] >= 0) { const uint64_t* c = &cols[(size_t)crow[fidx[j]] * W]; for (uint32_t k = 0; k < W; ++k) E.ucols[(size_t)i * W + k] ^= c[k]; }
        });
    E.t = (uint32_t)term_vars.size();
    E.mcols.assign((size_t)E.t * W, 0);
    for (uint32_t k = 0; k < E.t; ++k) std::copy_n(&cols[(size_t)crow[term_vars[k]] * W], W, (E.mcols.data() + (size_t)k * W));

    // Diag (normal_form.rs:354-391) from the basis vectors' Z4 values d(p) =
    // q(p) mod 4 (normal_form.rs:339-348), carried through the elimination
    E.diag.assign(W, 0);
    E.o3.assign(L.Wo, 0);
    for (uint32_t j = 0; j < L.g; ++j) {
        if (nf.zpa[j] == 2) E.diag[L.off_a() + (j >> 6)] |= 1ull << (j & 63);
        if (nf.zpb[j] == 2) E.diag[L.off_b() + (j >> 6)] |= 1ull << (j & 63);
    }
    E.sigma = 0;
    for (uint32_t l = 0; l < L.nodd; ++l) {
        if (nf.zodd[l] == 1) { E.diag[L.off_o() + (l >> 6)] |= 1ull << (l & 63); E.sigma += 1; }
        else { E.o3[l >> 6] |= 1ull << (l & 63); E.sigma -= 1; }
    }
    for (uint32_t l = 0; l < L.neven; ++l)
        if (nf.zeven[l] == 2) E.diag[L.off_e() + (l >> 6)] |= 1ull << (l & 63);

    // RREF of K = M restricted to the radical coordinates (engine.rs:119-184),
    // rows over all t terms
    const uint32_t t = E.t, ne = L.neven;
    std::vector<Bits> krow(ne, Bits(t));
    std::vector<Bits> rrow(ne, Bits(ne));
    for (uint32_t e = 0; e < ne; ++e) {
        for (uint32_t c = 0; c < t; ++c) if ((E.mcols[(size_t)c * W + L.off_e() + (e >> 6)] >> (e & 63)) & 1) krow[e].set(c);
        rrow[e].set(e);
    }
    uint32_t rank = 0;
    std::vector<uint8_t> is_piv(t, 0);
    for (uint32_t col = 0; col < t; ++col) {
        uint32_t r = rank;
        while (r < ne && !krow[r].get(col)) ++r;
        if (r == ne) continue;
        std::swap(krow[rank], krow[r]); std::swap(rrow[rank], rrow[r]);
        for (uint32_t r2 = 0; r2 < ne; ++r2)
            if (r2 != rank && krow[r2].get(col)) { krow[r2] ^= krow[rank]; rrow[r2] ^= rrow[rank]; }
        E.pivot_col.push_back(col);
        is_piv[col] = 1;
        rank++;
    }
    E.rk = rank;
    // the R rows are read after the elimination finished (engine.rs:149)
    for (uint32_t i = 0; i < rank; ++i) E.pivot_row.push_back(rrow[i]);
    for (uint32_t r = rank; r < ne; ++r) E.consistency.push_back(rrow[r]);
    std::vector<Bits> kernel_full;
    std::vector<uint32_t> free_full;
    for (uint32_t j = 0; j < t; ++j) {
        if (is_piv[j]) continue;
        Bits nvec(t);
        nvec.set(j);
        for (uint32_t i = 0; i < rank; ++i) if (krow[i].get(j)) nvec.set(E.pivot_col[i]);
        kernel_full.push_back(nvec);
        free_full.push_back(j);
    }
    for (uint32_t k = 0; k < t; ++k) E.theta.push_back(ps.theta[freev[term_vars[k]]]);
    place_terms(E, kernel_full, free_full);

    for (uint32_t f : fixed) { E.fixed_v.push_back(ps.v[f]); E.fixed_theta.push_back(ps.theta[f]); }
    E.nh = ps.nh; E.global = ps.global; E.zero = ps.zero;
    finish_engine(E);
    return Ep;
}

void place_terms(HostEngine& E, const std::vector<Bits>& kernel_full, const std::vector<uint32_t>& free_full) {
    const uint32_t t = E.t, W = E.L.words();
    // perm[new position] = term; device terms (those a kernel vector flips)
    // first when t > 64, else the reference's order
    std::vector<uint32_t> perm;
    if (t <= 64) {
        for (uint32_t k = 0; k < t; ++k) perm.push_back(k);
        E.tdev = t;
    } else {
        std::vector<uint8_t> active(t, 0);
        for (const Bits& n : kernel_full) n.each([&](uint32_t k) { active[k] = 1; });
        for (uint32_t k = 0; k < t; ++k) if (active[k]) perm.push_back(k);
        E.tdev = (uint32_t)perm.size();
        if (E.tdev > 64) throw TermLimit(EngineStats{E.m, E.L.g, E.L.nodd, E.L.neven, t, (uint64_t)kernel_full.size()});
        for (uint32_t k = 0; k < t; ++k) if (!active[k]) perm.push_back(k);
    }
    std::vector<uint32_t> pos(t);
    for (uint32_t i = 0; i < t; ++i) pos[perm[i]] = i;
    E.term_pos = pos;
    std::vector<uint64_t> mc((size_t)t * W);
    std::vector<Phase> th(t);
    for (uint32_t i = 0; i < t; ++i) {
        std::copy_n((E.mcols.data() + (size_t)perm[i] * W), W, (mc.data() + (size_t)i * W));
        th[i] = E.theta[perm[i]];
    }
    E.mcols.swap(mc);
    E.theta.swap(th);
    // pivots on device terms first (their index is a bit of a u64 set)
    std::vector<uint32_t> order(E.rk);
    for (uint32_t i = 0; i < E.rk; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return pos[E.pivot_col[a]] < pos[E.pivot_col[b]]; });
    std::vector<uint32_t> pc;
    std::vector<Bits> pr;
    E.rkdev = 0;
    for (uint32_t i : order) {
        pc.push_back(pos[E.pivot_col[i]]);
        pr.push_back(E.pivot_row[i]);
        if (pc.back() < E.tdev) E.rkdev++;
    }
    E.pivot_col.swap(pc);
    E.pivot_row.swap(pr);
    E.kernel.clear(); E.free_col.clear();
    for (size_t j = 0; j < kernel_full.size(); ++j) {
        uint64_t nv = 0;
        kernel_full[j].each([&](uint32_t k) { nv |= 1ull << pos[k]; });
        E.kernel.push_back(nv);
        E.free_col.push_back(pos[free_full[j]]);
    }
    E.d = (uint32_t)E.kernel.size();
    E.kernel_u.assign((size_t)E.d * W, 0);
    for (uint32_t j = 0; j < E.d; ++j)
        for (uint32_t k = 0; k < E.tdev; ++k) if ((E.kernel[j] >> k) & 1)
            for (uint32_t w = 0; w < W; ++w) E.kernel_u[(size_t)j * W + w] ^= E.mcols[(size_t)k * W + w];
}

// Everything the evaluator derives from the engine's core fields (fixed
// values and edges, shift rows, Diag, RREF and kernel, term angles): the
// packed fixed-edge rows, the magnitude, the term multipliers and the Z8
// quadratic form.
void finish_engine(HostEngine& E) {
    const ULayout& L = E.L;
    const uint32_t W = L.words(), t = E.t;
    if (E.tdev > 64 || E.tdev > t) throw std::logic_error("terms not placed");
    E.nfw = (E.nf + 63) / 64;
    E.b_fixed_flat.assign((size_t)E.nf * E.nfw, 0);
    for (uint32_t i = 0; i < E.nf; ++i) std::copy(E.b_fixed[i].w.begin(), E.b_fixed[i].w.end(), (E.b_fixed_flat.data() + (size_t)i * E.nfw));
    E.log2mag = (double)L.g + (double)L.neven + (double)L.nodd / 2.0;   // normal_form.rs:465-467

    // Rotation terms.  Reference-style FP64 multipliers (engine.rs:221-238) for
    // the plain kernel, and the factorised form for the Z8 kernel.
    E.a0.clear(); E.a1.clear(); E.up.clear(); E.down.clear(); E.r0.clear(); E.r1.clear();
    E.Bmag = 1.0; E.neg_mask = 0;
    E.neg_rest.assign(t - E.tdev, 0);
    uint64_t tsum = 0;
    for (uint32_t k = 0; k < t; ++k) {
        const Phase th = E.theta[k];
        std::complex<double> e = phase_complex(th), one(1, 0);
        std::complex<double> a0 = (one + e) * 0.5, a1 = (one - e) * 0.5;
        E.a0.push_back(a0); E.a1.push_back(a1);
        // num-complex division formula (conj multiply over the squared norm)
        auto cdiv = [](std::complex<double> x, std::complex<double> y) {
            double n = y.real() * y.real() + y.imag() * y.imag();
            return std::complex<double>((x.real() * y.real() + x.imag() * y.imag()) / n, (x.imag() * y.real() - x.real() * y.imag()) / n);
        };
        E.up.push_back(cdiv(a1, a0)); E.down.push_back(cdiv(a0, a1));
        double half = M_PI * (double)th / 4294967296.0;   // theta/2 in (0, pi)
        double cs = std::fabs(std::cos(half)), sn = std::sin(half);
        double mx = std::max(cs, sn);
        E.r0.push_back(cs / mx); E.r1.push_back(sn / mx);
        E.Bmag *= mx;
        if (th > kPi) {
            if (k < E.tdev) E.neg_mask |= 1ull << k;
            else E.neg_rest[k - E.tdev] = 1;
        }
        tsum += th;
    }
    {
        // E = e^{i sum theta_k / 2}: the half-angle sum is periodic mod 2^33 units
        uint64_t hs = tsum & ((1ull << 33) - 1);
        double ang = M_PI * (double)hs / 4294967296.0;
        E.Ephase = std::complex<double>(std::cos(ang), std::sin(ang));
    }

    // Z8 quadratic form of the per-branch phase over the kernel coordinates.
    const uint32_t d = E.d;
    E.Q.assign(d, 0);
    E.lam.assign(d, 0);
    E.dual.assign((size_t)d * W, 0);
    for (uint32_t j = 0; j < d; ++j) {
        const uint64_t* U = (E.kernel_u.data() + (size_t)j * W);
        E.lam[j] = z8_lambda(E, U, E.kernel[j]);
        z8_dual(E, U, (E.dual.data() + (size_t)j * W));
        for (uint32_t k = j + 1; k < d; ++k)
            if (z8_cross(E, U, E.kernel[j], (E.kernel_u.data() + (size_t)k * W), E.kernel[k])) { E.Q[j] |= 1ull << k; E.Q[k] |= 1ull << j; }
    }

    E.mag = std::pow(2.0, E.log2mag - (double)E.nh / 2.0) * E.Bmag;
    E.ucol_dq.assign(E.nf, 0);
    for (uint32_t f = 0; f < E.nf; ++f)
        for (uint32_t j = 0; j < d; ++j)
            if (row_parity((E.ucols.data() + (size_t)f * W), (E.dual.data() + (size_t)j * W), W)) E.ucol_dq[f] |= 1ull << j;
    E.mcol_dq.assign(t, 0);
    for (uint32_t k = 0; k < t; ++k)
        for (uint32_t j = 0; j < d; ++j)
            if (row_parity((E.mcols.data() + (size_t)k * W), (E.dual.data() + (size_t)j * W), W) ^ (k < E.tdev && ((E.kernel[j] >> k) & 1) != 0))
                E.mcol_dq[k] |= 1ull << j;
}

// -------------------------------------------------------------- prologue --
// The prologue from the fixed values packed as a bit row over the E.nf fixed
// positions (nfw words).
static bool prepare_ones(const HostEngine& E, const uint64_t* __restrict__ on, Prepared& P);

bool prepare(const HostEngine& E, const uint8_t* values, Prepared& P) {
    thread_local std::vector<uint64_t> ones_tls;
    std::vector<uint64_t>& ones = ones_tls;             // one TLS lookup per call
    ones.assign(E.nfw, 0);
    for (uint32_t i = 0; i < E.nf; ++i) ones[i >> 6] |= (uint64_t)(values[i] & 1) << (i & 63);
    return prepare_ones(E, ones.data(), P);
}

BoundaryProg boundary_prog(const PathSum& ps, const HostEngine& E) {
    BoundaryProg B;
    B.nq = ps.nq;
    B.in_pos.assign(ps.nq, 0);
    B.out_pos.assign(ps.nq, -1);
    B.out_neg.assign(ps.nq, 0);
    B.out_chk.assign(ps.nq, -1);
    std::vector<int32_t> owner(ps.nv, -1);   // source of a variable's value: input q (q) or output q (64 + q)
    for (uint32_t q = 0; q < ps.nq; ++q) {
        if (E.fixed_pos[q] < 0) throw std::logic_error("input variable is not fixed in the engine");
        B.in_pos[q] = (uint32_t)E.fixed_pos[q];
        owner[q] = (int32_t)q;
    }
    for (uint32_t q = 0; q < ps.nq; ++q) {
        const uint32_t v = ps.out_var[q];
        B.out_neg[q] = ps.out_neg[q];
        if (owner[v] >= 0) { B.out_chk[q] = owner[v]; continue; }
        if (E.fixed_pos[v] < 0) throw std::logic_error("output variable is not fixed in the engine");
        B.out_pos[q] = E.fixed_pos[v];
        owner[v] = (int32_t)(64 + q);
    }
    // runs of consecutive qubits on consecutive positions within one word
    auto add_bit = [&](uint8_t src, uint32_t q, uint32_t pos) {
        if (!B.runs.empty()) {
            auto& r = B.runs.back();
            if (r.src == src && r.q0 + r.len == q && r.word == pos >> 6 && r.shift + r.len == (pos & 63)) { ++r.len; return; }
        }
        B.runs.push_back(BoundaryProg::Run{src, (uint8_t)q, 1, (uint8_t)(pos >> 6), (uint8_t)(pos & 63)});
    };
    for (uint32_t q = 0; q < ps.nq; ++q) add_bit(0, q, B.in_pos[q]);
    for (uint32_t q = 0; q < ps.nq; ++q) {
        if (B.out_neg[q]) B.yneg |= 1ull << q;
        if (B.out_chk[q] >= 0) B.checks.push_back({q, B.out_chk[q]});
        else add_bit(1, q, (uint32_t)B.out_pos[q]);
    }
    return B;
}

bool prepare_bits(const HostEngine& E, const BoundaryProg& B, uint64_t x, uint64_t y, Prepared& P, bool* contradictory) {
    // pathsum.rs:430-450 with the (variable, value) pairs written straight
    // into the fixed-position bit row
    if (E.nfw <= 2 && B.nq <= 64) {
        // the fixed row from the runs (at most two words: nf <= 2 nq)
        uint64_t on[2] = {0, 0};
        const uint64_t yy = y ^ B.yneg;
        for (const auto& r : B.runs) on[r.word] |= (((r.src ? yy : x) >> r.q0) & (~0ull >> (64 - r.len))) << r.shift;
        for (const auto& c : B.checks) {
            const uint64_t prev = c.second < 64 ? ((x >> c.second) & 1) : ((yy >> (c.second - 64)) & 1);
            if (prev != ((yy >> c.first) & 1)) { P.nonzero = false; if (contradictory) *contradictory = true; return false; }
        }
        if (contradictory) *contradictory = false;
        return prepare_ones(E, on, P);
    }
    thread_local std::vector<uint64_t> ones_tls;
    std::vector<uint64_t>& ones = ones_tls;
    ones.assign(E.nfw, 0);
    uint64_t* on = ones.data();
    // data-dependent bits are OR-ed in without branches
    for (uint32_t q = 0; q < B.nq; ++q) on[B.in_pos[q] >> 6] |= ((x >> q) & 1) << (B.in_pos[q] & 63);
    for (uint32_t q = 0; q < B.nq; ++q) {
        const uint64_t v = ((y >> q) & 1) ^ B.out_neg[q];
        const int32_t c = B.out_chk[q];
        if (c >= 0) {
            const uint64_t prev = c < 64 ? ((x >> c) & 1) : (((y >> (c - 64)) & 1) ^ B.out_neg[c - 64]);
            if (prev != v) { P.nonzero = false; if (contradictory) *contradictory = true; return false; }
        } else {
            on[B.out_pos[q] >> 6] |= v << (B.out_pos[q] & 63);
        }
    }
    if (contradictory) *contradictory = false;
    return prepare_ones(E, on, P);
}

static bool prepare_ones(const HostEngine& E, const uint64_t* __restrict__ on, Prepared& P) {
    P.nonzero = false;
    if (E.zero) return false;                                   // engine.rs:570-572
    const ULayout& L = E.L;
    const uint32_t W = L.words();
    uint32_t z = 0;
    uint64_t dq = 0;
    Phase th = E.global;
    P.u0.assign(W, 0);
    // Z4 value of the fixed assignment: linear part plus 2 per adjacent pair of
    // ones.  b_fixed is symmetric with a zero diagonal, so twice the pair count
    // is sum over ones a of |B_a & ones|.
    const uint32_t nfw = E.nfw;
    uint64_t* __restrict__ u0 = P.u0.data();
    for (uint32_t wi = 0; wi < nfw; ++wi) {
        for (uint64_t x = on[wi]; x; x &= x - 1) {
            const uint32_t f = wi * 64 + (uint32_t)__builtin_ctzll(x);
            const uint64_t* __restrict__ b = (E.b_fixed_flat.data() + (size_t)f * nfw);
            uint32_t pc = 0;
            for (uint32_t k = 0; k < nfw; ++k) pc += (uint32_t)__builtin_popcountll(b[k] & on[k]);
            z += E.fixed_v[f] + pc;
            th += E.fixed_theta[f];
            dq ^= E.ucol_dq[f];
            const uint64_t* __restrict__ c = (E.ucols.data() + (size_t)f * W);
            for (uint32_t k = 0; k < W; ++k) u0[k] ^= c[k];
        }
    }
    P.phase = th + z4(z);
    // kappa = w_e + u0_e; particular solution of K s = kappa (engine.rs:188-201)
    thread_local Bits kappa_tls;
    Bits& kappa = kappa_tls;
    if (kappa.n != L.neven) kappa = Bits(L.neven);
    for (uint32_t k = 0; k < L.We; ++k) kappa.w[k] = E.diag[L.off_e() + k] ^ P.u0[L.off_e() + k];
    for (auto& r : E.consistency) if (r.dot(kappa)) return false;
    uint64_t s0 = 0;
    P.s0_rest.clear();
    for (uint32_t i = 0; i < E.rk; ++i)
        if (E.pivot_row[i].dot(kappa)) {
            if (i < E.rkdev) s0 |= 1ull << E.pivot_col[i];
            else P.s0_rest.push_back(E.pivot_col[i]);
        }
    P.s0 = s0;
    P.u_start = P.u0;
    uint64_t* __restrict__ us = P.u_start.data();
    auto add_term = [&](uint32_t k) {
        const uint64_t* __restrict__ c = (E.mcols.data() + (size_t)k * W);
        for (uint32_t w = 0; w < W; ++w) us[w] ^= c[w];
        dq ^= E.mcol_dq[k];
    };
    for (uint64_t b = s0; b; b &= b - 1) add_term((uint32_t)__builtin_ctzll(b));
    for (uint32_t k : P.s0_rest) add_term(k);
    P.K0 = total_phase(E, us, s0);
    // ell_j = lam_j + 4 (parity(u_start . dual_j) ^ parity(s0 . kernel_j)):
    // the query part is the mask dq (prepared_ell spells out ell)
    P.dq = dq;
    P.scale = phase_complex(P.phase) * E.Ephase * E.mag;
    if (E.tdev < E.t) {
        // terms no kernel vector flips keep their s0 value on every branch:
        // their phase 6 s_k + 4 [theta_k > pi, s_k = 0] and magnitude r_k[s_k]
        // are constants of the query
        thread_local std::vector<uint8_t> in_s;
        in_s.assign(E.t - E.tdev, 0);
        for (uint32_t k : P.s0_rest) in_s[k - E.tdev] = 1;
        uint32_t kr = 0;
        double mr = 1.0;
        for (uint32_t k = E.tdev; k < E.t; ++k) {
            const bool sk = in_s[k - E.tdev];
            kr += sk ? 6u : (E.neg_rest[k - E.tdev] ? 4u : 0u);
            mr *= sk ? E.r1[k] : E.r0[k];
        }
        P.K0 = (P.K0 + kr) & 7u;
        P.scale *= mr;
    }
    P.nonzero = true;
    return true;
}

bool prepare_start(const HostEngine& E, const uint64_t* u, const uint64_t* s_ref, Prepared& P) {
    P.nonzero = false;
    if (E.zero) return false;
    const uint32_t W = E.L.words();
    P.phase = 0;
    P.u0.assign(u, u + W);
    P.u_start.assign(u, u + W);
    uint64_t s0 = 0;
    P.s0_rest.clear();
    for (uint32_t k = 0; k < E.t; ++k) {
        if (!((s_ref[k >> 6] >> (k & 63)) & 1)) continue;
        const uint32_t i = E.term_pos[k];
        if (i < E.tdev) s0 |= 1ull << i;
        else P.s0_rest.push_back(i);
    }
    // The walk's tables take the start on the coset's point whose non-pivot
    // columns are all 0, as prepare's particular solution is: kernel vector
    // j is the only one with a 1 in its column free_col[j], so XOR-ing in
    // the kernel vectors whose column is set moves the start there without
    // leaving the coset (the sum over the coset is the same).
    uint64_t* us = P.u_start.data();
    for (uint32_t j = 0; j < E.d; ++j) {
        if (!((s0 >> E.free_col[j]) & 1)) continue;
        s0 ^= E.kernel[j];
        const uint64_t* ku = (E.kernel_u.data() + (size_t)j * W);
        for (uint32_t w = 0; w < W; ++w) us[w] ^= ku[w];
    }
    u = us;
    P.s0 = s0;
    // ell_j = lam_j + 4 (parity(u . dual_j) ^ parity(s0 . kernel_j)), the
    // parities prepare accumulates column by column
    uint64_t dq = 0;
    for (uint32_t j = 0; j < E.d; ++j)
        if (row_parity(u, (E.dual.data() + (size_t)j * W), W) ^ ((__builtin_popcountll(s0 & E.kernel[j]) & 1) != 0)) dq |= 1ull << j;
    P.dq = dq;
    P.K0 = total_phase(E, u, s0);
    P.scale = E.Ephase * E.Bmag;
    if (E.tdev < E.t) {
        // the constant terms, as in prepare
        std::vector<uint8_t> in_s(E.t - E.tdev, 0);
        for (uint32_t k : P.s0_rest) in_s[k - E.tdev] = 1;
        uint32_t kr = 0;
        double mr = 1.0;
        for (uint32_t k = E.tdev; k < E.t; ++k) {
            const bool sk = in_s[k - E.tdev];
            kr += sk ? 6u : (E.neg_rest[k - E.tdev] ? 4u : 0u);
            mr *= sk ? E.r1[k] : E.r0[k];
        }
        P.K0 = (P.K0 + kr) & 7u;
        P.scale *= mr;
    }
    P.nonzero = true;
    return true;
}

std::complex<double> weight_ref(const HostEngine& E, const uint64_t* s_ref) {
    std::complex<double> c(1, 0);
    for (uint32_t k = 0; k < E.t; ++k) {
        const uint32_t i = E.term_pos[k];
        c *= ((s_ref[k >> 6] >> (k & 63)) & 1) ? E.a1[i] : E.a0[i];
    }
    return c;
}

void prepared_ell(const HostEngine& E, Prepared& P) {
    P.ell.resize(E.d);
    for (uint32_t j = 0; j < E.d; ++j) P.ell[j] = (uint8_t)((E.lam[j] + (uint32_t)((P.dq >> j) & 1) * 4) & 7);
}

std::complex<double> plain_c_start(const HostEngine& E, uint64_t s0) {
    std::complex<double> c(1, 0);
    for (uint32_t k = 0; k < E.tdev; ++k) c *= ((s0 >> k) & 1) ? E.a1[k] : E.a0[k];
    return c;
}

std::complex<double> plain_c_rest(const HostEngine& E, const Prepared& P) {
    std::complex<double> c(1, 0);
    if (E.tdev == E.t) return c;
    std::vector<uint8_t> in_s(E.t - E.tdev, 0);
    for (uint32_t k : P.s0_rest) in_s[k - E.tdev] = 1;
    for (uint32_t k = E.tdev; k < E.t; ++k) c *= in_s[k - E.tdev] ? E.a1[k] : E.a0[k];
    return c;
}

bool prepare_assignment(const HostEngine& E, const std::vector<std::pair<uint32_t, uint8_t>>& asg, Prepared& P) {
    thread_local std::vector<uint8_t> vals_tls;
    std::vector<uint8_t>& vals = vals_tls;
    vals.assign(E.nf, 0);
    size_t seen = 0;
    for (auto& a : asg) {
        int32_t p = a.first < E.fixed_pos.size() ? E.fixed_pos[a.first] : -1;
        if (p < 0) throw std::invalid_argument("variable is not fixed in this engine");   // engine.rs:527
        vals[p] = a.second; seen++;
    }
    if (seen < E.nf) throw std::invalid_argument("assignment does not cover all fixed variables");  // engine.rs:531
    return prepare(E, vals.data(), P);
}

// ---------------------------------------------------------------- rank.rs --
std::vector<std::pair<uint32_t, uint32_t>> pair_candidates(const PathSum& ps, const std::vector<uint32_t>& fixed, uint32_t& n_tclass) {
    const uint32_t M30 = (1u << 30) - 1;
    std::map<uint32_t, std::vector<uint32_t>> groups;   // ascending canonical residue (rank.rs:62-84)
    for (uint32_t k = 0; k < ps.nv; ++k) {
        if (ps.dead[k] || ps.theta[k] == 0 || std::find(fixed.begin(), fixed.end(), k) != fixed.end()) continue;
        uint32_t r = ps.theta[k] & M30;
        uint32_t o = (uint32_t)(((uint64_t)M30 + 1 - r) & M30);
        groups[std::min(r, o)].push_back(k);
    }
    std::vector<std::pair<uint32_t, uint32_t>> tp, gp;
    for (auto& g : groups) {
        auto& dst = g.first == (1u << 29) ? tp : gp;
        for (size_t i = 0; i + 1 < g.second.size(); i += 2) dst.push_back({g.second[i], g.second[i + 1]});
    }
    n_tclass = (uint32_t)tp.size();
    tp.insert(tp.end(), gp.begin(), gp.end());
    return tp;
}

// The same decision with GPU costs (the cost model of rank.rs:205-222 refit):
// each of the 2^p engines costs a host build (normal form, O(m^3/64) word ops
// at ~1 ns each plus a fixed upload), and every branch costs the measured
// walk rate of the Z8 kernel instead of the CPU's per-branch word operations.
uint32_t best_pairs_gpu(const EngineStats& st, uint32_t avail, uint32_t tclass, double branch_s) {
    const double m = (double)st.free_vars;
    const double build_s = 2e-4 + 1e-9 * (2.0 * m * m * m / 64.0 + 50.0 * m);
    uint32_t best = 0; double bc = INFINITY;
    for (uint32_t p = 0; p <= std::min(avail, kMaxPairs); ++p) {
        const uint32_t removed = 2 * std::min(p, tclass) + (p > tclass ? p - tclass : 0);
        const double dd = st.branch_dim > removed ? (double)(st.branch_dim - removed) : 0.0;
        const double cost = std::pow(2.0, (double)p) * ((p ? build_s : 0.0) + std::pow(2.0, dd) * branch_s);
        if (cost < bc) { bc = cost; best = p; }
    }
    return best;
}

uint32_t best_pairs(const EngineStats& st, uint32_t avail, uint32_t tclass) {   // rank.rs:205-222
    double m = (double)st.free_vars, words = (double)(st.free_vars / 64 + 1);
    double ldl = 2.0 * m * m * m / 64.0 + 50.0 * m + 2000.0;
    double step = 20.0 + 4.0 * words + (double)st.terms / 2.0;
    uint32_t best = 0; double bc = INFINITY;
    for (uint32_t p = 0; p <= std::min(avail, kMaxPairs); ++p) {
        uint32_t removed = 2 * std::min(p, tclass) + (p > tclass ? p - tclass : 0);
        double dd = st.branch_dim > removed ? (double)(st.branch_dim - removed) : 0.0;
        double cost = std::pow(2.0, (double)p) * ((p ? ldl : 0.0) + std::pow(2.0, dd) * step);
        if (cost < bc) { bc = cost; best = p; }
    }
    return best;
}

}  // namespace pq
