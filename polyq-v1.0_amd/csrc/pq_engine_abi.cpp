// Engine-level drop-in boundary: an Engine (src/engine.rs:273-284) handed to
// the GPU by the fields it already holds, with no circuit and no recompilation.
// A Rust RankEngine (src/rank.rs:47-51) keeps one pq_engine per branch engine
// and evaluates them with pq_engine_eval_batch; the sum over its branch
// engines is RankEngine::eval_assignment (src/rank.rs:184-190).
//
// The descriptor carries Core + Branching + Engine as evaluation needs them:
// the fixed values and fixed-fixed edges (prepare's boundary constants,
// engine.rs:569-586), the shift rows ucols / mcols and Diag of the normal form
// (normal_form.rs:161-218), the RREF of K (Elim: pivots, consistency rows,
// kernel, kernel_u, engine.rs:106-184) and the term angles (Term,
// engine.rs:221-238).  Everything else the kernels read is derived here.
#include "pq_api_internal.hpp"
#include "../../include/polyq_amd.h"
#include <cstring>

namespace pq {
namespace {

uint32_t words_of(uint32_t n) { return (n + 63) / 64; }

// A packed row has no bits at or beyond n (tail bits zero, bits.rs:10-19).
bool tail_clear(const uint64_t* w, uint32_t n) {
    const uint32_t nw = words_of(n);
    return nw == 0 || (n % 64) == 0 || (w[nw - 1] >> (n % 64)) == 0;
}

Bits bits_from(const uint64_t* w, uint32_t n) {
    Bits b(n);
    std::copy(w, w + b.w.size(), b.w.begin());
    return b;
}

std::shared_ptr<HostEngine> engine_from_desc(const pq_engine_desc& D) {
    auto Ep = std::make_shared<HostEngine>();
    HostEngine& E = *Ep;
    const uint32_t nf = D.nf, t = D.terms, d = D.branch_dim, rk = D.rank, ne = D.radical;
    if (rk + d != t) throw std::invalid_argument("RREF rank plus kernel dimension differs from the number of terms");
    if (rk + D.consistency_rows != ne) throw std::invalid_argument("RREF rank plus consistency rows differs from the radical size");
    if (!D.ucols && nf) throw std::invalid_argument("ucols missing");
    if (!D.mcols && t) throw std::invalid_argument("mcols missing");
    if (!D.diag || !D.o3) throw std::invalid_argument("diag rows missing");
    if (nf && (!D.fixed_v || !D.fixed_theta || !D.b_fixed)) throw std::invalid_argument("fixed-variable data missing");
    if (t && !D.theta) throw std::invalid_argument("term angles missing");
    if ((rk && (!D.pivot_col || !D.pivot_row)) || (D.consistency_rows && !D.consistency) || (d && !D.kernel))
        throw std::invalid_argument("RREF rows missing");
    ULayout& L = E.L;
    L.g = D.planes; L.nodd = D.lines; L.neven = ne;
    L.Wg = words_of(L.g); L.Wo = words_of(L.nodd); L.We = words_of(ne);
    const uint32_t W = L.words();
    auto check_urow = [&](const uint64_t* r, const char* what) {
        if (!tail_clear(r + L.off_a(), L.g) || !tail_clear(r + L.off_b(), L.g) || !tail_clear(r + L.off_o(), L.nodd) ||
            !tail_clear(r + L.off_e(), ne))
            throw std::invalid_argument(std::string(what) + " has bits beyond its part length");
    };
    E.m = D.free_vars; E.t = t; E.d = d; E.nf = nf; E.rk = rk;
    E.nh = D.num_h; E.global = D.global_phase; E.zero = D.zero != 0;
    // values arrive aligned with fixed() (engine.rs:428-430, :551)
    E.fixed.resize(nf);
    E.fixed_pos.resize(nf);
    for (uint32_t i = 0; i < nf; ++i) { E.fixed[i] = i; E.fixed_pos[i] = (int32_t)i; }
    E.fixed_v.assign(D.fixed_v, D.fixed_v + nf);
    E.fixed_theta.assign(D.fixed_theta, D.fixed_theta + nf);
    const uint32_t nfw = words_of(nf);
    E.b_fixed.assign(nf, Bits(nf));
    for (uint32_t i = 0; i < nf; ++i) {
        if (!tail_clear(D.b_fixed + (size_t)i * nfw, nf)) throw std::invalid_argument("b_fixed row has bits beyond nf");
        E.b_fixed[i] = bits_from(D.b_fixed + (size_t)i * nfw, nf);
        if (E.b_fixed[i].get(i)) throw std::invalid_argument("b_fixed has a nonzero diagonal");
    }
    for (uint32_t i = 0; i < nf; ++i)
        for (uint32_t j = 0; j < nf; ++j)
            if (E.b_fixed[i].get(j) != E.b_fixed[j].get(i)) throw std::invalid_argument("b_fixed is not symmetric");
    E.ucols.assign(D.ucols, D.ucols + (size_t)nf * W);
    for (uint32_t i = 0; i < nf; ++i) check_urow((E.ucols.data() + (size_t)i * W), "ucols row");
    E.mcols.assign(D.mcols, D.mcols + (size_t)t * W);
    for (uint32_t k = 0; k < t; ++k) check_urow((E.mcols.data() + (size_t)k * W), "mcols row");
    E.diag.assign(D.diag, D.diag + W);
    check_urow(E.diag.data(), "diag row");
    E.o3.assign(D.o3, D.o3 + L.Wo);
    if (!tail_clear(E.o3.data(), L.nodd)) throw std::invalid_argument("o3 has bits beyond the lines");
    int32_t sigma = 0;
    for (uint32_t w = 0; w < L.Wo; ++w) {
        const uint64_t o1 = E.diag[L.off_o() + w], o3 = E.o3[w];
        if (o1 & o3) throw std::invalid_argument("a line is both o1 and o3");
        const uint64_t all = (w + 1 < L.Wo || L.nodd % 64 == 0) ? ~0ull : ((1ull << (L.nodd % 64)) - 1);
        if ((o1 | o3) != all) throw std::invalid_argument("a line is neither o1 nor o3");
        sigma += __builtin_popcountll(o1) - __builtin_popcountll(o3);
    }
    if (sigma != D.sigma) throw std::invalid_argument("sigma differs from #o1 - #o3");
    E.sigma = sigma;
    // RREF of K (engine.rs:119-184): pivot columns increasing, R rows over the radical
    const uint32_t tw = words_of(t);
    std::vector<uint8_t> is_piv(t, 0);
    for (uint32_t i = 0; i < rk; ++i) {
        const uint32_t c = D.pivot_col[i];
        if (c >= t || (i && c <= D.pivot_col[i - 1])) throw std::invalid_argument("pivot columns must increase and lie below the number of terms");
        is_piv[c] = 1;
        E.pivot_col.push_back(c);
        if (!tail_clear(D.pivot_row + (size_t)i * L.We, ne)) throw std::invalid_argument("pivot row has bits beyond the radical");
        E.pivot_row.push_back(bits_from(D.pivot_row + (size_t)i * L.We, ne));
    }
    for (uint32_t r = 0; r < D.consistency_rows; ++r) {
        if (!tail_clear(D.consistency + (size_t)r * L.We, ne)) throw std::invalid_argument("consistency row has bits beyond the radical");
        E.consistency.push_back(bits_from(D.consistency + (size_t)r * L.We, ne));
    }
    // kernel vectors over all t terms, each with one non-pivot column, in
    // increasing order of that column (engine.rs:159-171)
    std::vector<Bits> kernel_full;
    std::vector<uint32_t> free_full;
    for (uint32_t j = 0; j < d; ++j) {
        const uint64_t* row = D.kernel + (size_t)j * tw;
        if (!tail_clear(row, t)) throw std::invalid_argument("kernel vector has bits beyond the terms");
        Bits n = bits_from(row, t);
        int32_t free_col = -1;
        n.each([&](uint32_t k) {
            if (is_piv[k]) return;
            if (free_col >= 0) throw std::invalid_argument("kernel vector with two non-pivot columns");
            free_col = (int32_t)k;
        });
        if (free_col < 0 || (j && (uint32_t)free_col <= free_full.back()))
            throw std::invalid_argument("kernel vectors must carry increasing non-pivot columns");
        kernel_full.push_back(n);
        free_full.push_back((uint32_t)free_col);
    }
    E.theta.assign(D.theta, D.theta + t);
    place_terms(E, kernel_full, free_full);
    // kernel_u[j] = xor of mcols over the support (engine.rs:172-182): the
    // e part vanishes inside ker K, and a supplied kernel_u must agree
    for (uint32_t j = 0; j < d; ++j) {
        const uint64_t* u = (E.kernel_u.data() + (size_t)j * W);
        for (uint32_t w = 0; w < L.We; ++w)
            if (u[L.off_e() + w]) throw std::invalid_argument("kernel vector is not in the kernel of K");
        if (D.kernel_u && std::memcmp(u, D.kernel_u + (size_t)j * W, (size_t)W * 8) != 0)
            throw std::invalid_argument("kernel_u differs from the xor of mcols over the kernel support");
    }
    finish_engine(E);
    return Ep;
}

}  // namespace
}  // namespace pq

using namespace pq;

struct pq_engine { std::shared_ptr<RankEng> r; };

#define PQ_TRY(...)                                                                                   \
    try { __VA_ARGS__ } catch (std::exception & e) { g_err = e.what(); return -1; } catch (...) { g_err = "unknown error"; return -1; }

namespace {
// Engine::eval (engine.rs:551-561) of query i = (engines[i], values[i]) for
// every i: the batched forms as one launch over all queries, the plain form
// one launch per query.  A vanishing prologue leaves the value 0.
void eval_engines(size_t n, pq_engine* const* engines, const uint8_t* const* values, int form, std::complex<double>* out) {
    for (size_t i = 0; i < n; ++i) {
        if (!engines[i]) throw std::invalid_argument("null engine");
        if (engines[i]->r->host[0]->nf && !values[i]) throw std::invalid_argument("null values");
        engines[i]->r->ensure_device();
        out[i] = {0, 0};
    }
    if (form == PQ_FORM_PLAIN_FP64) {
        for (size_t i = 0; i < n; ++i) {
            const RankEng& R = *engines[i]->r;
            const HostEngine& E = *R.host[0];
            Prepared P;
            if (!prepare(E, values[i], P)) continue;
            // engine.rs:560: phase * total * 2^{log2mag - h/2}
            out[i] = phase_complex(P.phase) * dev_plain_sum(*R.dev[0], P, nullptr) * std::pow(2.0, E.log2mag - (double)E.nh / 2.0);
        }
        return;
    }
    const int walk = walk_of_form(form);
    std::vector<Z8Job> jobs;
    std::vector<size_t> owner;
    Prepared P;
    for (size_t i = 0; i < n; ++i) {
        const RankEng& R = *engines[i]->r;
        if (!prepare(*R.host[0], values[i], P)) continue;
        Z8Job J;
        J.eng = R.dev[0].get();
        if (walk == kWalkSeg) J.sq = seg_query(*R.host[0], *R.seg[0], P);
        else J.q = z8_query(*R.host[0], *R.z8[0], P);
        J.scale = P.scale;
        J.mid_bits = 0;
        jobs.push_back(J);
        owner.push_back(i);
    }
    std::vector<std::complex<double>> res;
    run_jobs(jobs, owner, n, walk, res);
    for (size_t i = 0; i < n; ++i) out[i] = res[i];
}
}  // namespace

extern "C" {

int pq_engine_create(const pq_engine_desc* desc, pq_engine** out) {
    PQ_TRY({
        if (!desc || !out) throw std::invalid_argument("null descriptor or output handle");
        auto he = engine_from_desc(*desc);
        auto R = std::make_shared<RankEng>();
        auto z = std::make_shared<Z8Image>(build_z8_image(*he));
        auto g = std::make_shared<SegImage>(build_seg_image(*he));
        R->image_bytes = dev_engine_image_bytes(*he, *z, *g);
        R->host.push_back(he); R->z8.push_back(z); R->seg.push_back(g);
        auto h = std::make_unique<pq_engine>();
        h->r = R;
        *out = h.release();
        return 0;
    })
}

void pq_engine_destroy(pq_engine* e) { delete e; }

int pq_engine_stats(const pq_engine* e, uint64_t* o) {
    PQ_TRY({
        const EngineStats s = e->r->host[0]->stats();
        o[0] = s.free_vars; o[1] = s.planes; o[2] = s.lines; o[3] = s.radical; o[4] = s.terms; o[5] = s.branch_dim;
        return 0;
    })
}

int pq_engine_eval(pq_engine* e, const uint8_t* values, size_t nf, int form, double* out) {
    PQ_TRY({
        if (nf != e->r->host[0]->nf) throw std::invalid_argument("values do not match the engine's fixed variables");
        std::complex<double> v;
        eval_engines(1, &e, &values, form, &v);
        out[0] = v.real(); out[1] = v.imag();
        return 0;
    })
}

int pq_engine_eval_batch(size_t n, pq_engine* const* engines, const uint8_t* const* values, int form, double* out) {
    PQ_TRY({
        std::vector<std::complex<double>> v(n);
        eval_engines(n, engines, values, form, v.data());
        for (size_t i = 0; i < n; ++i) { out[2 * i] = v[i].real(); out[2 * i + 1] = v[i].imag(); }
        return 0;
    })
}

int pq_engine_sum_subspace(pq_engine* e, const uint64_t* u_start, const uint64_t* s_start, const double* c_start, int form, double* out) {
    PQ_TRY({
        if (!e || !u_start || !s_start || !c_start || !out) throw std::invalid_argument("null argument");
        const RankEng& R = *e->r;
        const HostEngine& E = *R.host[0];
        const ULayout& L = E.L;
        if (!tail_clear(u_start + L.off_a(), L.g) || !tail_clear(u_start + L.off_b(), L.g) || !tail_clear(u_start + L.off_o(), L.nodd) ||
            !tail_clear(u_start + L.off_e(), L.neven))
            throw std::invalid_argument("u_start has bits beyond its part lengths");
        if (!tail_clear(s_start, E.t)) throw std::invalid_argument("s_start has bits beyond the number of terms");
        // c_start is prepare's weight at s_start (engine.rs:600-606); the
        // walk derives every branch's weight from the term angles, as the
        // reference does for every chunk but the first (engine.rs:677-684)
        const std::complex<double> w = weight_ref(E, s_start), c(c_start[0], c_start[1]);
        if (std::abs(c - w) > 1e-9 * std::max(std::abs(w), 1e-300))
            throw std::invalid_argument("c_start is not the weight of s_start");
        out[0] = out[1] = 0.0;
        Prepared P;
        if (!prepare_start(E, u_start, s_start, P)) return 0;
        R.ensure_device();
        std::complex<double> v;
        if (form == PQ_FORM_PLAIN_FP64) {
            v = dev_plain_sum(*R.dev[0], P, nullptr);
        } else {
            const int walk = walk_of_form(form);
            Z8Job J;
            J.eng = R.dev[0].get();
            if (walk == kWalkSeg) J.sq = seg_query(E, *R.seg[0], P);
            else J.q = z8_query(E, *R.z8[0], P);
            J.scale = P.scale;
            J.mid_bits = 0;
            std::vector<Z8Job> jobs{J};
            std::vector<size_t> owner{0};
            std::vector<std::complex<double>> res;
            run_jobs(jobs, owner, 1, walk, res);
            v = res[0];
        }
        out[0] = v.real(); out[1] = v.imag();
        return 0;
    })
}

int pq_engine_branch_dump(pq_engine* e, const uint8_t* values, size_t n, const uint64_t* s, uint8_t* zero, uint8_t* k8, uint8_t* ktot,
                          uint32_t* r) {
    PQ_TRY({
        const RankEng& R = *e->r;
        const HostEngine& E = *R.host[0];
        if (E.zero) throw std::invalid_argument("path sum vanishes identically");
        if (E.tdev != E.t) throw std::invalid_argument("per-branch diagnostics need at most 64 non-Clifford terms");
        Prepared P;
        prepare(E, values, P);          // u0 is filled even when K s = kappa is inconsistent
        if (P.u0.empty()) throw std::logic_error("prepare did not produce the boundary shift");
        if (r) *r = 2 * E.L.g + 2 * E.L.neven + E.L.nodd;
        R.ensure_device();
        dev_dump_s(*R.dev[0], P, s, n, zero, k8, ktot);
        return 0;
    })
}

}  // extern "C"
