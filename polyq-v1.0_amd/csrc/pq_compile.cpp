// Circuit -> path-sum compiler (host, once per circuit).
// Gate rules follow src/pathsum.rs:136-224: X and SWAP are free (wire negation
// flags and relabelling), CX = H.CZ.H, rotations on one path variable merge and
// fold into the Z4 form when they become Clifford.
#include "pq_host.hpp"
#include <cmath>
#include <algorithm>

namespace pq {

static const double kS2 = 0.70710678118654752440;

std::complex<double> phase_complex(Phase p) {
    // exact eighth roots of unity on the pi/4 grid (phase.rs:60-66, :96-108)
    static const std::complex<double> root8[8] = {{1, 0}, {kS2, kS2}, {0, 1}, {-kS2, kS2}, {-1, 0}, {-kS2, -kS2}, {0, -1}, {kS2, -kS2}};
    if ((p & ((1u << 29) - 1)) == 0) return root8[p >> 29];
    double t = 2.0 * M_PI * (double)p / 4294967296.0;
    return {std::cos(t), std::sin(t)};
}

Phase phase_from_radians(double r) {
    double turns = r / (2.0 * M_PI);
    double frac = turns - std::floor(turns);
    return (Phase)((uint64_t)std::llround(frac * 4294967296.0) & 0xFFFFFFFFull);
}

void Circuit::add(uint8_t op, uint32_t qa, uint32_t qb, Phase ph) {
    if (op >= G_COUNT) throw std::invalid_argument("unknown gate code");
    if (op != G_GPHASE) n = std::max(n, qa + 1);
    if (op == G_CZ || op == G_CX || op == G_SWAP) {
        if (qa == qb) throw std::invalid_argument("two-qubit gate on one wire");
        n = std::max(n, qb + 1);
    }
    gates.push_back(Gate{op, qa, qb, ph});
}

uint32_t Circuit::hadamards() const {
    uint32_t c = 0;
    for (auto& g : gates) c += g.op == G_H ? 1u : (g.op == G_CX ? 2u : 0u);
    return c;
}

Circuit Circuit::inverse() const {
    Circuit r; r.n = n;
    r.gates.reserve(gates.size());
    for (auto it = gates.rbegin(); it != gates.rend(); ++it) {
        Gate g = *it;
        if (g.op == G_S) g.op = G_SDG; else if (g.op == G_SDG) g.op = G_S;
        else if (g.op == G_T) g.op = G_TDG; else if (g.op == G_TDG) g.op = G_T;
        else if (g.op == G_P || g.op == G_GPHASE) g.ph = 0u - g.ph;
        r.gates.push_back(g);
    }
    return r;
}

namespace {
// Incremental compiler state (pathsum.rs:85-267).  Variables get adjacency rows
// sized for the final variable count, known up front from the gate list.
struct Builder {
    uint32_t n, cap, nv;
    std::vector<uint32_t> wire;
    std::vector<uint8_t> neg;
    std::vector<Bits> adj;
    std::vector<uint8_t> v;
    std::vector<Phase> th;
    uint32_t nh = 0;
    Phase global = 0;

    Builder(uint32_t nq, uint32_t extra) : n(nq), cap(nq + extra), nv(nq), wire(nq), neg(nq, 0), adj(nq + extra, Bits(nq + extra)), v(nq, 0), th(nq, 0) {
        for (uint32_t q = 0; q < nq; ++q) wire[q] = q;
    }
    uint32_t fresh() {
        if (nv == cap) throw std::logic_error("variable capacity exceeded");
        v.push_back(0); th.push_back(0);
        return nv++;
    }
    void link(uint32_t a, uint32_t b) { adj[a].flip(b); adj[b].flip(a); }
    void addv(uint32_t z, uint32_t k) { v[z] = (uint8_t)((v[z] + k) & 3); }
    void rot(uint32_t q, Phase p) {
        uint32_t z = wire[q];
        if (neg[q]) { th[z] -= p; global += p; } else th[z] += p;
    }
    void hadamard(uint32_t q) {
        uint32_t z = wire[q], zp = fresh();
        link(z, zp); nh++;
        if (neg[q]) { addv(zp, 2); neg[q] = 0; }   // (-1)^{(z+1) z'} = (-1)^{z z'} (-1)^{z'}
        wire[q] = zp;
    }
    void cz(uint32_t a, uint32_t b) {
        link(wire[a], wire[b]);
        if (neg[b]) addv(wire[a], 2);
        if (neg[a]) addv(wire[b], 2);
        if (neg[a] && neg[b]) global += kPi;
    }
    void zgate(uint32_t q) { addv(wire[q], 2); if (neg[q]) global += kPi; }
    void apply(const Gate& g) {
        if (g.op != G_GPHASE && (g.a >= n || ((g.op == G_CZ || g.op == G_CX || g.op == G_SWAP) && g.b >= n)))
            throw std::invalid_argument("gate touches a wire beyond the circuit");
        switch (g.op) {
            case G_H: hadamard(g.a); break;
            case G_X: neg[g.a] ^= 1; break;
            case G_Z: zgate(g.a); break;
            case G_S: if (neg[g.a]) { addv(wire[g.a], 3); global += kHalfPi; } else addv(wire[g.a], 1); break;
            case G_SDG: if (neg[g.a]) { addv(wire[g.a], 1); global -= kHalfPi; } else addv(wire[g.a], 3); break;
            case G_T: rot(g.a, kQuarterPi); break;
            case G_TDG: rot(g.a, 0u - kQuarterPi); break;
            case G_P: rot(g.a, g.ph); break;
            case G_Y: zgate(g.a); neg[g.a] ^= 1; global += kHalfPi; break;   // Y = i X Z
            case G_CZ: cz(g.a, g.b); break;
            case G_CX: hadamard(g.b); cz(g.a, g.b); hadamard(g.b); break;
            case G_SWAP: std::swap(wire[g.a], wire[g.b]); std::swap(neg[g.a], neg[g.b]); break;
            case G_GPHASE: global += g.ph; break;
        }
    }
    PathSum finish() const {
        PathSum ps;
        ps.nq = n; ps.nv = nv; ps.nh = nh;
        ps.adj.assign(nv, Bits(nv));
        for (uint32_t i = 0; i < nv; ++i) adj[i].each([&](uint32_t j) { if (j < nv) ps.adj[i].set(j); });
        ps.v = v; ps.theta = th; ps.global = global;
        ps.dead.assign(nv, 0);
        ps.out_var = wire; ps.out_neg = neg;
        // fold Clifford angles into the Z4 form (pathsum.rs:317-324)
        for (uint32_t k = 0; k < nv; ++k) {
            int l = clifford_level(ps.theta[k]);
            if (l >= 0) { ps.v[k] = (uint8_t)((ps.v[k] + l) & 3); ps.theta[k] = 0; }
        }
        return ps;
    }
};
}  // namespace

PathSum compile(const Circuit& c) {
    Builder b(c.n, c.hadamards());
    for (auto& g : c.gates) b.apply(g);
    return b.finish();
}

Sandwich compile_sandwich(const Circuit& c) {
    Builder b(c.n, 2 * c.hadamards());
    for (auto& g : c.gates) b.apply(g);
    Sandwich s;
    s.junction = b.wire; s.jneg = b.neg;
    Circuit inv = c.inverse();
    for (auto& g : inv.gates) b.apply(g);
    s.ps = b.finish();
    return s;
}

std::vector<uint32_t> PathSum::io_vars() const {
    std::vector<uint32_t> f(nq);
    for (uint32_t q = 0; q < nq; ++q) f[q] = q;
    for (uint32_t o : out_var) if (std::find(f.begin(), f.end(), o) == f.end()) f.push_back(o);
    return f;
}

void PathSum::substitute(uint32_t a, uint32_t b, bool c) {
    if (a == b || dead[a] || dead[b]) throw std::invalid_argument("substitute: variables must be live and distinct");
    bool bab = adj[a].get(b);
    uint8_t vb = v[b];
    Phase tb = theta[b];
    Bits nb = adj[b];
    nb.put(a, false);
    if (!c) {
        v[a] = (uint8_t)((v[a] + vb + (bab ? 2 : 0)) & 3);
        theta[a] += tb;
    } else {
        v[a] = (uint8_t)((v[a] + 4 - vb) & 3);
        theta[a] -= tb;
        global += z4(vb) + tb;
        nb.each([&](uint32_t k) { v[k] = (uint8_t)((v[k] + 2) & 3); });
    }
    nb.each([&](uint32_t k) { adj[a].flip(k); adj[k].flip(a); });
    adj[b].each([&](uint32_t k) { adj[k].put(b, false); });
    adj[b].clear();
    v[b] = 0; theta[b] = 0; dead[b] = 1;
    int l = clifford_level(theta[a]);
    if (l >= 0) { v[a] = (uint8_t)((v[a] + l) & 3); theta[a] = 0; }
}

// (var, value) pairs put into out; false if a variable gets two values.
// extra_var / extra_val append k more pairs (the marginal junction).
bool PathSum::boundary_bits(uint64_t x, uint64_t y, std::vector<std::pair<uint32_t, uint8_t>>& out, size_t k, const uint32_t* extra_var,
                            const uint8_t* extra_val) const {
    out.clear();
    // value already given to a variable: 0 / 1, or 255 if none yet
    thread_local std::vector<uint8_t> given_tls;
    std::vector<uint8_t>& given_v = given_tls;          // one TLS lookup per call
    if (given_v.size() < nv) given_v.assign(nv, 255);
    uint8_t* given = given_v.data();
    out.reserve(2 * nq + k);
    bool ok = true;
    auto put = [&](uint32_t var, uint8_t val) {
        if (given[var] != 255) { ok = ok && given[var] == val; return; }
        given[var] = val;
        out.push_back({var, val});
    };
    for (uint32_t q = 0; q < nq; ++q) put(q, (uint8_t)((x >> q) & 1));
    for (uint32_t q = 0; q < nq; ++q) put(out_var[q], (uint8_t)(((y >> q) & 1) ^ out_neg[q]));
    for (size_t i = 0; i < k; ++i) put(extra_var[i], extra_val[i]);
    for (auto& e : out) given[e.first] = 255;
    if (!ok) out.clear();
    return ok;
}

bool PathSum::boundary(const uint8_t* x, const uint8_t* y, std::vector<std::pair<uint32_t, uint8_t>>& out, size_t k, const uint32_t* extra_var,
                       const uint8_t* extra_val) const {
    if (nq <= 64) {
        uint64_t xb = 0, yb = 0;
        for (uint32_t q = 0; q < nq; ++q) { xb |= (uint64_t)(x[q] ? 1 : 0) << q; yb |= (uint64_t)(y[q] ? 1 : 0) << q; }
        return boundary_bits(xb, yb, out, k, extra_var, extra_val);
    }
    // any number of qubits: the same (variable, value) pairs from the bytes
    out.clear();
    std::vector<uint8_t> given(nv, 255);
    out.reserve(2 * nq + k);
    bool ok = true;
    auto put = [&](uint32_t var, uint8_t val) {
        if (given[var] != 255) { ok = ok && given[var] == val; return; }
        given[var] = val;
        out.push_back({var, val});
    };
    for (uint32_t q = 0; q < nq; ++q) put(q, x[q] ? 1 : 0);
    for (uint32_t q = 0; q < nq; ++q) put(out_var[q], (uint8_t)((y[q] ? 1 : 0) ^ out_neg[q]));
    for (size_t i = 0; i < k; ++i) put(extra_var[i], extra_val[i]);
    if (!ok) out.clear();
    return ok;
}

}  // namespace pq
