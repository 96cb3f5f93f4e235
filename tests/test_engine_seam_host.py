"""Engine-level seam on the host: engines built from the oracle's Engine
fields (pq_engine_create) have the statistics of the library's own build of
the same circuit, and inconsistent descriptors are refused with an error.
No device is needed to create an engine."""
import numpy as np
import pytest

import oracle_lib as O
import polyq_amd as pq
from polyq_amd import workloads as W


def test_engines_from_oracle_fields_match_library_stats():
    rng = W.Rng(31)
    built = 0
    for trial in range(40):
        n = 3 + rng.below(10)
        c = W.random_clifford_t(n, 20 + rng.below(80), rng.below(12), trial % 4 == 0, rng)
        o = O.OSim(O.OCircuit.from_circuit(c), "cached")
        f = O.engine_fields(o)
        e = pq.Engine(f)
        assert e.stats() == pq.Simulator(c, strategy="cached").stats()[:6], trial
        built += 1
    for seed in (1, 2, 3):
        c, _, _ = W.config2(seed)
        f = O.engine_fields(O.OSim(O.OCircuit.from_circuit(c), "cached"))
        assert pq.Engine(f).stats() == pq.Simulator(c, strategy="cached").stats()[:6]
        built += 1
    assert built == 43


def _fields():
    c, _, _ = W.config2(1)
    return O.engine_fields(O.OSim(O.OCircuit.from_circuit(c), "cached"))


def test_descriptor_checks():
    f = _fields()
    assert f["rank"] >= 1 and f["branch_dim"] >= 2 and f["nf"] >= 2
    pq.Engine(dict(f))                          # the exported fields are consistent
    pq.Engine(dict(f, kernel_u=None))           # kernel_u is derived when absent

    def bad(match, **change):
        g = dict(f)
        g.update(change)
        with pytest.raises(pq.PolyqError, match=match):
            pq.Engine(g)

    ku = f["kernel_u"].copy(); ku[0] ^= 1
    bad("kernel_u differs", kernel_u=ku)
    bad("sigma", sigma=f["sigma"] + 2)
    bf = f["b_fixed"].copy()
    bf[0] ^= 2                                  # flips edge (0, 1) in row 0 only
    bad("not symmetric", b_fixed=bf)
    bad("RREF rank", rank=f["rank"] + 1)
    kn = f["kernel"].copy(); kn[0] |= 1 << 63
    bad("kernel vector", kernel=kn)
    o3 = f["o3"].copy(); o3[0] ^= 1
    bad("line", o3=o3)


def _kernel_rows(f):
    """The packed kernel vectors and their shift rows of an engine's fields."""
    t, d = f["terms"], f["branch_dim"]
    sw = (t + 63) // 64
    uw = 2 * ((f["planes"] + 63) // 64) + (f["lines"] + 63) // 64 + (f["radical"] + 63) // 64
    return f["kernel"][:d * sw].reshape(d, sw), f["kernel_u"][:d * uw].reshape(d, uw)


def test_oracle_sum_subspace_start_and_modulus():
    # The oracle's sum_subspace from prepare's own start gives |eval| once
    # multiplied by 2^(log2_magnitude - h/2) (engine.rs:557-565), and the
    # same sum from another start point of the subspace (u ^ kernel_u[j],
    # s ^ kernel[j], with that point's weight): the sum is over the coset,
    # whatever point the walk starts from.
    c, xs, ys = W.config1(1)
    o = O.OSim(O.OCircuit.from_circuit(c), "cached")
    f = O.engine_fields(o)
    K, KU = _kernel_rows(f)
    assert len(K) >= 3
    mag = 2.0 ** (f["planes"] + f["radical"] + f["lines"] / 2.0 - f["num_h"] / 2.0)
    checked = 0
    for x, y in zip(xs[:8], ys[:8]):
        v = O.engine_values(o, x, y)
        st = O.engine_prepare(o, 0, v)
        if st is None:
            continue
        u, s, cs = st
        assert abs(cs - O.engine_weight(o, 0, s)) <= 1e-15
        total = O.engine_sum_subspace(o, 0, u, s, cs)
        assert abs(abs(total) * mag - abs(O.engine_eval(o, 0, v))) <= 1e-12
        for j in (0, len(K) - 1):
            s2, u2 = s ^ K[j], u ^ KU[j]
            other = O.engine_sum_subspace(o, 0, u2, s2, O.engine_weight(o, 0, s2))
            assert abs(other - total) <= 1e-12 * max(1.0, abs(total))
        checked += 1
    assert checked >= 4


def test_sum_subspace_entry_checks_its_start():
    # pq_engine_sum_subspace refuses a c_start that is not the weight of
    # s_start, and branch bits beyond the terms, before any device work
    c, xs, ys = W.config1(1)
    o = O.OSim(O.OCircuit.from_circuit(c), "cached")
    e = pq.Engine(O.engine_fields(o))
    u, s, cs = O.engine_prepare(o, 0, O.engine_values(o, xs[0], ys[0]))
    with pytest.raises(pq.PolyqError, match="weight"):
        e.sum_subspace(u, s, cs * 1.001)
    t = e.stats()[4]
    bad = s.copy()
    bad[-1] |= np.uint64(1) << np.uint64(63) if t % 64 else np.uint64(0)
    if t % 64:
        with pytest.raises(pq.PolyqError, match="beyond"):
            e.sum_subspace(u, bad, cs)
    with pytest.raises(ValueError):
        e.sum_subspace(u[:0], s, cs)
