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
