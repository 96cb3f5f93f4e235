"""Product host logic and the C ABI surface, without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import polyq_amd as pq
from polyq_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "polyq_amd.h")).read()
    names = set(re.findall(r"\b(pq_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(pq.lib, n), n


def test_generators_match_the_reference_generator():
    for seed in range(6):
        rng = W.Rng(seed)
        c = W.random_clifford_t(4, 30, 6, seed % 2 == 0, rng)
        oc, st = O.random_clifford_t((seed + W.GAMMA) & W.M64, 4, 30, 6, seed % 2 == 0)
        assert list(c.ops) == list(oc.ops) and list(c.q0) == list(oc.q0) and list(c.q1) == list(oc.q1) and list(c.ph) == list(oc.ph)
        assert st == rng.s
        rng = W.Rng(seed)
        c = W.random_layered(3, 2, rng)
        oc, st = O.random_layered((seed + W.GAMMA) & W.M64, 3, 2)
        assert list(c.ops) == list(oc.ops) and list(c.q0) == list(oc.q0) and st == rng.s


def test_engine_stats_match_the_oracle():
    rng = W.Rng(31)
    for trial in range(120):
        n = 1 + rng.below(6)
        c = W.random_clifford_t(n, 10 + rng.below(40), rng.below(12), trial % 3 == 0, rng) if trial % 2 else W.random_layered(n, 1 + rng.below(3), rng)
        oc = O.OCircuit.from_circuit(c)
        for strat in ("cached", "auto"):
            assert pq.Simulator(c, strategy=strat).stats() == O.OSim(oc, strat).stats()
    for cfg in (W.config1, W.config2):
        c, _, _ = cfg(5)
        assert pq.Simulator(c, strategy="cached").stats() == O.OSim(O.OCircuit.from_circuit(c), "cached").stats()


def test_z8_quadratic_form_matches_direct_phase():
    # k(c) = K0 + sum c_j l_j + 4 sum_{j<k} c_j c_k Q_jk  vs  phase8(u(c)) + 6|s| + 4|neg\s|
    rng = W.Rng(8)
    for trial in range(150):
        n = 1 + rng.below(6)
        c = W.random_clifford_t(n, 10 + rng.below(40), 1 + rng.below(14), trial % 2 == 0, rng)
        s = pq.Simulator(c, strategy="cached")
        for _ in range(4):
            bad = s.selfcheck(rng.below(1 << n), rng.below(1 << n), trials=64, seed=trial)
            assert bad in (None, 0)
    for cfg in (W.config2, W.config3):
        c, xs, ys = cfg(5) if cfg is W.config2 else cfg(1, n_amps=4)
        s = pq.Simulator(c, strategy="cached")
        for x, y in zip(xs[:4], ys[:4]):
            assert s.selfcheck(x, y, trials=4000) in (None, 0)


def test_pair_strategy_engines_are_host_consistent():
    rng = W.Rng(12)
    for trial in range(30):
        n = 2 + rng.below(3)
        c = W.random_layered(n, 3, rng)
        for p in (1, 3):
            s = pq.Simulator(c, pairs=p)
            st = s.stats()
            assert st[6] <= p
            assert s.selfcheck(0, rng.below(1 << n), trials=16) in (None, 0)


def test_invalid_inputs_raise_clean_errors():
    c = pq.Circuit(2).h(0).t(0).cx(0, 1)
    with pytest.raises(ValueError):
        pq.Simulator(c, strategy="bogus")
    bad = pq.Circuit(2)
    bad.ops.append(99); bad.q0.append(0); bad.q1.append(0); bad.ph.append(0)
    with pytest.raises(pq.PolyqError):
        pq.Simulator(bad)
    cz = pq.Circuit(2)
    cz.ops.append(pq.CZ); cz.q0.append(1); cz.q1.append(1); cz.ph.append(0)
    with pytest.raises(pq.PolyqError):
        pq.Simulator(cz)


def _have_gpu():
    try:
        pq.device_info()
        return True
    except pq.PolyqError:
        return False


@pytest.mark.skipif(_have_gpu(), reason="checks the no-device error path")
def test_device_calls_fail_loudly_without_a_gpu():
    c = pq.Circuit(2).h(0).t(0).cx(0, 1).h(1)
    s = pq.Simulator(c)
    with pytest.raises(pq.PolyqError, match="HIP|device"):
        s.amplitude(0, 1)
