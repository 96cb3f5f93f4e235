"""HIP path vs the CPU oracle through the C ABI (MI355X)."""
import math

import numpy as np
import pytest

import oracle_lib as O
import polyq_amd as pq
from polyq_amd import workloads as W

pytestmark = pytest.mark.gpu

FP32_REL = 1e-6      # FP32-compensated Z8 forms: relative to max(|a|, scale of the amplitude set)
FP64_ABS = 1e-10     # plain FP64 form


def _close(got, want, rel, floor):
    return abs(got - want) <= rel * max(abs(want), floor)


def test_device_present():
    info = pq.device_info()
    print("device:", info)
    assert "gfx950" in info


@pytest.mark.parametrize("form", ["z8", "z8_accurate", "plain_fp64"])
def test_engine_matches_brute_force_all_boundaries(form):
    # engine.rs:745-774 pattern on the GPU: every (x, y) of small circuits
    rng = W.Rng(21)
    checked = 0
    for trial in range(60):
        n = 1 + rng.below(4)
        c = W.random_clifford_t(n, 8 + rng.below(14), rng.below(6), trial % 3 == 0, rng)
        oc = O.OCircuit.from_circuit(c)
        if O.pathsum_info(oc)["live"] > 20:
            continue
        s = pq.Simulator(c, strategy="cached", form=form)
        xs = [x for x in range(1 << n) for _ in range(1 << n)]
        ys = [y for _ in range(1 << n) for y in range(1 << n)]
        got = s.amplitudes(xs, ys)
        for a, x, y in zip(got, xs, ys):
            want = O.brute(oc, x, y)
            tol = FP64_ABS if form == "plain_fp64" else 1e-6
            assert abs(a - want) < tol, (trial, x, y, a, want)
        checked += 1
    assert checked >= 30


def test_config1_against_dense():
    c, xs, ys = W.config1(1)
    oc = O.OCircuit.from_circuit(c)
    want = np.array([O.dense(oc, x)[y] for x, y in zip(xs, ys)])
    for form in ("z8", "plain_fp64"):
        s = pq.Simulator(c, strategy="cached", form=form)
        got = s.amplitudes(xs, ys)
        tol = 1e-10 if form == "plain_fp64" else 1e-7
        np.testing.assert_allclose(got, want, atol=tol)


def test_branch_outputs_bit_exact_against_oracle():
    # per-branch (zero, r, k8) and total phase, keyed by the t-bit branch vector s
    rng = W.Rng(404)
    checked = 0
    for trial in range(40):
        n = 2 + rng.below(4)
        c = W.random_clifford_t(n, 20 + rng.below(30), 2 + rng.below(12), trial % 2 == 0, rng)
        oc = O.OCircuit.from_circuit(c)
        s = pq.Simulator(c, strategy="cached")
        o = O.OSim(oc, "cached")
        t = s.stats()[4]
        if t == 0:
            continue
        x, y = rng.below(1 << n), rng.below(1 << n)
        S = np.array([rng.next_u64() & ((1 << t) - 1) for _ in range(512)], np.uint64)
        surv = o.surviving(x, y, cap=512)
        S = np.concatenate([S, surv]).astype(np.uint64)
        d = s.branch_dump(x, y, S)
        od = o.branch_dump(x, y, S)
        if d is None:
            assert od is None
            continue
        zero, k8, kt, r = d
        ozero, orr, ok8, okt = od
        np.testing.assert_array_equal(zero, ozero)
        live = ozero == 0
        np.testing.assert_array_equal(k8[live], ok8[live])
        np.testing.assert_array_equal(kt[live], okt[live])
        assert np.all(orr == r)
        checked += 1
    assert checked >= 20


def test_walk_phases_bit_exact_against_oracle():
    # every branch the Z8 walk visits: integer phase k(c) equals the oracle's
    # phase8(u(s)) + 6|s| + 4|neg \ s| (mod 8) for the same s, bit for bit
    rng = W.Rng(505)
    checked = 0
    for trial in range(30):
        n = 3 + rng.below(4)
        c = W.random_clifford_t(n, 30 + rng.below(40), 6 + rng.below(10), trial % 2 == 0, rng)
        s = pq.Simulator(c, strategy="cached")
        info = s.z8_info()
        if info["slots"] > 16:
            continue
        o = O.OSim(O.OCircuit.from_circuit(c), "cached")
        x, y = rng.below(1 << n), rng.below(1 << n)
        for mid in (0, 2):
            dump = s.walk_dump(x, y, mid_bits=mid)
            if dump is None:
                continue
            k, mag, S = dump
            real = k != 255
            assert real.sum() == 1 << info["d"]
            assert len(set(S[real].tolist())) == 1 << info["d"]          # each surviving branch exactly once
            zero, r, k8, ktot = o.branch_dump(x, y, S[real])
            assert not zero.any()
            np.testing.assert_array_equal(k[real], ktot)
            checked += 1
    assert checked >= 10


def test_strategies_agree_on_gpu():
    rng = W.Rng(5)
    for trial in range(20):
        n = 2 + rng.below(3)
        c = W.random_clifford_t(n, 20, 4 + rng.below(5), trial % 4 == 0, rng) if trial % 2 == 0 else W.random_layered(n, 3, rng)
        oc = O.OCircuit.from_circuit(c)
        want = O.dense(oc, 0)
        for kw in ({"strategy": "cached"}, {"strategy": "auto"}, {"pairs": 2}, {"pairs": 6}):
            s = pq.Simulator(c, **kw)
            got = s.amplitudes([0] * (1 << n), list(range(1 << n)))
            np.testing.assert_allclose(got, want, atol=1e-6)


def test_marginals_match_dense_on_gpu():
    rng = W.Rng(3)
    for _ in range(8):
        n = 2 + rng.below(3)
        c = W.random_layered(n, 2, rng)
        sv = O.dense(O.OCircuit.from_circuit(c), 0)
        s = pq.Simulator(c)
        for _ in range(3):
            k = 1 + rng.below(n)
            meas = [(q, rng.next_f64() < 0.5) for q in range(k)]
            want = sum(abs(a) ** 2 for i, a in enumerate(sv) if all(((i >> q) & 1) == b for q, b in meas))
            for form in ("z8", "plain_fp64"):
                assert abs(s.marginal(0, meas, form=form) - want) < 1e-6


def test_config2_small_seed_against_oracle():
    # the config-2 generator at seed 1 (d = 19): all 16 amplitudes vs the oracle
    c, xs, ys = W.config2(1)
    o = O.OSim(O.OCircuit.from_circuit(c), "cached")
    o.set_threads(8)
    want = np.array([o.amplitude(x, y) for x, y in zip(xs, ys)])
    scale = np.max(np.abs(want))
    for form in ("z8", "z8_accurate", "plain_fp64"):
        got = pq.Simulator(c, strategy="cached", form=form).amplitudes(xs, ys)
        err = np.max(np.abs(got - want)) / scale
        print(form, "max err / max|a| =", err)
        assert err < (1e-12 if form == "plain_fp64" else FP32_REL)


def test_normalisation_sum_of_probabilities():
    # size-independent property: sum_y |<y|C|0>|^2 = 1
    rng = W.Rng(9)
    c = W.sandwich_layer(10, 12, 40, 10, rng)
    s = pq.Simulator(c, strategy="cached")
    a = s.amplitudes([0] * 1024, list(range(1024)))
    assert abs(np.sum(np.abs(a) ** 2) - 1) < 1e-6
