"""Statevector and weak simulation (sampling) on the GPU against the oracle."""
import numpy as np
import pytest

import oracle_lib as O
import polyq_amd as pq
from polyq_amd import workloads as W

pytestmark = pytest.mark.gpu


def _sampling_circuit():
    # sim.rs:383-398 (sampling_follows_the_distribution)
    c = pq.Circuit(3)
    c.h(0).t(0).h(0).cx(0, 1).h(2).t(2).h(2).cz(1, 2).h(1).t(1).h(1)
    return c


def test_sampling_follows_the_distribution():
    # sim.rs:383-398 on the device: 40000 shots (seed 1234) within 0.012 of
    # |amplitude|^2, and the same shots as the oracle's sample_many
    c = _sampling_circuit()
    oc = O.OCircuit.from_circuit(c)
    sv = O.dense(oc, 0)
    shots = 40000
    got = pq.Simulator(c).sample_many(0, shots, 1234)
    hist = np.bincount(np.asarray(got, np.int64), minlength=8)
    for y in range(8):
        assert abs(abs(sv[y]) ** 2 - hist[y] / shots) < 0.012, y
    want = O.OSim(oc, "auto").sample_many(0, 2000, 1234)
    np.testing.assert_array_equal(np.asarray(got[:2000], np.uint64), want)


@pytest.mark.parametrize("form", ["z8", "plain_fp64"])
def test_sample_many_equals_oracle(form):
    # same per-shot streams, same chain rule: the same outcomes (a flip would
    # need a uniform draw within ~1e-13 of a prefix probability)
    rng = W.Rng(41)
    for trial in range(5):
        n = 4 + rng.below(5)
        c = W.random_clifford_t(n, 30, 2 + rng.below(6), trial % 2 == 1, rng)
        x = rng.below(1 << n)
        got = pq.Simulator(c, strategy="cached", form=form).sample_many(x, 300, 77 + trial)
        want = O.OSim(O.OCircuit.from_circuit(c), "cached").sample_many(x, 300, 77 + trial)
        np.testing.assert_array_equal(np.asarray(got, np.uint64), want)


@pytest.mark.parametrize("form", ["z8", "branch_walk", "plain_fp64"])
def test_statevector_matches_dense(form):
    # sim.rs:329-342 shape: every amplitude against the dense statevector
    rng = W.Rng(99)
    for trial in range(8):
        n = 2 + rng.below(9)
        c = W.random_clifford_t(n, 25, rng.below(8), trial % 4 == 0, rng) if trial % 2 == 0 else W.random_layered(n, 2 + rng.below(2), rng)
        x = rng.below(1 << n)
        want = O.dense(O.OCircuit.from_circuit(c), x)
        got = pq.Simulator(c, strategy="cached", form=form).statevector_array(x)
        np.testing.assert_allclose(got, want, atol=1e-6 if form == "branch_walk" else 1e-10)
        assert abs(np.sum(np.abs(got) ** 2) - 1) < 1e-6


def test_statevector_across_batches():
    # 2^17 outputs: two batches of 2^16 queries; normalisation and a dense spot check
    rng = W.Rng(5)
    c = W.random_layered(17, 2, rng)
    got = pq.Simulator(c, strategy="cached").statevector_array(0)
    assert abs(np.sum(np.abs(got) ** 2) - 1) < 1e-9
    want = O.dense(O.OCircuit.from_circuit(c), 0)
    np.testing.assert_allclose(got, want, atol=1e-10)


def _estimate_cases():
    # as tests/test_oracle.py: H-sandwiched Clifford+T layers, d up to ~12
    rng = W.Rng(23)
    out = []
    for trial in range(6):
        n = 9 + rng.below(4)
        c = W.sandwich_layer(n, 11 + rng.below(5), (15 * n) // 2, (5 * n) // 2, rng)
        out.append((c, 0, rng.below(1 << n)))
    return out


@pytest.mark.parametrize("form", ["z8", "plain_fp64"])
@pytest.mark.parametrize("strategy", ["cached", "pairs"])
def test_amplitude_estimate_equals_oracle(strategy, form):
    # the same random points (per-chunk SplitMix64 streams, jumped to on the
    # device) or the same random branch engines: the same estimate and error
    for c, x, y in _estimate_cases():
        oc = O.OCircuit.from_circuit(c)
        o = O.OSim(oc, strategy, 3)
        s = pq.Simulator(c, strategy=strategy, pairs=3 if strategy == "pairs" else None, form=form)
        for samples, seed in ((200, 1), (3000, 9)):
            re, im, se = s.amplitude_estimate(x, y, samples, seed)
            want, wse = o.amplitude_estimate(x, y, samples, seed)
            scale = abs(want) + wse + 1e-300
            assert abs(complex(re, im) - want) <= 1e-9 * scale, (samples, seed, complex(re, im), want)
            assert abs(se - wse) <= 1e-9 * scale


def test_amplitude_estimate_on_the_headline_workload():
    # n = 40, 36 T (d = 30): 2^20 uniform points; within 6 standard errors of
    # the exact amplitude
    c, xs, ys = W.config2(5)
    s = pq.Simulator(c, strategy="cached")
    exact = s.amplitudes(xs[:1], ys[:1])[0]
    re, im, se = s.amplitude_estimate(xs[0], ys[0], 1 << 20, 12345)
    print("estimate", complex(re, im), "+-", se, "exact", exact)
    assert se > 0
    assert abs(complex(re, im) - exact) < 6 * se


def test_statevector_bytes_as_the_reference_binding():
    # pybind.rs:158-168: raw little-endian complex128 bytes, re then im
    c, _, _ = W.config1(2)
    s = pq.Simulator(c, strategy="cached")
    b = s.statevector(5)
    assert isinstance(b, bytes) and len(b) == 16 << c.num_qubits
    a = s.statevector_array(5)
    np.testing.assert_array_equal(np.frombuffer(b, dtype="<f8").reshape(-1, 2), np.stack([a.real, a.imag], axis=1))
