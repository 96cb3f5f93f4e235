"""Batches of at most 16 queries travel whole as kernel arguments (one
workgroup per (item, query), each query with its own engine descriptor);
larger ones are staged.  Same kernels, so a query's result must be the same
bit for bit either way."""
import numpy as np
import pytest

import polyq_amd as pq
from polyq_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_tuning():
    tuning = pq.get_walk_tuning()
    yield
    pq.set_walk_tuning(*tuning)


def _staged(sims, X, Y):
    """The same queries inside a batch of 17 or more (the staged path)."""
    n = len(sims)
    pad = max(0, 17 - n)
    out = pq.batch_amplitudes(list(sims) + [sims[0]] * pad, list(X) + [X[0]] * pad, list(Y) + [Y[0]] * pad)
    return out[:n]


@pytest.mark.parametrize("tuning", [(12, 8), (11, 4), (14, 2), (10, 6), (13, 1), (16, 0)])
def test_argument_batches_equal_staged_batches(tuning, restore_tuning):
    # lane walks of 2^M points for M = 0..4 (short-walk instance) and 6 (long)
    pq.set_walk_tuning(*tuning)
    c, xs, ys = W.config2(5)
    sim = pq.Simulator(c, strategy="cached")
    for n in (1, 3, 16):
        X = [int(x) for x in xs[:n]]
        Y = [int(y) for y in ys[:n]]
        np.testing.assert_array_equal(pq.batch_amplitudes([sim] * n, X, Y), _staged([sim] * n, X, Y))


def test_argument_batch_of_many_engines_and_both_walk_variants(restore_tuning):
    # engines of different shapes in one kernel-argument batch, whose queries
    # take the short-walk and the long-walk kernel (two launches)
    pq.set_walk_tuning(8, 8)
    c1, xs1, ys1 = W.config2(5)      # d = 30: 14 outer slots, M = 6
    c2, xs2, ys2 = W.config2(2)      # d = 24: 10 outer slots, M = 2
    rng = W.Rng(7)
    c3 = W.sandwich_layer(16, 12, 120, 40, rng)    # no outer slot at all
    c4 = W.sandwich_layer(16, 12, 120, 40, rng)
    sims = [pq.Simulator(c, strategy="cached") for c in (c1, c2, c3, c4)]
    assert sims[0].seg_info()["mid_bits"] > 4 and sims[1].seg_info()["mid_bits"] <= 4
    S = [sims[0]] * 4 + [sims[1]] * 4 + [sims[2]] * 4 + [sims[3]] * 4
    X = [int(x) for x in xs1[:4]] + [int(x) for x in xs2[:4]] + [rng.below(1 << 16) for _ in range(8)]
    Y = [int(y) for y in ys1[:4]] + [int(y) for y in ys2[:4]] + [rng.below(1 << 16) for _ in range(8)]
    a = pq.batch_amplitudes(S, X, Y)
    np.testing.assert_array_equal(a, _staged(S, X, Y))
    alone = np.array([S[i].amplitudes([X[i]], [Y[i]])[0] for i in range(16)])
    np.testing.assert_array_equal(a, alone)


def test_buffer_protocol_call_equals_ctypes_call():
    # Simulator.amplitudes passes uint64 arrays through the CPython fast path
    # and anything else through ctypes: the same C ABI call either way
    assert pq._pqfast is not None, "the CPython fast path was not built"
    c, xs, ys = W.config2(5)
    sim = pq.Simulator(c, strategy="cached")
    X = np.asarray(xs, np.uint64)
    Y = np.asarray(ys, np.uint64)
    a = sim.amplitudes(X, Y)
    b = sim.amplitudes([int(x) for x in xs], [int(y) for y in ys])
    np.testing.assert_array_equal(a, b)
    with pytest.raises(ValueError):
        sim.amplitudes(X, Y[:3])


def test_timed_argument_batches_equal_untimed(restore_tuning):
    # timing events bound to the walk dispatches (one launch, then two: the
    # short-walk and the long-walk kernel) change no result and give a
    # kernel time inside the batch's total
    c1, xs1, ys1 = W.config2(5)
    c2, xs2, ys2 = W.config2(2)
    s1, s2 = pq.Simulator(c1, strategy="cached"), pq.Simulator(c2, strategy="cached")
    for tuning, S, X, Y in (((12, 8), [s1] * 16, xs1[:16], ys1[:16]),
                            ((8, 8), [s1] * 3 + [s2] * 3, list(xs1[:3]) + list(xs2[:3]), list(ys1[:3]) + list(ys2[:3]))):
        pq.set_walk_tuning(*tuning)
        X = [int(x) for x in X]
        Y = [int(y) for y in Y]
        plain = pq.batch_amplitudes(S, X, Y)
        with pq.timed():
            a = pq.batch_amplitudes(S, X, Y)
            t = pq.last_timing()
        np.testing.assert_array_equal(a, plain)
        assert np.isfinite(t["kernel_ms"]) and t["kernel_ms"] > 0
        assert t["kernel_ms"] <= t["total_ms"] + 1e-6
    pq.batch_amplitudes(S, X, Y)
    assert np.isnan(pq.last_timing()["kernel_ms"])      # untimed again outside the block
