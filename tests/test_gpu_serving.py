"""Batched, graph-replayed and served evaluation on the GPU: every result is
the same alone and in the mix."""
import random
import time

import numpy as np
import pytest

import oracle_lib as O
import polyq_amd as pq
from polyq_amd import workloads as W

pytestmark = pytest.mark.gpu


def _mixed_circuits(seed, count, nmax=24, tmax=16):
    rng = W.Rng(seed)
    out = []
    for _ in range(count):
        n = 6 + rng.below(nmax - 5)
        t = 2 + rng.below(tmax)
        c = W.sandwich_layer(n, t, (15 * n) // 2, (5 * n) // 2, rng)
        pairs = [(rng.below(1 << n), rng.below(1 << n)) for _ in range(4)]
        out.append((c, pairs))
    return out


def test_batch_results_equal_alone_results_bitwise():
    work = _mixed_circuits(11, 24)
    sims = [pq.Simulator(c, strategy="cached") for c, _ in work]
    S, X, Y, alone = [], [], [], []
    for s, (c, pairs) in zip(sims, work):
        for x, y in pairs:
            S.append(s); X.append(x); Y.append(y)
            alone.append(s.amplitudes([x], [y])[0])
    mixed = pq.batch_amplitudes(S, X, Y)
    np.testing.assert_array_equal(mixed, np.array(alone))
    # and against the FP64 oracle on the small ones
    for s, x, y, a in list(zip(S, X, Y, mixed))[:24]:
        o = O.OSim(O.OCircuit.from_circuit(s.circuit), "cached")
        if s.stats()[5] <= 18:
            assert abs(o.amplitude(x, y) - a) < 1e-10


@pytest.mark.parametrize("form", ["z8", "branch_walk"])
def test_graph_replays_match_eager(form):
    work = _mixed_circuits(12, 12)
    sims = [pq.Simulator(c, strategy="cached") for c, _ in work]
    S, X, Y = [], [], []
    for s, (c, pairs) in zip(sims, work):
        for x, y in pairs:
            S.append(s); X.append(x); Y.append(y)
    eager = pq.batch_amplitudes(S, X, Y, form=form)
    g = pq.Graph(S, X, Y, form=form)
    for rep in range(4):
        got, ms = g.run()
        np.testing.assert_array_equal(got, eager)
        # eager work and a synchronisation between replays
        S[4 * rep].amplitudes(X[4 * rep:4 * rep + 2], Y[4 * rep:4 * rep + 2])


def test_server_results_equal_alone_results():
    work = _mixed_circuits(13, 10, nmax=16)
    srv = pq.Server(hbm_budget_bytes=64 << 30, workers=4)
    rnd = random.Random(5)
    tickets = []
    for rep in range(3):
        for c, pairs in work:
            kind = rnd.choice(["amplitude", "amplitude", "probability", "marginal"])
            x, y = pairs[rnd.randrange(4)]
            meas = [(q, rnd.random() < 0.5) for q in rnd.sample(range(c.num_qubits), 2)] if kind == "marginal" else ()
            tickets.append((srv.submit(c, kind, x, y, meas), c, kind, x, y, meas))
    for t, c, kind, x, y, meas in tickets:
        val, lat = srv.wait(t, timeout=120.0)
        s = pq.Simulator(c, strategy="cached")
        if kind == "amplitude":
            want = complex(*s.amplitude(x, y))
            assert val == want
        elif kind == "probability":
            assert val.real == s.probability(x, y)
        else:
            assert val.real == s.marginal(x, meas)
        assert lat > 0
    st = srv.stats()
    assert st["completed"] == len(tickets) and st["failed"] == 0
    assert st["builds"] == len(work)          # repeats hit resident engines
    srv.close()


def test_server_tiny_budget_evicts_and_stays_correct():
    work = _mixed_circuits(14, 8, nmax=14)
    probe = pq.Server(hbm_budget_bytes=64 << 30, workers=1)
    c0, p0 = work[0]
    probe.wait(probe.submit(c0, "amplitude", *p0[0]), timeout=120.0)
    one = probe.stats()["resident_bytes"]
    probe.close()
    assert one > 0
    # room for about two engines: later circuits must evict earlier ones
    srv = pq.Server(hbm_budget_bytes=int(2.5 * one), workers=2)
    tickets = []
    for rep in range(2):
        for c, pairs in work:
            x, y = pairs[0]
            tickets.append((srv.submit(c, "amplitude", x, y), c, x, y))
    got = {}
    for t, c, x, y in tickets:
        try:
            got[t] = srv.wait(t, timeout=120.0)[0]
        except pq.PolyqError as e:
            assert "HBM budget" in str(e)      # an engine larger than the whole budget is refused, not crashed
    st = srv.stats()
    assert st["evictions"] > 0
    assert st["resident_bytes"] <= 2.5 * one
    assert len(got) >= len(tickets) // 2
    for t, c, x, y in tickets:
        if t in got:
            assert got[t] == complex(*pq.Simulator(c, strategy="cached").amplitude(x, y))
    srv.close()


def test_large_batch_reduction_equals_fused_small_batches_bitwise():
    # A batch above the fused-reduction size (here 320 queries x 16 items of
    # the n=40 workload circuit) takes the separate reduction launch; small
    # batches finish queries inside the walk.  Same reduction order: same bits.
    c, _, _ = W.config2(5)
    s = pq.Simulator(c, strategy="cached")
    rng = W.Rng(123)
    xs = [rng.below(1 << 40) for _ in range(320)]
    ys = [0] * 320
    big = s.amplitudes(xs, ys)
    assert pq.last_timing()["items"] > 4096
    small = np.concatenate([s.amplitudes(xs[i:i + 8], ys[i:i + 8]) for i in range(0, 320, 8)])
    np.testing.assert_array_equal(big, small)


def test_branch_walk_host_finish_equals_device_reduction_bitwise():
    # The per-branch walk of the n=40 workload circuit at a lane walk of 3
    # mid slots: 512 items per query.  16 queries (8192 items) reduce on the
    # device; one query at a time (512 items) is finished on the host.  Same
    # reduction order: same bits.
    c, xs, ys = W.config2(5)
    s = pq.Simulator(c, strategy="cached", form="branch_walk")
    prev = pq.get_walk_tuning("branch_walk")
    try:
        pq.set_walk_tuning(27, 6, "branch_walk")
        big = s.amplitudes(xs, ys)
        assert pq.last_timing()["items"] == 16 * 512
        small = [s.amplitudes(xs[i:i + 1], ys[i:i + 1]) for i in range(16)]
        assert pq.last_timing()["items"] == 512
        np.testing.assert_array_equal(big, np.concatenate(small))
    finally:
        pq.set_walk_tuning(*prev, "branch_walk")


def test_host_finish_equals_device_reduction_many_items_per_query():
    # 1024 items per query (2^18 lanes of one point each): more items than
    # reduction threads, finished on the host in a batch of one query and on
    # the device in a batch of five.  Same order: same bits.
    c, _, _ = W.config2(26)
    bits = pq.get_segment_bits()
    pq.set_segment_bits(14)
    try:
        s = pq.Simulator(c, strategy="cached")
    finally:
        pq.set_segment_bits(bits)
    assert s.seg_info()["outer"] == 18
    prev = pq.get_walk_tuning()
    try:
        pq.set_walk_tuning(30, 0)
        rng = W.Rng(321)
        xs = [rng.below(1 << 40) for _ in range(5)]
        ys = [0] * 5
        big = s.amplitudes(xs, ys)
        assert pq.last_timing()["items"] == 5 * 1024
        small = np.concatenate([s.amplitudes(xs[i:i + 1], ys[i:i + 1]) for i in range(5)])
        np.testing.assert_array_equal(big, small)
    finally:
        pq.set_walk_tuning(*prev)


def _resident_bytes_of(c, x, y):
    probe = pq.Server(hbm_budget_bytes=64 << 30, workers=1)
    probe.wait(probe.submit(c, "amplitude", x, y), timeout=120.0)
    one = probe.stats()["resident_bytes"]
    probe.close()
    assert one > 0
    return one


def test_server_marginal_on_resident_circuit_beyond_budget_fails_fast():
    # The amplitude engine is resident; the marginal (sandwich) engine of the
    # same circuit would need about as much again, which the budget cannot
    # hold next to it while the request pins its own circuit: the request
    # fails with the HBM error at once instead of waiting for room.
    (c, pairs), = _mixed_circuits(15, 1, nmax=14)
    one = _resident_bytes_of(c, *pairs[0])
    srv = pq.Server(hbm_budget_bytes=int(1.5 * one), workers=1)
    a, _ = srv.wait(srv.submit(c, "amplitude", *pairs[0]), timeout=60.0)
    assert a == complex(*pq.Simulator(c, strategy="cached").amplitude(*pairs[0]))
    t0 = time.time()
    with pytest.raises(pq.PolyqError, match="HBM budget"):
        srv.wait(srv.submit(c, "marginal", pairs[0][0], 0, [(0, True), (1, False)]), timeout=60.0)
    assert time.time() - t0 < 30.0
    # the server keeps serving afterwards
    a, _ = srv.wait(srv.submit(c, "amplitude", *pairs[1]), timeout=60.0)
    assert a == complex(*pq.Simulator(c, strategy="cached").amplitude(*pairs[1]))
    srv.close()


def test_server_concurrent_marginals_under_tight_budget_never_hang():
    # Two resident circuits and a budget with room for one more engine: two
    # marginals submitted together each pin their own circuit and want the
    # other evicted.  Neither waits forever: each either runs or fails with
    # the HBM error, and the ones that run are right.
    work = _mixed_circuits(16, 2, nmax=14)
    one = max(_resident_bytes_of(c, *p[0]) for c, p in work)
    srv = pq.Server(hbm_budget_bytes=int(2.6 * one), workers=2)
    for c, pairs in work:
        srv.wait(srv.submit(c, "amplitude", *pairs[0]), timeout=60.0)
    meas = [(0, True), (1, True)]
    tickets = [(srv.submit(c, "marginal", pairs[0][0], 0, meas), c, pairs[0][0]) for c, pairs in work for _ in range(3)]
    ok = 0
    for t, c, x in tickets:
        try:
            v, _ = srv.wait(t, timeout=60.0)
            assert v.real == pq.Simulator(c, strategy="cached").marginal(x, meas)
            ok += 1
        except pq.PolyqError as e:
            assert "HBM budget" in str(e)
    st = srv.stats()
    assert st["completed"] + st["failed"] == len(tickets) + len(work)
    assert st["resident_bytes"] <= 2.6 * one
    srv.close()


def test_server_host_lru_drops_idle_builds_and_rebuilds():
    # A host budget below one build: every idle circuit is dropped when the
    # next one is built, comes back as a fresh build, and answers the same.
    work = _mixed_circuits(17, 6, nmax=14)
    srv = pq.Server(hbm_budget_bytes=64 << 30, workers=2, host_budget_bytes=1)
    got = []
    for rep in range(2):
        for c, pairs in work:
            for kind in ("amplitude", "marginal"):
                x, y = pairs[rep]
                meas = [(0, rep == 0), (1, True)] if kind == "marginal" else ()
                v, _ = srv.wait(srv.submit(c, kind, x, y, meas), timeout=60.0)
                got.append((c, kind, x, y, meas, v))
    hs = srv.host_stats()
    st = srv.stats()
    assert hs["host_drops"] >= len(work)
    assert st["builds"] >= 2 * len(work) - 1
    assert st["resident_circuits"] <= 2
    for c, kind, x, y, meas, v in got:
        s = pq.Simulator(c, strategy="cached")
        want = complex(*s.amplitude(x, y)) if kind == "amplitude" else complex(s.marginal(x, meas), 0.0)
        assert v == want
    srv.close()


def test_server_phase_timings_add_up_to_latency():
    # the six phases of a request run from its submission to its result, so
    # their means add up to the mean latency the waits report; the slowest 1%
    # add up to at least that mean
    work = _mixed_circuits(19, 12, nmax=16)
    srv = pq.Server(hbm_budget_bytes=64 << 30, workers=4)
    rnd = random.Random(8)
    tickets = []
    for rep in range(4):
        for c, pairs in work:
            kind = rnd.choice(["amplitude", "probability", "marginal"])
            x, y = pairs[rnd.randrange(4)]
            meas = [(q, rnd.random() < 0.5) for q in rnd.sample(range(c.num_qubits), 2)] if kind == "marginal" else ()
            tickets.append(srv.submit(c, kind, x, y, meas))
    lat = [srv.wait(t, timeout=120.0)[1] for t in tickets]
    ph = srv.phase_stats()
    srv.close()
    assert ph["requests"] == len(tickets)
    assert set(ph["all"]) == set(pq.Server.PHASES)
    assert all(v >= 0 for v in ph["all"].values()) and all(v >= 0 for v in ph["slowest_1pct"].values())
    total = sum(ph["all"].values())
    assert abs(total - float(np.mean(lat))) <= 1e-3 * float(np.mean(lat)) + 1e-4, (total, np.mean(lat))
    assert sum(ph["slowest_1pct"].values()) >= total
    assert ph["all"]["build"] > 0          # the first request of each circuit builds it
