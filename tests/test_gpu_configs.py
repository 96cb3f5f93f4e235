"""The benchmark workloads at their full sizes on the GPU, each checked
against the FP64 GPU form per amplitude and against the CPU oracle on a
sample (src/engine.rs:745-774 and src/sim.rs:329-342 patterns at scale).

  n=40, 36 T (seed 5, d = 30): 16 amplitudes, 4 of them against the oracle
  n=48, 40 Rz (seed 4, d = 30): 1024 amplitudes, 32 against the oracle
  256 circuits x 16 queries in one graph: replays == eager, all vs FP64,
      64 spread over d against the oracle
  Poisson serving (config-4 circuits, 70/20/10 mix, 64 GB): every result
      bitwise equal to its alone result, 64 against the oracle
  the two heaviest marginals of the 3,200 req/s serving schedule (d = 37,
      39; large segment tables) against the dense statevector
"""
import os
import random
import time

import numpy as np
import pytest

import oracle_lib as O
import polyq_amd as pq
from polyq_amd import workloads as W

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
PER_AMP_REL = 1e-9          # per amplitude, relative to its own magnitude
ORACLE_REL = 1e-9


def _rel(got, want):
    got, want = np.asarray(got), np.asarray(want)
    return np.abs(got - want) / np.maximum(np.abs(want), 1e-300)


def _oracle(c):
    return O.OSim(O.OCircuit.from_circuit(c), "cached", threads=THREADS)


def test_config2_seed5_all_forms_per_amplitude_and_oracle():
    c, xs, ys = W.config2(5)
    s = pq.Simulator(c, strategy="cached")
    assert s.stats()[5] == 30
    fp64 = pq.Simulator(c, strategy="cached", form="plain_fp64").amplitudes(xs, ys)
    for form in ("z8", "branch_walk"):
        got = s.amplitudes(xs, ys, form=form)
        rel = _rel(got, fp64)
        print("config2 %s vs FP64: max rel %.3e" % (form, rel.max()), flush=True)
        assert rel.max() < PER_AMP_REL
    o = _oracle(c)
    for i in (0, 5, 10, 15):
        want = o.amplitude(xs[i], ys[i])
        for form in ("z8", "branch_walk", "plain_fp64"):
            got = s.amplitudes(xs[i:i + 1], ys[i:i + 1], form=form)[0]
            assert abs(got - want) <= ORACLE_REL * abs(want), (i, form, got, want)


def test_config3_every_amplitude_against_fp64_and_32_against_oracle():
    c, xs, ys = W.config3(4, n_amps=1024)
    s = pq.Simulator(c, strategy="cached")
    assert s.stats()[5] == 30
    fp64 = pq.Simulator(c, strategy="cached", form="plain_fp64").amplitudes(xs, ys)
    assert np.all(np.abs(fp64) > 0)
    for form in ("z8", "branch_walk"):
        got = s.amplitudes(xs, ys, form=form)
        rel = _rel(got, fp64)
        small = np.abs(fp64) < 2.0 ** -20
        print("config3 %s vs FP64: max rel %.3e (%d below 2^-20, max rel there %.3e)"
              % (form, rel.max(), small.sum(), rel[small].max() if small.any() else 0), flush=True)
        assert rel.max() < PER_AMP_REL
    o = _oracle(c)
    z = s.amplitudes(xs, ys)
    for i in range(0, 1024, 32):
        want = o.amplitude(xs[i], ys[i])
        assert abs(fp64[i] - want) <= ORACLE_REL * abs(want), (i, fp64[i], want)
        assert abs(z[i] - want) <= ORACLE_REL * abs(want), (i, z[i], want)


def test_config4_graph_replays_fp64_and_oracle():
    work = W.config4(1)
    assert len(work) == 256
    sims = [pq.Simulator(c, strategy="cached") for c, _ in work]
    S, X, Y, D = [], [], [], []
    for s, (c, pairs) in zip(sims, work):
        d = s.stats()[5]
        for x, y in pairs:
            S.append(s); X.append(x); Y.append(y); D.append(d)
    assert len(S) == 4096
    eager = pq.batch_amplitudes(S, X, Y)
    g = pq.Graph(S, X, Y)
    for rep in range(3):
        got, ms = g.run()
        np.testing.assert_array_equal(got, eager)
        # eager work and a synchronisation between replays
        S[rep].amplitudes(X[rep:rep + 2], Y[rep:rep + 2])
    plain = {id(s): pq.Simulator(s.circuit, strategy="cached", form="plain_fp64") for s in sims}
    ref = np.array([plain[id(s)].amplitudes([x], [y])[0] for s, x, y in zip(S, X, Y)])
    scale = float(np.max(np.abs(ref)))
    err = np.abs(eager - ref)
    print("config4: max |z8 - FP64| / max|a| = %.3e" % (err.max() / scale), flush=True)
    assert err.max() <= 1e-11 * scale
    # 64 queries spread over the branch dimensions, against the oracle
    D = np.array(D)
    order = np.argsort(D, kind="stable")
    picks = [int(order[k]) for k in np.linspace(0, len(order) - 1, 64).astype(int)]
    cache = {}
    for i in picks:
        s = S[i]
        o = cache.get(id(s))
        if o is None:
            o = cache[id(s)] = _oracle(s.circuit)
        want = o.amplitude(X[i], Y[i])
        assert abs(eager[i] - want) <= 1e-10 * max(abs(want), 1e-3 * scale), (i, D[i], eager[i], want)
    assert len({int(D[i]) for i in picks}) >= 16


def test_config5_serving_results_equal_alone_and_oracle():
    rnd = random.Random(2025)
    rng = W.Rng(78)
    rate, n_req = 200.0, 2000
    t, sched, circuits = 0.0, [], []
    while len(sched) < n_req:
        t += rnd.expovariate(rate)
        if not circuits or rnd.random() < 0.5:
            c = W.config4_circuit(rng)
            circuits.append(c)
        else:
            c = circuits[rnd.randrange(len(circuits))]
        u = rnd.random()
        n = c.num_qubits
        x, y = rnd.getrandbits(n), rnd.getrandbits(n)
        if u < 0.7:
            sched.append((t, c, "amplitude", x, y, ()))
        elif u < 0.9:
            sched.append((t, c, "probability", x, y, ()))
        else:
            sched.append((t, c, "marginal", x, 0, tuple((q, rnd.random() < 0.5) for q in rnd.sample(range(n), 4))))
    srv = pq.Server(hbm_budget_bytes=64 << 30, workers=8)
    tickets = []
    t0 = time.perf_counter()
    for (ta, c, kind, x, y, meas) in sched:
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        tickets.append(srv.submit(c, kind, x, y, meas))
    vals, lat = [], []
    for tk in tickets:
        v, l = srv.wait(tk, timeout=120.0)
        vals.append(v); lat.append(l)
    st = srv.stats()
    srv.close()
    print("config5: %d requests, p50 %.3f ms, p99 %.3f ms, max batch %d, overlapped %d"
          % (len(vals), np.percentile(lat, 50), np.percentile(lat, 99), st["max_batch"], st["overlapped_batches"]), flush=True)
    assert st["completed"] == n_req and st["failed"] == 0
    kinds = {k: sum(1 for r in sched if r[2] == k) for k in ("amplitude", "probability", "marginal")}
    assert kinds["marginal"] > 100 and kinds["probability"] > 250
    # every result bitwise equal to the same request evaluated alone
    sims = {}
    for i, (_, c, kind, x, y, meas) in enumerate(sched):
        s = sims.get(id(c))
        if s is None:
            s = sims[id(c)] = pq.Simulator(c, strategy="cached")
        if kind == "amplitude":
            want = complex(*s.amplitude(x, y))
        elif kind == "probability":
            want = complex(s.probability(x, y), 0.0)
        else:
            want = complex(s.marginal(x, meas), 0.0)
        assert vals[i] == want, (i, kind, vals[i], want)
    # 64 of them against the CPU oracle (branch dimension <= 24, both kinds of engine)
    checked, oc = 0, {}
    for i, (_, c, kind, x, y, meas) in enumerate(sched):
        s = sims[id(c)]
        if kind == "marginal":
            if s.marginal_stats([q for q, _ in meas])[5] > 24:
                continue
        elif s.stats()[5] > 24:
            continue
        o = oc.get(id(c))
        if o is None:
            o = oc[id(c)] = _oracle(c)
        if kind == "amplitude":
            want = o.amplitude(x, y)
            assert abs(vals[i] - want) <= 1e-10 * max(abs(want), 1e-8), (i, vals[i], want)
        elif kind == "probability":
            want = abs(o.amplitude(x, y)) ** 2
            assert abs(vals[i].real - want) <= 1e-9 * max(want, 1e-12), (i, vals[i], want)
        else:
            want = o.marginal(x, list(meas))
            assert abs(vals[i].real - want) <= 1e-9 * max(want, 1e-9), (i, vals[i], want)
        checked += 1
        if checked == 64:
            break
    assert checked == 64


def _serve_circuits(rate, secs, rseed, cseed):
    """The request stream of the serving benchmark's Poisson schedule (the
    same draws in the same order): (kind, circuit, x, measured) per request."""
    rnd = random.Random(rseed)
    rng = W.Rng(cseed)
    t, out, circuits = 0.0, [], []
    while True:
        t += rnd.expovariate(rate)
        if t >= secs:
            return out
        if not circuits or rnd.random() < 0.5:
            c = W.config4_circuit(rng)
            circuits.append(c)
        else:
            c = circuits[rnd.randrange(len(circuits))]
        u = rnd.random()
        n = c.num_qubits
        x, y = rnd.getrandbits(n), rnd.getrandbits(n)
        if u < 0.9:
            out.append(("amplitude" if u < 0.7 else "probability", c, x, ()))
        else:
            out.append(("marginal", c, x, tuple((q, rnd.random() < 0.5) for q in rnd.sample(range(n), 4))))


def test_pivot_heavy_marginal_against_dense():
    # The heaviest 4-qubit marginals of the 3,200 req/s serving schedule:
    # sandwich engines of d = 37 and 39 whose inner directions touch many
    # pivots (they get segment tables of up to 2^20 entries), on 23 and 21
    # qubits, against the marginal of the dense statevector
    reqs = _serve_circuits(3200.0, 5.0, 3200, 3201)
    heavy = []
    for kind, c, x, meas in reqs:
        if kind == "marginal" and c.num_qubits <= 23:
            s = pq.Simulator(c, strategy="cached")
            d = s.marginal_stats([q for q, _ in meas])[5]
            if d >= 37:
                heavy.append((s, c, x, meas, d))
    assert sorted(h[4] for h in heavy) == [37, 39]
    for s, c, x, meas, d in heavy:
        got = s.marginal(x, list(meas))
        sv = O.dense(O.OCircuit.from_circuit(c), x)
        idx = np.arange(1 << c.num_qubits)
        keep = np.ones(idx.shape, bool)
        for q, b in meas:
            keep &= ((idx >> q) & 1) == (1 if b else 0)
        want = float(np.sum(np.abs(sv[keep]) ** 2))
        print("d=%d n=%d marginal %.12g dense %.12g" % (d, c.num_qubits, got, want), flush=True)
        assert abs(got - want) <= 1e-10, (d, got, want)
