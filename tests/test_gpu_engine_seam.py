"""The engine-level seam on the GPU: engines handed over by the oracle's
Engine fields (what a Rust Engine holds) evaluate to the oracle's Engine::eval
and Simulator::amplitude, per branch bit-exactly, and a RankEngine's branch
engines summed through one batched launch give RankEngine::eval_assignment
(src/engine.rs:551-561, src/rank.rs:184-190)."""
import numpy as np
import pytest

import oracle_lib as O
import polyq_amd as pq
from polyq_amd import workloads as W

pytestmark = pytest.mark.gpu


def _osim(c, strategy="cached", pairs=0):
    return O.OSim(O.OCircuit.from_circuit(c), strategy, pairs=pairs)


@pytest.mark.parametrize("form", ["z8", "branch_walk", "plain_fp64"])
def test_config1_engine_from_oracle_fields(form):
    c, xs, ys = W.config1(1)
    o = _osim(c)
    e = pq.Engine(O.engine_fields(o))
    oc = O.OCircuit.from_circuit(c)
    for x, y in zip(xs, ys):
        v = O.engine_values(o, x, y)
        assert v is not None
        got = e.eval(v, form=form)
        assert abs(got - O.engine_eval(o, 0, v)) < 1e-10
        assert abs(got - O.dense(oc, x)[y]) < 1e-10


def test_config2_seed1_batch_and_per_branch_bits():
    c, xs, ys = W.config2(1)
    o = _osim(c)
    e = pq.Engine(O.engine_fields(o))
    vals = [O.engine_values(o, x, y) for x, y in zip(xs, ys)]
    want = np.array([O.engine_eval(o, 0, v) for v in vals])
    scale = float(np.max(np.abs(want)))
    for form in ("z8", "branch_walk"):
        got = pq.engine_eval_batch([e] * len(vals), vals, form=form)
        assert np.max(np.abs(got - want)) <= 1e-11 * scale, form
        # the same amplitudes as the library's own build from the circuit
        sim = pq.Simulator(c, strategy="cached").amplitudes(xs, ys, form=form)
        assert np.max(np.abs(got - sim)) <= 1e-11 * scale
    # per-branch (zero, k8, total phase), keyed by the t-bit branch vector s
    rng = W.Rng(8)
    t = e.stats()[4]
    S = np.array([rng.next_u64() & ((1 << t) - 1) for _ in range(256)], np.uint64)
    S = np.concatenate([S, o.surviving(xs[0], ys[0], cap=256)]).astype(np.uint64)
    zero, k8, kt, r = e.branch_dump(vals[0], S)
    ozero, orr, ok8, okt = o.branch_dump(xs[0], ys[0], S)
    np.testing.assert_array_equal(zero, ozero)
    live = ozero == 0
    assert live.sum() >= 256
    np.testing.assert_array_equal(k8[live], ok8[live])
    np.testing.assert_array_equal(kt[live], okt[live])
    assert np.all(orr == r)


def test_rank_engine_of_pair_branches():
    # Strategy::Pairs(3): 8 branch engines built by the oracle, handed over
    # one by one, evaluated in one launch; their sum is the amplitude
    c, xs, ys = W.config2(1)
    o = _osim(c, "pairs", pairs=3)
    assert o.stats()[6] == 3
    engines = [pq.Engine(O.engine_fields(o, b)) for b in range(8)]
    for x, y in list(zip(xs, ys))[:6]:
        v = O.engine_values(o, x, y)
        vals = pq.engine_eval_batch(engines, [v] * 8)
        for b in range(8):
            assert abs(vals[b] - O.engine_eval(o, b, v)) <= 1e-12 + 1e-10 * abs(vals[b])
        want = o.amplitude(x, y)
        assert abs(vals.sum() - want) <= 1e-10 * max(abs(want), 1e-12)


def test_small_engines_against_brute_force_all_boundaries():
    rng = W.Rng(61)
    checked = 0
    for trial in range(40):
        n = 1 + rng.below(4)
        c = W.random_clifford_t(n, 8 + rng.below(14), rng.below(6), trial % 3 == 0, rng)
        oc = O.OCircuit.from_circuit(c)
        if O.pathsum_info(oc)["live"] > 20:
            continue
        o = _osim(c)
        e = pq.Engine(O.engine_fields(o))
        for x in range(1 << n):
            for y in range(1 << n):
                v = O.engine_values(o, x, y)
                want = O.brute(oc, x, y)
                got = 0j if v is None else e.eval(v)
                assert abs(got - want) < 1e-10, (trial, x, y, got, want)
        checked += 1
    assert checked >= 20


def _kernel_rows(f):
    t, d = f["terms"], f["branch_dim"]
    sw = (t + 63) // 64
    uw = 2 * ((f["planes"] + 63) // 64) + (f["lines"] + 63) // 64 + (f["radical"] + 63) // 64
    return f["kernel"][:d * sw].reshape(d, sw), f["kernel_u"][:d * uw].reshape(d, uw)


@pytest.mark.parametrize("config", [1, 2])
def test_sum_subspace_from_prepare_start(config):
    # pq_engine_sum_subspace on the start prepare leaves (sc.u, sc.s0,
    # c_start, engine.rs:596-606) against the oracle's Engine::sum_subspace
    # in every form, and from another point of the same coset (the sum does
    # not depend on where the walk starts)
    c, xs, ys = W.config1(1) if config == 1 else W.config2(1)
    o = _osim(c)
    f = O.engine_fields(o)
    e = pq.Engine(f)
    K, KU = _kernel_rows(f)
    checked = 0
    for x, y in list(zip(xs, ys))[:6]:
        v = O.engine_values(o, x, y)
        st = O.engine_prepare(o, 0, v)
        if st is None:
            continue
        u, s, cs = st
        want = O.engine_sum_subspace(o, 0, u, s, cs)
        tol = 1e-11 * max(abs(want), 1e-3 * abs(O.engine_weight(o, 0, s)) * 2.0 ** (f["branch_dim"] / 2))
        for form in ("z8", "branch_walk", "plain_fp64"):
            got = e.sum_subspace(u, s, cs, form=form)
            assert abs(got - want) <= tol, (config, form, got, want)
        j = len(K) // 2
        s2, u2 = s ^ K[j], u ^ KU[j]
        moved = e.sum_subspace(u2, s2, O.engine_weight(o, 0, s2))
        assert abs(moved - want) <= tol
        checked += 1
    assert checked >= 4
