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


@pytest.mark.parametrize("bits", [14, 8, 3])
def test_segment_tables_match_direct_inner_sums(bits):
    # table entry x lane magnitude == sum of the 2^I inner branches with the
    # packed-row phase, for random outer points (host only)
    old = pq.get_segment_bits()
    pq.set_segment_bits(bits)
    try:
        rng = W.Rng(77 + bits)
        checked = 0
        for trial in range(40):
            n = 2 + rng.below(5)
            if trial % 2:
                c = W.random_clifford_t(n, 20 + rng.below(40), 2 + rng.below(14), trial % 4 == 1, rng)
            else:
                c = W.sandwich_layer(n + 3, 4 + rng.below(10), 3 * (n + 3), n + 3, rng)
            for strat in ("cached", "auto"):
                s = pq.Simulator(c, strategy=strat)
                info = s.seg_info()
                assert info["inner"] + info["inner_pivots"] <= bits
                assert info["inner"] + info["outer"] == info["d"]
                nq = c.num_qubits
                err = s.seg_selfcheck(rng.below(1 << nq), rng.below(1 << nq), trials=6, seed=trial)
                if err is not None:
                    assert err < 1e-12, (trial, strat, info, err)
                    checked += 1
        for cfg in (W.config2, W.config3):
            c, xs, ys = cfg(5) if cfg is W.config2 else cfg(4, n_amps=2)
            s = pq.Simulator(c, strategy="cached")
            assert s.seg_selfcheck(xs[0], ys[0], trials=2) < 1e-12
        assert checked > 30
    finally:
        pq.set_segment_bits(old)


@pytest.mark.parametrize("cap", [16, 17])
def test_large_segment_tables_match_direct_inner_sums(cap):
    # engines of branch dimension 30 and 32 get min(cap, d - 14) table bits
    old = pq.get_segment_bits()
    pq.set_segment_bits(cap)
    try:
        for seed, d in ((5, 30), (26, 32)):
            c, xs, ys = W.config2(seed)
            s = pq.Simulator(c, strategy="cached")
            info = s.seg_info()
            assert info["d"] == d
            assert info["inner"] + info["inner_pivots"] == min(cap, d - 14)
            assert s.seg_selfcheck(xs[0], ys[0], trials=3) < 1e-12
    finally:
        pq.set_segment_bits(old)


def test_segment_lane_state_matches_quadratic_form():
    # the host restatement of the device walk state: K at every outer point
    # equals K0 + sum l + 4 sum Q over the outer coordinates, and walking
    # (mid bits > 0) visits the same points as seeding each one directly
    rng = W.Rng(91)
    old = pq.get_segment_bits()
    pq.set_segment_bits(4)
    try:
        for trial in range(20):
            n = 3 + rng.below(4)
            c = W.sandwich_layer(n, 6 + rng.below(8), 3 * n, n, rng)
            s = pq.Simulator(c, strategy="cached")
            info = s.seg_info()
            if info["outer"] > 14:
                continue
            x, y = rng.below(1 << n), rng.below(1 << n)
            base = s.seg_lane_ref(x, y, 0)
            if base is None:
                continue
            walk = s.seg_lane_ref(x, y, 3)
            mb = min(3, info["walkable"], info["outer"] - info["lane_low"])
            # reorder the walked dump to outer-point order: point (lane, o)
            # covers the coordinates of lane bits and gray(o) on the walk slots
            K0, H0, P0 = base
            K1, H1, P1 = walk
            L6, nout = info["lane_low"], info["outer"]
            for idx in range(len(K1)):
                lane, o = idx >> mb, idx & ((1 << mb) - 1)
                g = o ^ (o >> 1)
                cc = 0
                for b in range(nout - mb):
                    if (lane >> b) & 1:
                        cc |= 1 << (b if b < L6 else L6 + mb + (b - L6))
                for m in range(mb):
                    if (g >> m) & 1:
                        cc |= 1 << (L6 + m)
                assert (K1[idx], H1[idx], P1[idx]) == (K0[cc], H0[cc], P0[cc])
    finally:
        pq.set_segment_bits(old)


def test_direct_prologue_matches_assignment_prologue():
    # the batched forms map (x, y) straight to the fixed positions; the plain
    # form goes through (variable, value) pairs as the reference does
    rng = W.Rng(1234)
    checked = 0
    for trial in range(60):
        n = 1 + rng.below(6)
        if trial % 3 == 0:
            c = W.random_clifford_t(n, 10 + rng.below(40), rng.below(12), trial % 2 == 0, rng)
        elif trial % 3 == 1:
            c = W.random_layered(n, 1 + rng.below(3), rng)
        else:
            c = W.uniform_gates(n, 5 + rng.below(30), rng) if n > 1 else W.random_layered(n, 2, rng)
        for kw in ({"strategy": "cached"}, {"pairs": 2}):
            s = pq.Simulator(c, **kw)
            xs = [x for x in range(1 << n) for _ in range(1 << n)]
            ys = [y for _ in range(1 << n) for y in range(1 << n)]
            assert s.prologue_selfcheck(xs, ys) == 0, (trial, kw)
            checked += 1
    c, xs, ys = W.config2(5)
    assert pq.Simulator(c, strategy="cached").prologue_selfcheck(xs, ys) == 0
    assert checked >= 100


@pytest.mark.parametrize("fast", [True, False])
def test_index_marshalling_to_uint64(fast, monkeypatch):
    # basis indices reach the C-ABI as one uint64 array, from lists, tuples
    # or arrays; a negative index is an error, not a wrapped value.  Lists go
    # through the CPython extension's checked conversion, or the Python one
    # where the extension is absent: both give the same values and errors.
    if fast:
        assert pq._pqfast is not None
    else:
        monkeypatch.setattr(pq, "_pqfast", None)
    for v in ([1, 2, 2**64 - 1], (0, 7), np.array([3, 4], dtype=np.int64), np.arange(5, dtype=np.uint64)[::2]):
        a = pq._u64_arg(v)
        assert list(a) == [int(x) for x in v]
    assert len(pq._u64_arg([])) == 0
    with pytest.raises(OverflowError):
        pq._u64_arg([5, -1])
    # numpy inputs: a signed array or scalar with a negative value raises
    # instead of wrapping to 2^64 - 1, and floats are not truncated
    with pytest.raises(OverflowError):
        pq._u64_arg(np.array([-1], np.int64))
    with pytest.raises(OverflowError):
        pq._u64_arg([np.int64(-1)])
    with pytest.raises(OverflowError):
        pq._index(np.int64(-1))
    with pytest.raises(OverflowError):
        pq._index(1 << 64)
    with pytest.raises(TypeError):
        pq._u64_arg(np.array([1.5, 2.0]))
    with pytest.raises(TypeError):
        pq._index(2.0)
    with pytest.raises(TypeError):
        pq._u64_arg(np.array([True]))
    assert pq._index(np.uint64(2**64 - 1)) == 2**64 - 1
    assert list(pq._u64_arg(np.array([7, 0], np.int32))) == [7, 0]
    # lists mixing Python ints and numpy integer scalars; bools, floats and
    # other objects inside a list are refused
    assert list(pq._u64_array([np.uint64(2**64 - 1), np.int32(7), 2**63, 0])) == [2**64 - 1, 7, 2**63, 0]
    assert list(pq._u64_array(range(3))) == [0, 1, 2]
    with pytest.raises(OverflowError):
        pq._u64_array([1, 1 << 64])
    for bad in ([True, 2], [np.bool_(True)], [1.0], [np.float64(2)], ["a"], [None]):
        with pytest.raises(TypeError):
            pq._u64_array(bad)


def test_circuit_extend_matches_builders_and_rejects_bad_ops():
    # pybind.rs:101-121: (name, qubits) pairs
    a = pq.Circuit(3).extend([("H", [0]), ("CX", [0, 1]), ("T", [2]), ("SDG", [1]), ("SWAP", [0, 2]), ("CZ", (1, 2)), ("TDG", [0])])
    b = pq.Circuit(3).h(0).cx(0, 1).t(2).sdg(1).swap(0, 2).cz(1, 2).tdg(0)
    for x, y in zip(a.arrays(), b.arrays()):
        np.testing.assert_array_equal(x, y)
    assert a.num_qubits == 3 and len(a) == 7
    for bad in ([("CCX", [0, 1, 2])], [("H", [0, 1])], [("CX", [0])], [("X", [-1])]):
        with pytest.raises(ValueError, match="bad op"):
            pq.Circuit(3).extend(bad)
    # the cached arrays follow later gates
    n0 = len(a.arrays()[0])
    a.h(1)
    assert len(a.arrays()[0]) == n0 + 1


def test_wide_circuits_build_and_refuse_u64_indices():
    # above 64 qubits the engine builds (no device needed), and the u64
    # basis-index entry points refuse instead of truncating the index
    c = pq.Circuit(80)
    c.h(70).cx(70, 3).t(3).h(79).cz(79, 70).t(70).h(70)
    s = pq.Simulator(c, strategy="cached")
    small = pq.Circuit(3)
    small.h(1).cx(1, 0).t(0).h(2).cz(2, 1).t(1).h(1)
    assert s.stats()[:6] == pq.Simulator(small, strategy="cached").stats()[:6]
    with pytest.raises(pq.PolyqError, match="more than 64 qubits"):
        pq._check(pq.lib.pq_sim_amplitude_index(s.handle, 0, 0, 0, (ctypes.c_double * 2)()))
    with pytest.raises(OverflowError):
        s._qubit_bytes(1 << 80)


def test_engines_with_more_than_64_terms_build_on_the_host():
    # 60 isolated H P(theta) H gadgets beside an 8-qubit Clifford+T circuit:
    # 68 terms, of which the gadgets' are constant over the branch subspace
    rng = W.Rng(3)
    main = W.random_clifford_t(8, 60, 12, False, rng)
    main.ccx(0, 1, 2)
    c = pq.Circuit(68)
    for op, a, b, ph in zip(main.ops, main.q0, main.q1, main.ph):
        c._push(op, a, b, ph)
    for q in range(8, 68):
        c.h(q).t(q).h(q)
    st = pq.Simulator(c, strategy="cached").stats()
    assert st[4] == 68 and st[5] == pq.Simulator(main, strategy="cached").stats()[5]
    assert st == O.OSim(O.OCircuit.from_circuit(c), "cached").stats()
    # the prologue of the basis-index path and of the assignment path agree
    # (the constant terms folded into each query)
    core = W.sandwich_layer(12, 40, 90, 30, rng)
    small = pq.Circuit(64)
    for op, a, b, ph in zip(core.ops, core.q0, core.q1, core.ph):
        small._push(op, a, b, ph)
    for q in range(12, 64):
        small.h(q).tdg(q).h(q)
    s = pq.Simulator(small, strategy="cached")
    assert s.stats()[4] > 64
    xs = [rng.next_u64() for _ in range(32)]
    ys = [rng.next_u64() for _ in range(32)]
    assert s.prologue_selfcheck(xs, ys) == 0


def test_best_pairs_at_hand_evaluated_points():
    # rank.rs:205-222 evaluated by hand at config-2 sizes (m = 340 free
    # variables, t = 34 terms, 17 T-class pairs available): p = 5, 7, 8, 10
    # for branch dimensions 24, 28, 30, 34
    for d, p in ((24, 5), (28, 7), (30, 8), (34, 10)):
        assert pq.best_pairs((340, 0, 0, 0, 34, d), 17, 17) == p
    # small subspaces are cheaper to walk than to split; the cap is MAX_PAIRS
    # (rank.rs:44) however large d is
    assert pq.best_pairs((340, 0, 0, 0, 34, 10), 17, 17) == 0
    assert pq.best_pairs((40, 0, 0, 0, 120, 120), 60, 60) == 14
    # the GPU refit keeps p = 0 on the headline engine
    assert pq.best_pairs((342, 2, 338, 0, 30, 30), 15, 15, gpu_model=True) == 0
