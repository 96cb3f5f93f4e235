"""The CPU oracle pinned against the reference's own tests, known answers and
independent mathematics (no GPU)."""
import math

import numpy as np
import pytest

import oracle_lib as O
from polyq_amd import workloads as W


# ---------------------------------------------------------------- phase.rs
def test_phase_known_answers():
    # src/phase.rs:154-182
    L = O.lib()
    assert L.orc_phase_from_radians(-math.pi / 4) == (2**32 - 2**29)
    for k in [0, 1, 12345, 1 << 29, (1 << 31) + 7, 2**32 - 1]:
        assert L.orc_phase_from_radians(2 * math.pi * k / 2**32) == k
    out = np.zeros(2)
    L.orc_phase_to_complex(1 << 31, out.ctypes.data_as(O._f64p))
    assert abs(out[0] + 1) < 1e-15 and abs(out[1]) < 1e-15
    L.orc_phase_to_complex(1 << 29, out.ctypes.data_as(O._f64p))
    assert out[0] == out[1] == 0.70710678118654752440   # exact ROOT8 (phase.rs:96-108)


def test_splitmix64_published_answer():
    # SplitMix64 from state 0: first output 0xE220A8397B1DCDAF (published algorithm)
    import ctypes
    st = ctypes.c_uint64(0)
    assert O.lib().orc_splitmix_next(ctypes.byref(st)) == 0xE220A8397B1DCDAF
    # Rng::new(seed) pre-adds gamma (rng.rs:13-15)
    assert O.lib().orc_rng_u64(0) == W.Rng(0).next_u64()


# ---------------------------------------------------------------- dense.rs
def _np_statevector(c, x):
    """Independent numpy statevector by explicit 2x2 / 4x4 gate matrices."""
    n = c.n
    psi = np.zeros((2,) * n, complex)
    psi[tuple((x >> q) & 1 for q in reversed(range(n)))] = 1
    h = 1 / np.sqrt(2)

    def one(m, q):
        nonlocal psi
        ax = n - 1 - q
        psi = np.moveaxis(np.tensordot(m, psi, axes=([1], [ax])), 0, ax)

    def ph(p):
        return np.exp(2j * np.pi * int(p) / 2**32)

    for op, a, b, p in zip(c.ops, c.q0, c.q1, c.ph):
        a, b = int(a), int(b)
        if op == 0: one(np.array([[h, h], [h, -h]]), a)
        elif op == 1: one(np.array([[0, 1], [1, 0]]), a)
        elif op == 2: one(np.array([[0, -1j], [1j, 0]]), a)
        elif op == 3: one(np.diag([1, -1]), a)
        elif op == 4: one(np.diag([1, 1j]), a)
        elif op == 5: one(np.diag([1, -1j]), a)
        elif op == 6: one(np.diag([1, np.exp(1j * np.pi / 4)]), a)
        elif op == 7: one(np.diag([1, np.exp(-1j * np.pi / 4)]), a)
        elif op == 8: one(np.diag([1, ph(p)]), a)
        elif op == 12: psi = psi * ph(p)
        else:
            flat = psi.reshape(-1).copy()
            idx = np.arange(1 << n)
            ba = (idx >> a) & 1; bb = (idx >> b) & 1
            if op == 9:
                flat[(ba & bb) == 1] *= -1
            elif op == 10:
                src = np.where(ba == 1, idx ^ (1 << b), idx)
                flat = flat[src]
            else:
                src = np.where(ba != bb, idx ^ (1 << a) ^ (1 << b), idx)
                flat = flat[src]
            psi = flat.reshape((2,) * n)
    return psi.reshape(-1)


def test_dense_matches_independent_numpy():
    st = 4242
    for trial in range(40):
        c, st = O.random_clifford_t(st, 1 + trial % 5, 25, trial % 6, trial % 2 == 0)
        for x in range(min(4, 1 << c.n)):
            np.testing.assert_allclose(O.dense(c, x), _np_statevector(c, x), atol=1e-12)


# ------------------------------------------------------------- pathsum.rs
def test_pathsum_matches_dense_random():
    # src/pathsum.rs:504-530: seed 7, 60 trials
    rng = W.Rng(7)
    checked = 0
    for trial in range(60):
        n = 1 + rng.below(3)
        c = W.random_clifford_t(n, 6 + rng.below(8), rng.below(4), trial % 2 == 0, rng)
        oc = O.OCircuit.from_circuit(c)
        if O.pathsum_info(oc)["live"] > 22:
            continue
        checked += 1
        for x in range(1 << n):
            sv = O.dense(oc, x)
            for y in range(1 << n):
                assert abs(O.brute(oc, x, y) - sv[y]) < 1e-9
    assert checked >= 40


# --------------------------------------------------------- normal_form.rs
def _random_form(m, density, rng):
    B = np.zeros((m, m), np.uint8)
    for i in range(m):
        for j in range(i + 1, m):
            if rng.next_f64() < density:
                B[i, j] ^= 1; B[j, i] ^= 1
    v = np.array([rng.below(4) for _ in range(m)], np.uint8)
    return B, v


def test_basis_block_diagonalises_gram():
    # normal_form.rs:554-606 (seed 1, 200 forms, m <= 40, both pivot rules)
    rng = W.Rng(1)
    for trial in range(200):
        m = 1 + rng.below(40)
        B, v = _random_form(m, [0.05, 0.2, 0.5][trial % 3], rng)
        for rule in (0, 1):
            (g, odd, even), _, _, ok = O.normal_form_check(B, v, np.zeros(m, np.uint8), rule, brute=False)
            assert 2 * g + odd + even == m and ok


def test_gauss_sum_matches_brute_force():
    # normal_form.rs:608-625 (seed 2, 400 forms, m <= 11, random shift)
    rng = W.Rng(2)
    for trial in range(400):
        m = 1 + rng.below(11)
        B, v = _random_form(m, [0.1, 0.3, 0.6][trial % 3], rng)
        shift = np.array([1 if rng.next_f64() < 0.4 else 0 for _ in range(m)], np.uint8)
        for rule in (0, 1):
            _, got, want, _ = O.normal_form_check(B, v, shift, rule)
            assert abs(got - want) < 1e-9


def test_alternating_forms_have_no_lines():
    # normal_form.rs:627-645
    rng = W.Rng(3)
    for _ in range(100):
        m = 2 + rng.below(9)
        B, v = _random_form(m, 0.4, rng)
        v &= 2
        (g, odd, even), got, want, _ = O.normal_form_check(B, v, np.zeros(m, np.uint8), 1)
        assert odd == 0 and abs(got - want) < 1e-9


# -------------------------------------------------------------- engine.rs
def test_engine_matches_brute_force():
    # engine.rs:745-774 (seed 21, 120 trials, >= 60 checked)
    rng = W.Rng(21)
    checked = 0
    for trial in range(120):
        n = 1 + rng.below(4)
        c = W.random_clifford_t(n, 8 + rng.below(14), rng.below(6), trial % 3 == 0, rng)
        oc = O.OCircuit.from_circuit(c)
        if O.pathsum_info(oc)["live"] > 20:
            continue
        sim = O.OSim(oc, "cached")
        checked += 1
        for x in range(1 << n):
            for y in range(1 << n):
                assert abs(sim.amplitude(x, y) - O.brute(oc, x, y)) < 1e-8
    assert checked >= 60


def test_bell_ghz_and_htzh_known_answers():
    # engine.rs:776-802
    from polyq_amd import Circuit
    c = Circuit(3); c.h(0).cx(0, 1).cx(1, 2)
    s = O.OSim(O.OCircuit.from_circuit(c), "cached")
    for y in range(8):
        want = 1 / math.sqrt(2) if y in (0, 7) else 0.0
        assert abs(s.amplitude(0, y) - want) < 1e-12
    c = Circuit(1); c.h(0).t(0).z(0).h(0)
    s = O.OSim(O.OCircuit.from_circuit(c), "cached")
    w = complex(math.cos(math.pi / 4), math.sin(math.pi / 4))
    assert abs(s.amplitude(0, 0) - (1 - w) / 2) < 1e-12
    assert abs(s.amplitude(0, 1) - (1 + w) / 2) < 1e-12


def test_branch_gauss_sums_match_brute_force():
    # per-branch (zero, r, k8) of the oracle vs the 2^m brute-force Gauss sum at
    # the same shift, on every t-bit branch vector of small engines
    rng = W.Rng(77)
    checked = 0
    for trial in range(60):
        n = 1 + rng.below(3)
        c = W.random_clifford_t(n, 10 + rng.below(10), 1 + rng.below(5), trial % 2 == 0, rng)
        oc = O.OCircuit.from_circuit(c)
        sim = O.OSim(oc, "cached")
        free, t = sim.stats()[0], sim.stats()[4]
        if free > 16 or t == 0 or t > 6:
            continue
        for x in range(1 << n):
            y = rng.below(1 << n)
            S = np.arange(1 << t, dtype=np.uint64)
            d = sim.branch_dump(x, y, S)
            if d is None:
                continue
            zero, r, k8, _ = d
            G = O.branch_brute(sim, x, y, S)
            for i in range(len(S)):
                if zero[i]:
                    assert abs(G[i]) < 1e-9
                else:
                    want = 2 ** (r[i] / 2) * complex(math.cos(math.pi * k8[i] / 4), math.sin(math.pi * k8[i] / 4))
                    assert abs(G[i] - want) < 1e-9 * max(1, abs(want))
            checked += 1
    assert checked >= 20


# ---------------------------------------------------------------- rank.rs
def test_rank_engine_matches_plain_engine():
    # rank.rs:243-276 (seed 5)
    rng = W.Rng(5)
    with_pairs = 0
    for trial in range(60):
        n = 2 + rng.below(3)
        c = W.random_clifford_t(n, 20, 4 + rng.below(5), trial % 4 == 0, rng) if trial % 2 == 0 else W.random_layered(n, 3, rng)
        oc = O.OCircuit.from_circuit(c)
        if O.pathsum_info(oc)["live"] > 20:
            continue
        plain = O.OSim(oc, "cached")
        for pairs in (1, 2, 3, 6):
            r = O.OSim(oc, "pairs", pairs)
            if r.stats()[6] > 0:
                with_pairs += 1
            for x in range(1 << n):
                for y in range(1 << n):
                    assert abs(plain.amplitude(x, y) - r.amplitude(x, y)) < 1e-8
    assert with_pairs >= 20


# ----------------------------------------------------------------- sim.rs
def test_statevector_matches_dense_all_strategies():
    # sim.rs:329-342 (seed 99, 50 circuits)
    rng = W.Rng(99)
    for trial in range(50):
        n = 1 + rng.below(5)
        c = W.random_clifford_t(n, 25, rng.below(8), trial % 4 == 0, rng) if trial % 2 == 0 else W.random_layered(n, 2 + rng.below(2), rng)
        x = sum((1 << q) for q in range(n) if rng.next_f64() < 0.5)
        oc = O.OCircuit.from_circuit(c)
        want = O.dense(oc, x)
        for strat, p in (("cached", 0), ("auto", 0), ("pairs", 2), ("pairs", 100)):
            s = O.OSim(oc, strat, p)
            got = np.array([s.amplitude(x, y) for y in range(1 << n)])
            np.testing.assert_allclose(got, want, atol=1e-8)


def test_marginals_match_dense():
    # sim.rs:354-379 (seed 3)
    rng = W.Rng(3)
    for _ in range(12):
        n = 2 + rng.below(3)
        c = W.random_layered(n, 2, rng)
        oc = O.OCircuit.from_circuit(c)
        sv = O.dense(oc, 0)
        s = O.OSim(oc, "auto")
        for _ in range(4):
            k = 1 + rng.below(n)
            qs = list(range(n))
            for i in range(n - 1, 0, -1):
                j = rng.below(i + 1)
                qs[i], qs[j] = qs[j], qs[i]
            meas = [(q, rng.next_f64() < 0.5) for q in qs[:k]]
            want = sum(abs(a) ** 2 for i, a in enumerate(sv) if all(((i >> q) & 1) == b for q, b in meas))
            assert abs(s.marginal(0, meas) - want) < 1e-9


def _sampling_circuit():
    # sim.rs:383-398 (sampling_follows_the_distribution)
    import polyq_amd as pq
    c = pq.Circuit(3)
    c.h(0).t(0).h(0).cx(0, 1).h(2).t(2).h(2).cz(1, 2).h(1).t(1).h(1)
    return c


def test_sampling_follows_the_distribution():
    # sim.rs:383-398: 40000 shots (seed 1234) within 0.012 of |amplitude|^2
    oc = O.OCircuit.from_circuit(_sampling_circuit())
    sv = O.dense(oc, 0)
    shots = 40000
    got = O.OSim(oc, "auto").sample_many(0, shots, 1234)
    hist = np.bincount(got.astype(np.int64), minlength=8)
    for y in range(8):
        assert abs(abs(sv[y]) ** 2 - hist[y] / shots) < 0.012, y


def test_sampling_draws_only_possible_outcomes_and_is_reproducible():
    # every drawn outcome has nonzero probability; the same seed gives the
    # same shots, another seed other shots (wider random circuits)
    rng = W.Rng(17)
    for trial in range(6):
        n = 3 + rng.below(4)
        c = W.random_layered(n, 2, rng)
        oc = O.OCircuit.from_circuit(c)
        x = rng.below(1 << n)
        sv = O.dense(oc, x)
        s = O.OSim(oc, "cached")
        a = s.sample_many(x, 300, 5 + trial)
        assert all(abs(sv[int(y)]) ** 2 > 1e-12 for y in a)
        np.testing.assert_array_equal(a, s.sample_many(x, 300, 5 + trial))
        assert not np.array_equal(a, s.sample_many(x, 300, 6 + trial)) or np.max(np.abs(sv) ** 2) > 0.999


def _estimate_cases():
    # H-sandwiched Clifford+T layers (the config-2 generator shape): the branch
    # dimension follows the T count and the amplitudes are nonzero
    rng = W.Rng(23)
    out = []
    for trial in range(6):
        n = 9 + rng.below(4)
        c = W.sandwich_layer(n, 11 + rng.below(5), (15 * n) // 2, (5 * n) // 2, rng)
        out.append((c, 0, rng.below(1 << n)))
    return out


def test_estimate_is_exact_when_the_subspace_fits():
    # engine.rs:622-624: 2^d <= samples gives the exact sum and error 0
    for c, x, y in _estimate_cases()[:4]:
        oc = O.OCircuit.from_circuit(c)
        s = O.OSim(oc, "cached")
        d = s.stats()[5]
        est, se = s.amplitude_estimate(x, y, 1 << d, 3)
        assert se == 0.0
        assert abs(est - s.amplitude(x, y)) < 1e-12


@pytest.mark.parametrize("strategy", ["cached", "pairs"])
def test_estimate_is_unbiased_within_its_error(strategy):
    # engine.rs:612-659 / rank.rs:149-171: over many seeds the estimates
    # scatter around the exact amplitude with the stated standard error
    for c, x, y in _estimate_cases():
        oc = O.OCircuit.from_circuit(c)
        s = O.OSim(oc, strategy, 3)
        exact = s.amplitude(x, y)
        d = s.stats()[5]
        if d < 8:
            continue
        zs = []
        for seed in range(12):
            est, se = s.amplitude_estimate(x, y, 64, seed)
            if se > 0:
                zs.append(abs(est - exact) / se)
        if zs:
            assert max(zs) < 6 and np.mean(zs) < 2.5, zs


def test_toffoli_and_circuit_known_answers():
    # circuit.rs:233-244 and sim.rs:344-351
    from polyq_amd import Circuit
    for inp in range(8):
        c = Circuit(3); c.ccx(0, 1, 2)
        sv = O.dense(O.OCircuit.from_circuit(c), inp)
        expect = inp ^ 4 if (inp & 1 and inp & 2) else inp
        assert abs(abs(sv[expect]) - 1) < 1e-9
        assert c.t_count() == 7
    c = Circuit(4); c.h(0).h(1).ccx(0, 1, 2).cx(2, 3).ccz(0, 2, 3).h(3)
    oc = O.OCircuit.from_circuit(c)
    s = O.OSim(oc, "auto")
    np.testing.assert_allclose([s.amplitude(0, y) for y in range(16)], O.dense(oc, 0), atol=1e-9)


def test_timed_full_range_is_the_amplitude():
    # the bench's CPU leg times whole coset ranges; over [0, nhi) it must be the amplitude
    rng = W.Rng(31)
    for trial in range(6):
        c = W.sandwich_layer(10, 8 + trial, 40, 12, rng)
        o = O.OSim(O.OCircuit.from_circuit(c), "cached")
        d = o.stats()[5]
        x, y = rng.below(1 << 10), rng.below(1 << 10)
        a, sec, br = o.chunk_range_timed(x, y, 0, 1 << (d - min(d, 14)), 2)
        assert abs(a - o.amplitude(x, y)) < 1e-12
        assert br == (1 << d) or br == 0


def test_parallel_rank_engine_equals_serial():
    # rank.rs:92-99 / :184-186 build and evaluate 8 or more branch engines in
    # parallel; the oracle's threaded build and evaluation must give the same
    # engines and (engine-ordered) sums, bit for bit, as the serial ones
    rng = W.Rng(77)
    checked = 0
    for trial in range(8):
        c = W.sandwich_layer(10, 10 + trial, 40, 12, rng)
        oc = O.OCircuit.from_circuit(c)
        for strat, p in (("pairs", 4), ("auto", 0)):
            serial = O.OSim(oc, strat, p)
            par = O.OSim(oc, strat, p, threads=4)
            assert par.stats() == serial.stats()
            assert par.build_seconds is not None and par.build_seconds >= 0
            xs = [rng.below(1 << 10) for _ in range(4)]
            ys = [rng.below(1 << 10) for _ in range(4)]
            got, _ = par.amplitudes_timed(xs, ys, 4)
            for i in range(4):
                assert got[i] == serial.amplitude(xs[i], ys[i])
            checked += serial.stats()[6] >= 3
    assert checked >= 4
