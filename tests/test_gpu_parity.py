"""HIP path vs the CPU oracle through the C ABI (MI355X)."""
import math

import numpy as np
import pytest

import oracle_lib as O
import polyq_amd as pq
from polyq_amd import workloads as W

pytestmark = pytest.mark.gpu

DF_REL = 1e-11       # segment walk (FP64-built tables, double-float lanes): relative to the largest amplitude
BW_REL = 1e-12       # per-branch walk (FP64 accumulation): relative to the largest amplitude
FP64_ABS = 1e-10     # plain FP64 form
ABS_TOL = {"z8": 1e-10, "branch_walk": 1e-10, "plain_fp64": FP64_ABS}
REL_TOL = {"z8": DF_REL, "branch_walk": BW_REL, "plain_fp64": 1e-12}


def _close(got, want, rel, floor):
    return abs(got - want) <= rel * max(abs(want), floor)


def test_device_present():
    info = pq.device_info()
    print("device:", info)
    assert "gfx950" in info


@pytest.mark.parametrize("form", ["z8", "branch_walk", "plain_fp64"])
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
            assert abs(a - want) < ABS_TOL[form], (trial, x, y, a, want)
        checked += 1
    assert checked >= 30


def test_config1_against_dense():
    c, xs, ys = W.config1(1)
    oc = O.OCircuit.from_circuit(c)
    want = np.array([O.dense(oc, x)[y] for x, y in zip(xs, ys)])
    for form in ("z8", "branch_walk", "plain_fp64"):
        s = pq.Simulator(c, strategy="cached", form=form)
        got = s.amplitudes(xs, ys)
        np.testing.assert_allclose(got, want, atol=1e-10)


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
            np.testing.assert_allclose(got, want, atol=1e-10)


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
                assert abs(s.marginal(0, meas, form=form) - want) < 1e-10


def test_config2_small_seed_against_oracle():
    # the config-2 generator at seed 1 (d = 19): all 16 amplitudes vs the oracle
    c, xs, ys = W.config2(1)
    o = O.OSim(O.OCircuit.from_circuit(c), "cached")
    o.set_threads(8)
    want = np.array([o.amplitude(x, y) for x, y in zip(xs, ys)])
    scale = np.max(np.abs(want))
    for form in ("z8", "branch_walk", "plain_fp64"):
        got = pq.Simulator(c, strategy="cached", form=form).amplitudes(xs, ys)
        err = np.max(np.abs(got - want)) / scale
        print(form, "max err / max|a| =", err)
        assert err < REL_TOL[form]
        if form == "z8":
            # per amplitude, relative to its own magnitude
            rel = np.abs(got - want) / np.abs(want)
            print("z8 per-amplitude max rel =", rel.max())
            assert rel.max() < 1e-9


def test_branch_walk_paths_and_walk_lengths_agree():
    # the per-branch walk at several mid-walk lengths, on the kernel's paths:
    # pivot-free mid slots (fast), mid slots sharing pivots (generic), inner
    # pivots with the four low inner slots pivot-free (lane walk) or not (one
    # point at a time); force each walk length in turn on
    # circuits of branch dimension 7..14 and on the config-2 generator at seed 1
    # (d = 19), and check both against the FP64 form and the oracle
    rng = W.Rng(17)
    saved = pq.get_walk_tuning("branch_walk")
    paths = {"fast": 0, "generic": 0, "pivots_lane_walk": 0, "pivots_per_point": 0}
    try:
        cases = 0
        for trial in range(12):
            c = W.sandwich_layer(14 + rng.below(10), 14 + rng.below(20), 60, 20, rng)
            sims = {f: pq.Simulator(c, strategy="cached", form=f) for f in ("branch_walk", "plain_fp64")}
            d = sims["branch_walk"].stats()[5]
            zi = sims["branch_walk"].z8_info()
            if zi["inner_pivots"]:
                paths["pivots_lane_walk" if zi["low_slots_pivot_free"] else "pivots_per_point"] += 1
            for off, mx in ((8, 6), (8, 1)):
                m = pq.branch_walk_mid_bits(d, off, mx, zi["inner_slots"])
                if m:     # pivot-free mid slots take the kernel's fast path
                    paths["fast" if zi["inner_pivots"] == 0 and m <= zi["pivot_free_mid"] else "generic"] += 1
            xs = [rng.below(1 << 10) for _ in range(6)]
            ys = [rng.below(1 << 10) for _ in range(6)]
            want = sims["plain_fp64"].amplitudes(xs, ys)
            res = {}
            for tuning in ((8, 6), (8, 1), (64, 0)):      # pairs over up to 2^6 and 2^1 points, single points
                pq.set_walk_tuning(*tuning, form="branch_walk")
                res[tuning] = sims["branch_walk"].amplitudes(xs, ys)
            scale = max(np.max(np.abs(want)), 1e-300)
            for tuning, got in res.items():
                err = np.max(np.abs(got - want)) / scale
                assert err < BW_REL, (trial, tuning, err)
            if trial < 3:
                o = O.OSim(O.OCircuit.from_circuit(c), "cached")
                ref = np.array([o.amplitude(x, y) for x, y in zip(xs, ys)])
                assert np.max(np.abs(res[(8, 6)] - ref)) <= BW_REL * max(np.max(np.abs(ref)), 1e-300)
            cases += d > zi["inner_slots"]
        assert cases >= 6
        print("mid walks by kernel path:", paths)
        assert paths["fast"] >= 2 and paths["generic"] >= 2
        assert paths["pivots_lane_walk"] >= 2 and paths["pivots_per_point"] >= 2
        c, xs, ys = W.config2(1)
        want = pq.Simulator(c, strategy="cached", form="plain_fp64").amplitudes(xs, ys)
        s = pq.Simulator(c, strategy="cached", form="branch_walk")
        for tuning in ((8, 6), (12, 3), (64, 0)):
            pq.set_walk_tuning(*tuning, form="branch_walk")
            got = s.amplitudes(xs, ys)
            err = np.max(np.abs(got - want)) / np.max(np.abs(want))
            print("config2 seed 1, tuning", tuning, "max err / max|a| =", err)
            assert err < BW_REL
    finally:
        pq.set_walk_tuning(saved[0], saved[1], form="branch_walk")


def test_normalisation_sum_of_probabilities():
    # size-independent property: sum_y |<y|C|0>|^2 = 1
    rng = W.Rng(9)
    c = W.sandwich_layer(10, 12, 40, 10, rng)
    s = pq.Simulator(c, strategy="cached")
    a = s.amplitudes([0] * 1024, list(range(1024)))
    assert abs(np.sum(np.abs(a) ** 2) - 1) < 1e-10


def _seg_circuits(seed, count, wide=False):
    # wide: the config-2 densities (7.5n two-qubit, 2.5n S gates) at n = 12..17,
    # whose engines have enough pivot-free directions left for lane walks
    rng = W.Rng(seed)
    out = []
    for trial in range(count):
        if wide:
            n = 12 + rng.below(6)
            c = W.sandwich_layer(n, 14 + rng.below(8), (15 * n) // 2, (5 * n) // 2, rng)
        else:
            n = 4 + rng.below(5)
            if trial % 3 == 0:
                c = W.random_clifford_t(n, 40 + rng.below(40), 10 + rng.below(8), trial % 2 == 0, rng)
            else:
                c = W.sandwich_layer(n, 12 + rng.below(8), 3 * n, n, rng)
        out.append((c, rng.below(1 << n), rng.below(1 << n)))
    return out


@pytest.mark.parametrize("bits", [2, 4, 7])
def test_segment_walk_state_bit_exact(bits):
    # device seed + Gray flips == host restatement at every outer point, for
    # several lane walk lengths (integer phase, sign mask, inner pivots)
    old = pq.get_segment_bits()
    pq.set_segment_bits(bits)
    try:
        checked = walked = 0
        for c, x, y in _seg_circuits(50 + bits, 24) + _seg_circuits(70 + bits, 24, wide=True):
            s = pq.Simulator(c, strategy="cached")
            info = s.seg_info()
            if info["outer"] > 20:
                continue
            for mid in (0, 2, 3, 5):
                dev = s.seg_dump(x, y, mid)
                if dev is None:
                    break
                ref = s.seg_lane_ref(x, y, mid)
                for a, b in zip(dev, ref):
                    np.testing.assert_array_equal(a, b)
                checked += 1
                walked += int(min(mid, info["walkable"], info["outer"] - info["lane_low"]) > 0)
        assert checked >= 40 and walked >= 10, (checked, walked)
    finally:
        pq.set_segment_bits(old)


@pytest.mark.parametrize("bits", [0, 3, 6, 10, 14])
def test_segment_form_matches_fp64_for_every_table_size(bits):
    # the same amplitudes whatever split of inner / outer coordinates, pivots
    # in the table or not, walked or seeded
    old = pq.get_segment_bits()
    pq.set_segment_bits(bits)
    try:
        for c, x, y in _seg_circuits(60 + bits, 16) + _seg_circuits(80 + bits, 8, wide=True):
            n = c.num_qubits
            xs = [x] * 8
            ys = [(y + 37 * i) % (1 << n) for i in range(8)]
            z = pq.Simulator(c, strategy="cached")
            got = z.amplitudes(xs, ys)
            want = pq.Simulator(c, strategy="cached", form="plain_fp64").amplitudes(xs, ys)
            scale = max(np.max(np.abs(want)), 1e-300)
            assert np.max(np.abs(got - want)) <= DF_REL * scale, (bits, z.seg_info())
            for mid in (0, 3, 6):
                pq.set_walk_tuning(0, mid)
                np.testing.assert_allclose(z.amplitudes(xs, ys), want, atol=DF_REL * scale, rtol=0)
            pq.set_walk_tuning(12, 6)
    finally:
        pq.set_segment_bits(old)
        pq.set_walk_tuning(12, 6)


def test_segment_form_config3_shape_per_amplitude():
    # n = 48 with 40 Rz(theta): every amplitude within 1e-9 of FP64, relative to itself
    c, xs, ys = W.config3(4, n_amps=64)
    got = pq.Simulator(c, strategy="cached").amplitudes(xs, ys)
    want = pq.Simulator(c, strategy="cached", form="plain_fp64").amplitudes(xs, ys)
    rel = np.abs(got - want) / np.abs(want)
    print("config3 shape: max rel", rel.max(), "median", np.median(rel))
    assert rel.max() < 1e-9


def test_headline_workload_per_amplitude_with_large_tables():
    # n = 40, 36 T (d = 30: 2^16-entry tables by default): the 16 amplitudes
    # of a headline step within 1e-9 of FP64, each relative to itself
    c, xs, ys = W.config2(5)
    z = pq.Simulator(c, strategy="cached")
    assert z.seg_info()["inner"] == 16
    got = z.amplitudes(xs, ys)
    want = pq.Simulator(c, strategy="cached", form="plain_fp64").amplitudes(xs, ys)
    rel = np.abs(got - want) / np.abs(want)
    print("headline workload: max rel", rel.max(), "median", np.median(rel))
    assert rel.max() < 1e-9


def test_marginal_with_a_normal_form_beyond_the_plain_kernel():
    # A 56-qubit circuit of the batched-workload generator whose 4-qubit
    # marginal sandwich has 1030 lines: the plain FP64 image (<= 1024 planes or
    # lines) is left out of its device blob, the segment form still runs, and
    # the plain form reports the limit instead of failing the whole engine.
    rng = W.Rng(77)
    for _ in range(563):
        c = W.config4_circuit(rng)
    s = pq.Simulator(c, strategy="cached")
    qs = [32, 44, 0, 10]
    assert s.marginal_stats(qs)[2] > 1024
    meas = [(q, True) for q in qs]
    x = 23444544006173823
    got = s.marginal(x, meas)
    o = O.OSim(O.OCircuit.from_circuit(c), "cached")
    o.set_threads(8)
    want = o.marginal(x, meas)
    assert abs(got - want) <= 1e-11 * max(abs(want), 1e-300) + 1e-300, (got, want)
    with pytest.raises(pq.PolyqError, match="too large for the plain kernel"):
        s.marginal(x, meas, form="plain_fp64")
