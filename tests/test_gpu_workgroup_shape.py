"""Workgroup shapes of the segment walk: 1, 2 or 4 waves per workgroup, so
work items of 64, 128 or 256 lanes.  At every shape the walk gives the FP64
form's amplitudes to the segment walk's precision, kernel-argument and staged
batches agree bit for bit, a query's result is the same alone and in a mixed
batch, a hipGraph replay equals the eager batch, and the item count follows
the shape.  The autotuner picks a shape among those it times and applies it."""
import numpy as np
import pytest

import polyq_amd as pq
from polyq_amd import workloads as W

pytestmark = pytest.mark.gpu

REL = 1e-11      # per amplitude, against the FP64 form


@pytest.fixture
def restore():
    tuning, shape = pq.get_walk_tuning(), pq.get_walk_shape()
    yield
    pq.set_walk_tuning(*tuning)
    pq.set_walk_shape(shape)


@pytest.fixture(scope="module")
def work():
    """Two engines (d = 30 and d = 24, n = 40) with 16 queries each and their
    FP64-form amplitudes."""
    out = []
    for seed in (5, 2):
        c, xs, ys = W.config2(seed)
        sim = pq.Simulator(c, strategy="cached")
        X = [int(x) for x in xs]
        Y = [int(y) for y in ys]
        out.append((sim, X, Y, sim.amplitudes(X, Y, form="plain_fp64")))
    return out


def _staged(sims, X, Y):
    """The same queries inside a batch of 17 or more (the staged path)."""
    pad = max(0, 17 - len(sims))
    return pq.batch_amplitudes(list(sims) + [sims[0]] * pad, list(X) + [X[0]] * pad, list(Y) + [Y[0]] * pad)[: len(sims)]


@pytest.mark.parametrize("waves", [1, 2, 4])
@pytest.mark.parametrize("tuning", [(12, 2), (8, 6)])      # short and long lane walks
def test_shape_matches_fp64_and_is_batch_independent(work, waves, tuning, restore):
    pq.set_walk_tuning(*tuning)
    pq.set_walk_shape(waves)
    assert pq.get_walk_shape() == waves
    for sim, X, Y, want in work:
        got = pq.batch_amplitudes([sim] * 16, X, Y)
        rel = np.abs(got - want) / np.abs(want)
        assert rel.max() < REL, (waves, tuning, rel.max())
        np.testing.assert_array_equal(got, _staged([sim] * 16, X, Y))
    # both engines in one batch (staged: 24 queries), against each alone
    (s1, X1, Y1, _), (s2, X2, Y2, _) = work
    S = [s1] * 12 + [s2] * 12
    mixed = pq.batch_amplitudes(S, X1[:12] + X2[:12], Y1[:12] + Y2[:12])
    alone = np.array([s.amplitudes([x], [y])[0] for s, x, y in zip(S, X1[:12] + X2[:12], Y1[:12] + Y2[:12])])
    np.testing.assert_array_equal(mixed, alone)


def test_items_follow_the_shape_and_graph_replays(work, restore):
    sim, X, Y, want = work[0]
    pq.set_walk_tuning(12, 2)
    items = {}
    for waves in (4, 2, 1):
        pq.set_walk_shape(waves)
        with pq.timed():
            eager = pq.batch_amplitudes([sim] * 16, X, Y)
            items[waves] = pq.last_timing()["items"]
        g = pq.Graph([sim] * 16, X, Y)
        for _ in range(3):
            got, _ms = g.run()
            np.testing.assert_array_equal(got, eager)
            pq.batch_amplitudes([sim] * 3, X[:3], Y[:3])       # eager work between replays
        assert np.max(np.abs(eager - want) / np.abs(want)) < REL
    # 2^H lanes per query in items of 64 x waves lanes
    info = sim.seg_info()
    lanes = 1 << (info["outer"] - info["mid_bits"])
    assert lanes >= 256
    for waves in (4, 2, 1):
        assert items[waves] == 16 * (lanes // (64 * waves)), (waves, items)


def test_autotune_picks_and_applies_a_shape(work, restore):
    sim, X, Y, want = work[0]
    off, mx, table = pq.autotune([sim] * 16, X, Y, offsets=(12,), max_mids=(2, 6), shapes=(1, 2, 4), repeats=1, rounds=1,
                                 cache=False)
    assert sorted(table) == sorted("12/%d/%d" % (m, w) for m in (2, 6) for w in (1, 2, 4))
    assert pq.get_walk_shape() in (1, 2, 4)
    assert min(table.values()) == table["%d/%d/%d" % (off, mx, pq.get_walk_shape())]
    got = sim.amplitudes(X, Y)
    assert np.max(np.abs(got - want) / np.abs(want)) < REL
    with pytest.raises(pq.PolyqError):
        pq.set_walk_shape(3)
    with pytest.raises(pq.PolyqError):
        pq.set_walk_shape(2, form="branch_walk")
