"""The allocation-free entry point (pq_sim_amplitudes_noalloc) against
pq_sim_amplitudes through the C ABI: the same results bit for bit, on two
calls in a row (the second smaller, so stale scratch contents would show),
in both batched forms, and a batch beyond the scratch size is refused."""
import ctypes

import numpy as np
import pytest

import polyq_amd as pq
from polyq_amd import workloads as W

pytestmark = pytest.mark.gpu

lib = pq.lib
lib.pq_sim_amplitudes_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
lib.pq_noalloc_scratch_bytes.restype = ctypes.c_size_t


def _call(fn, sim, xs, ys, form):
    xs = np.ascontiguousarray(xs, dtype=np.uint64)
    ys = np.ascontiguousarray(ys, dtype=np.uint64)
    out = np.zeros(2 * len(xs), dtype=np.float64)
    rc = fn(sim.handle, len(xs), xs.ctypes.data, ys.ctypes.data, pq.FORMS[form], out.ctypes.data)
    if rc != 0:
        raise pq.PolyqError(lib.pq_last_error().decode())
    return out.view(np.complex128)


@pytest.mark.parametrize("form", ["z8", "branch_walk"])
def test_noalloc_equals_allocating_entry_point(form):
    assert lib.pq_noalloc_scratch_bytes() == 8 << 20
    c, xs, ys = W.config2(1)
    sim = pq.Simulator(c, strategy="cached")
    rng = W.Rng(9)
    big_x = [rng.below(1 << 40) for _ in range(200)]
    big_y = [rng.below(1 << 40) for _ in range(200)]
    want_big = _call(lib.pq_sim_amplitudes, sim, big_x, big_y, form)      # uploads the engines
    got_big = _call(lib.pq_sim_amplitudes_noalloc, sim, big_x, big_y, form)
    np.testing.assert_array_equal(got_big, want_big)
    # a second, smaller call right after: nothing of the first may leak in
    got_small = _call(lib.pq_sim_amplitudes_noalloc, sim, xs[:3], ys[:3], form)
    want_small = _call(lib.pq_sim_amplitudes, sim, xs[:3], ys[:3], form)
    np.testing.assert_array_equal(got_small, want_small)
    assert np.all(np.abs(want_small) > 0)


def test_noalloc_needs_resident_engines_and_bounds_the_batch():
    c, xs, ys = W.config2(2)
    sim = pq.Simulator(c, strategy="cached")
    with pytest.raises(pq.PolyqError, match="not resident"):
        _call(lib.pq_sim_amplitudes_noalloc, sim, xs[:1], ys[:1], "z8")
    _call(lib.pq_sim_amplitudes, sim, xs[:1], ys[:1], "z8")
    # 8192 queries of this engine (d = 24, 10 inner slots, no mid walk: 64
    # items of 256 lanes each) need 16 MiB of item partials
    rng = W.Rng(10)
    n = 8192
    bx = [rng.below(1 << 40) for _ in range(n)]
    prev = pq.get_walk_tuning("branch_walk")
    pq.set_walk_tuning(27, 6, "branch_walk")
    try:
        with pytest.raises(pq.PolyqError, match="scratch"):
            _call(lib.pq_sim_amplitudes_noalloc, sim, bx, [0] * n, "branch_walk")
    finally:
        pq.set_walk_tuning(*prev, form="branch_walk")
    # and a batch that fits still runs correctly afterwards
    got = _call(lib.pq_sim_amplitudes_noalloc, sim, xs[:4], ys[:4], "z8")
    np.testing.assert_array_equal(got, _call(lib.pq_sim_amplitudes, sim, xs[:4], ys[:4], "z8"))
