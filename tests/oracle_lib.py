"""ctypes binding of the CPU oracle (oracle/liboracle.so).  Test infrastructure:
used by tests/, the smoke check and bench.py's cpu_baseline leg only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# POLYQ_ORACLE_LIB selects another build of the oracle (the sanitizer run)
PATH = os.environ.get("POLYQ_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_rng_u64.restype = ctypes.c_uint64
        L.orc_rng_u64.argtypes = [ctypes.c_uint64]
        L.orc_splitmix_next.restype = ctypes.c_uint64
        L.orc_splitmix_next.argtypes = [_u64p]
        L.orc_phase_from_radians.restype = ctypes.c_uint32
        L.orc_phase_from_radians.argtypes = [ctypes.c_double]
        L.orc_phase_to_complex.argtypes = [ctypes.c_uint32, _f64p]
        L.orc_random_clifford_t.restype = ctypes.c_long
        L.orc_random_clifford_t.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _u8p, _u32p, _u32p, _u32p, ctypes.c_long]
        L.orc_random_layered.restype = ctypes.c_long
        L.orc_random_layered.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, _u8p, _u32p, _u32p, _u32p, ctypes.c_long]
        circ = [ctypes.c_uint32, ctypes.c_size_t, _u8p, _u32p, _u32p, _u32p]
        L.orc_dense_statevector.argtypes = circ + [_u8p, _f64p]
        L.orc_pathsum_brute.argtypes = circ + [_u8p, _u8p, _f64p]
        L.orc_pathsum_info.argtypes = circ + [_u64p]
        L.orc_normal_form_check.argtypes = [ctypes.c_uint32, _u8p, _u8p, _u8p, ctypes.c_int, ctypes.c_int, _u64p, _f64p, _f64p, ctypes.POINTER(ctypes.c_int)]
        L.orc_sim_create.argtypes = circ + [ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
        L.orc_sim_create_threaded.argtypes = circ + [ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_double)]
        L.orc_sim_destroy.argtypes = [ctypes.c_void_p]
        L.orc_sim_destroy.restype = None
        L.orc_sim_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_set_threads.restype = None
        L.orc_sim_stats.argtypes = [ctypes.c_void_p, _u64p]
        L.orc_sim_amplitude.argtypes = [ctypes.c_void_p, _u8p, _u8p, _f64p]
        L.orc_sim_marginal.argtypes = [ctypes.c_void_p, _u8p, _u32p, _u8p, ctypes.c_size_t, _f64p]
        L.orc_sim_sample_many.argtypes = [ctypes.c_void_p, _u8p, ctypes.c_size_t, ctypes.c_uint64, _u64p]
        L.orc_sim_amplitude_estimate.argtypes = [ctypes.c_void_p, _u8p, _u8p, ctypes.c_size_t, ctypes.c_uint64, _f64p]
        L.orc_branch_dump.argtypes = [ctypes.c_void_p, _u8p, _u8p, ctypes.c_size_t, _u64p, _u8p, _u16p, _u8p, _u8p]
        L.orc_surviving_branches.restype = ctypes.c_long
        L.orc_surviving_branches.argtypes = [ctypes.c_void_p, _u8p, _u8p, _u64p, ctypes.c_long]
        L.orc_sim_amplitudes_timed.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u8p, _u8p, ctypes.c_int, _f64p, _f64p]
        L.orc_sim_chunk_range_timed.argtypes = [ctypes.c_void_p, _u8p, _u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _f64p, _f64p, _u64p]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_engine_sizes.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p]
        L.orc_engine_export.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p, _u64p, _u64p, _u64p, _u8p, _u32p, _u64p, _u32p, _u32p,
                                        _u64p, _u64p, _u64p, _u64p]
        L.orc_engine_values.argtypes = [ctypes.c_void_p, _u8p, _u8p, _u8p]
        L.orc_engine_eval.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u8p, ctypes.c_size_t, _f64p]
        L.orc_engine_prepare.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u8p, ctypes.c_size_t, _u64p, _u64p, _f64p]
        L.orc_engine_sum_subspace.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p, _u64p, _f64p, _f64p]
        L.orc_engine_weight.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _u64p, _f64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _check(rc):
    if rc < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return rc


def bits(n, idx):
    return np.array([(idx >> q) & 1 for q in range(n)], dtype=np.uint8)


class OCircuit:
    """Gate arrays in the shared gate-code layout."""

    def __init__(self, n, ops, q0, q1, ph):
        self.n = n
        self.ops = np.asarray(ops, np.uint8)
        self.q0 = np.asarray(q0, np.uint32)
        self.q1 = np.asarray(q1, np.uint32)
        self.ph = np.asarray(ph, np.uint32)

    @classmethod
    def from_circuit(cls, c):
        ops, q0, q1, ph = c.arrays() if hasattr(c, "arrays") else (c.ops, c.q0, c.q1, c.ph)
        return cls(c.num_qubits, ops, q0, q1, ph)

    def args(self):
        return (self.n, len(self.ops), _p(self.ops, _u8p), _p(self.q0, _u32p), _p(self.q1, _u32p), _p(self.ph, _u32p))


def random_clifford_t(state, n, depth, tcount, generic):
    st = ctypes.c_uint64(state)
    cap = 4096
    ops = np.zeros(cap, np.uint8); q0 = np.zeros(cap, np.uint32); q1 = np.zeros(cap, np.uint32); ph = np.zeros(cap, np.uint32)
    k = lib().orc_random_clifford_t(ctypes.byref(st), n, depth, tcount, 1 if generic else 0, _p(ops, _u8p), _p(q0, _u32p), _p(q1, _u32p), _p(ph, _u32p), cap)
    _check(k)
    return OCircuit(n, ops[:k], q0[:k], q1[:k], ph[:k]), st.value


def random_layered(state, n, rounds):
    st = ctypes.c_uint64(state)
    cap = 4096
    ops = np.zeros(cap, np.uint8); q0 = np.zeros(cap, np.uint32); q1 = np.zeros(cap, np.uint32); ph = np.zeros(cap, np.uint32)
    k = lib().orc_random_layered(ctypes.byref(st), n, rounds, _p(ops, _u8p), _p(q0, _u32p), _p(q1, _u32p), _p(ph, _u32p), cap)
    _check(k)
    return OCircuit(n, ops[:k], q0[:k], q1[:k], ph[:k]), st.value


def dense(c, x):
    out = np.zeros(2 << c.n, np.float64)
    _check(lib().orc_dense_statevector(*c.args(), _p(bits(c.n, x), _u8p), _p(out, _f64p)))
    return out[0::2] + 1j * out[1::2]


def brute(c, x, y):
    out = np.zeros(2, np.float64)
    _check(lib().orc_pathsum_brute(*c.args(), _p(bits(c.n, x), _u8p), _p(bits(c.n, y), _u8p), _p(out, _f64p)))
    return complex(out[0], out[1])


def pathsum_info(c):
    o = np.zeros(4, np.uint64)
    _check(lib().orc_pathsum_info(*c.args(), _p(o, _u64p)))
    return dict(zip(("num_vars", "live", "num_h", "rotations"), (int(v) for v in o)))


STRAT = {"auto": 0, "cached": 1, "pairs": 2}


class OSim:
    def __init__(self, c, strategy="auto", pairs=0, threads=None):
        """threads=None: single-threaded build and evaluation.  threads=k: the
        branch engines are built and evaluated by k OpenMP workers and the
        build's wall time is kept in `build_seconds`."""
        self.c = c
        h = ctypes.c_void_p()
        self.build_seconds = None
        if threads is None:
            _check(lib().orc_sim_create(*c.args(), STRAT[strategy], pairs, ctypes.byref(h)))
        else:
            sec = ctypes.c_double()
            _check(lib().orc_sim_create_threaded(*c.args(), STRAT[strategy], pairs, int(threads), ctypes.byref(h), ctypes.byref(sec)))
            self.build_seconds = sec.value
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def set_threads(self, t):
        lib().orc_sim_set_threads(self.h, t)

    def stats(self):
        o = np.zeros(7, np.uint64)
        _check(lib().orc_sim_stats(self.h, _p(o, _u64p)))
        return tuple(int(v) for v in o)

    def amplitude(self, x, y):
        out = np.zeros(2, np.float64)
        _check(lib().orc_sim_amplitude(self.h, _p(bits(self.c.n, x), _u8p), _p(bits(self.c.n, y), _u8p), _p(out, _f64p)))
        return complex(out[0], out[1])

    def amplitude_estimate(self, x, y, samples, seed):
        out = np.zeros(3, np.float64)
        _check(lib().orc_sim_amplitude_estimate(self.h, _p(bits(self.c.n, x), _u8p), _p(bits(self.c.n, y), _u8p), samples, seed, _p(out, _f64p)))
        return complex(out[0], out[1]), float(out[2])

    def sample_many(self, x, shots, seed):
        out = np.zeros(shots, np.uint64)
        _check(lib().orc_sim_sample_many(self.h, _p(bits(self.c.n, x), _u8p), shots, seed, _p(out, _u64p)))
        return out

    def marginal(self, x, measured):
        qs = np.asarray([m[0] for m in measured], np.uint32)
        os_ = np.asarray([1 if m[1] else 0 for m in measured], np.uint8)
        out = ctypes.c_double()
        _check(lib().orc_sim_marginal(self.h, _p(bits(self.c.n, x), _u8p), _p(qs, _u32p), _p(os_, _u8p), len(qs), ctypes.byref(out)))
        return out.value

    def branch_dump(self, x, y, s_list):
        s = np.ascontiguousarray(s_list, np.uint64)
        n = len(s)
        zero = np.zeros(n, np.uint8); r = np.zeros(n, np.uint16); k8 = np.zeros(n, np.uint8); kt = np.zeros(n, np.uint8)
        rc = _check(lib().orc_branch_dump(self.h, _p(bits(self.c.n, x), _u8p), _p(bits(self.c.n, y), _u8p), n, _p(s, _u64p), _p(zero, _u8p), _p(r, _u16p), _p(k8, _u8p), _p(kt, _u8p)))
        if rc == 1:
            return None
        return zero, r, k8, kt

    def surviving(self, x, y, cap=1 << 16):
        s = np.zeros(cap, np.uint64)
        k = _check(lib().orc_surviving_branches(self.h, _p(bits(self.c.n, x), _u8p), _p(bits(self.c.n, y), _u8p), _p(s, _u64p), cap))
        return s[:k]

    def chunk_range_timed(self, x, y, hi0, hi1, threads):
        out = np.zeros(2, np.float64)
        sec = ctypes.c_double()
        br = ctypes.c_uint64()
        _check(lib().orc_sim_chunk_range_timed(self.h, _p(bits(self.c.n, x), _u8p), _p(bits(self.c.n, y), _u8p), hi0, hi1, threads, _p(out, _f64p), ctypes.byref(sec), ctypes.byref(br)))
        return complex(out[0], out[1]), sec.value, br.value

    def amplitudes_timed(self, xs, ys, threads):
        n = len(xs)
        X = np.stack([bits(self.c.n, x) for x in xs]).astype(np.uint8)
        Y = np.stack([bits(self.c.n, y) for y in ys]).astype(np.uint8)
        out = np.zeros(2 * n, np.float64)
        sec = ctypes.c_double()
        _check(lib().orc_sim_amplitudes_timed(self.h, n, _p(X, _u8p), _p(Y, _u8p), threads, _p(out, _f64p), ctypes.byref(sec)))
        return out[0::2] + 1j * out[1::2], sec.value


def normal_form_check(B, v, shift, rule, brute=True):
    m = len(v)
    Bm = np.ascontiguousarray(B, np.uint8)
    vv = np.ascontiguousarray(v, np.uint8)
    sh = np.ascontiguousarray(shift, np.uint8)
    counts = np.zeros(3, np.uint64)
    gs = np.zeros(2, np.float64); br = np.zeros(2, np.float64)
    ok = ctypes.c_int()
    _check(lib().orc_normal_form_check(m, _p(Bm, _u8p), _p(vv, _u8p), _p(sh, _u8p), rule, 1 if brute else 0, _p(counts, _u64p), _p(gs, _f64p), _p(br, _f64p), ctypes.byref(ok)))
    return tuple(int(c) for c in counts), complex(gs[0], gs[1]), complex(br[0], br[1]), bool(ok.value)


def branch_brute(sim, x, y, s_list):
    L = lib()
    L.orc_branch_brute.argtypes = [ctypes.c_void_p, _u8p, _u8p, ctypes.c_size_t, _u64p, _f64p]
    s = np.ascontiguousarray(s_list, np.uint64)
    out = np.zeros(2 * len(s), np.float64)
    rc = _check(L.orc_branch_brute(sim.h, _p(bits(sim.c.n, x), _u8p), _p(bits(sim.c.n, y), _u8p), len(s), _p(s, _u64p), _p(out, _f64p)))
    if rc == 1:
        return None
    return out[0::2] + 1j * out[1::2]


def _w(n):
    return (n + 63) // 64


def engine_fields(osim, branch=0):
    """The fields of branch engine `branch` of an oracle simulator, in the
    layout of the library's pq_engine_desc (a dict of sizes and arrays)."""
    sz = np.zeros(13, np.uint64)
    _check(lib().orc_engine_sizes(osim.h, branch, _p(sz, _u64p)))
    nf, g, nodd, ne, t, d, rk, nc, nh, glob, zero, sigma, m = (int(v) for v in sz)
    sigma = int(np.int64(np.uint64(sigma)))
    W = 2 * _w(g) + _w(nodd) + _w(ne)
    a = {
        "ucols": np.zeros(max(1, nf * W), np.uint64), "mcols": np.zeros(max(1, t * W), np.uint64),
        "diag": np.zeros(max(1, W), np.uint64), "o3": np.zeros(max(1, _w(nodd)), np.uint64),
        "fixed_v": np.zeros(max(1, nf), np.uint8), "fixed_theta": np.zeros(max(1, nf), np.uint32),
        "b_fixed": np.zeros(max(1, nf * _w(nf)), np.uint64), "theta": np.zeros(max(1, t), np.uint32),
        "pivot_col": np.zeros(max(1, rk), np.uint32), "pivot_row": np.zeros(max(1, rk * _w(ne)), np.uint64),
        "consistency": np.zeros(max(1, nc * _w(ne)), np.uint64), "kernel": np.zeros(max(1, d * _w(t)), np.uint64),
        "kernel_u": np.zeros(max(1, d * W), np.uint64),
    }
    _check(lib().orc_engine_export(osim.h, branch, _p(a["ucols"], _u64p), _p(a["mcols"], _u64p), _p(a["diag"], _u64p), _p(a["o3"], _u64p),
                                   _p(a["fixed_v"], _u8p), _p(a["fixed_theta"], _u32p), _p(a["b_fixed"], _u64p), _p(a["theta"], _u32p),
                                   _p(a["pivot_col"], _u32p), _p(a["pivot_row"], _u64p), _p(a["consistency"], _u64p), _p(a["kernel"], _u64p),
                                   _p(a["kernel_u"], _u64p)))
    a.update(nf=nf, planes=g, lines=nodd, radical=ne, terms=t, branch_dim=d, rank=rk, consistency_rows=nc, num_h=nh,
             global_phase=glob, zero=zero, sigma=sigma, free_vars=m)
    return a


def engine_values(osim, x, y):
    """The fixed values of <y|C|x> aligned with the engines' fixed variables,
    or None for a contradictory boundary."""
    sz = np.zeros(13, np.uint64)
    _check(lib().orc_engine_sizes(osim.h, 0, _p(sz, _u64p)))
    vals = np.zeros(max(1, int(sz[0])), np.uint8)
    rc = _check(lib().orc_engine_values(osim.h, _p(bits(osim.c.n, x), _u8p), _p(bits(osim.c.n, y), _u8p), _p(vals, _u8p)))
    return None if rc == 1 else vals[:int(sz[0])]


def engine_eval(osim, branch, vals):
    """Engine::eval (engine.rs:551-561) of one branch engine."""
    out = np.zeros(2, np.float64)
    v = np.ascontiguousarray(vals, np.uint8)
    _check(lib().orc_engine_eval(osim.h, branch, _p(v, _u8p), len(v), _p(out, _f64p)))
    return complex(out[0], out[1])


def _start_words(osim, branch):
    sz = np.zeros(13, np.uint64)
    _check(lib().orc_engine_sizes(osim.h, branch, _p(sz, _u64p)))
    g, nodd, ne, t = (int(v) for v in sz[1:5])
    return 2 * _w(g) + _w(nodd) + _w(ne), _w(t)


def engine_prepare(osim, branch, vals):
    """Engine::prepare (engine.rs:569-610): (u_start, s_start, c_start), the
    shift row and branch vector packed as in the descriptor layout, or None."""
    uw, sw = _start_words(osim, branch)
    u = np.zeros(max(1, uw), np.uint64); s = np.zeros(max(1, sw), np.uint64); c = np.zeros(2, np.float64)
    v = np.ascontiguousarray(vals, np.uint8)
    rc = _check(lib().orc_engine_prepare(osim.h, branch, _p(v, _u8p), len(v), _p(u, _u64p), _p(s, _u64p), _p(c, _f64p)))
    return None if rc == 1 else (u, s, complex(c[0], c[1]))


def engine_sum_subspace(osim, branch, u, s, c):
    """Engine::sum_subspace (engine.rs:669-693) from explicit start arguments."""
    out = np.zeros(2, np.float64)
    cc = np.array([complex(c).real, complex(c).imag], np.float64)
    u = np.ascontiguousarray(u, np.uint64); s = np.ascontiguousarray(s, np.uint64)
    _check(lib().orc_engine_sum_subspace(osim.h, branch, _p(u, _u64p), _p(s, _u64p), _p(cc, _f64p), _p(out, _f64p)))
    return complex(out[0], out[1])


def engine_weight(osim, branch, s):
    """Engine::weight (engine.rs:661-667) of a packed branch vector."""
    out = np.zeros(2, np.float64)
    s = np.ascontiguousarray(s, np.uint64)
    _check(lib().orc_engine_weight(osim.h, branch, _p(s, _u64p), _p(out, _f64p)))
    return complex(out[0], out[1])
