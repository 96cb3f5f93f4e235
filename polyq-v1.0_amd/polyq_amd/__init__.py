"""polyq_amd: MI355X branch evaluator for Clifford+T path sums.

Python mirror of the reference's Python surface (src/pybind.rs:15-234):
``Circuit`` builders, ``Simulator(circuit, strategy="auto"|"cached", pairs=None)``
with ``amplitude(input, output) -> (re, im)``, ``probability``, ``marginal`` and
``stats()``.  Every evaluation runs on the GPU through the C ABI of
``libpolyq_amd.so`` (include/polyq_amd.h); there is no CPU fallback, and a
missing library or device raises.
"""
import ctypes
import math
import os

import numpy as np

__all__ = ["Circuit", "Simulator", "Engine", "Graph", "batch_amplitudes", "engine_eval_batch", "rng_u64", "Phase", "lib", "last_timing",
           "device_info"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpolyq_amd.so")

if not os.path.exists(_LIB_PATH):
    raise ImportError("libpolyq_amd.so is not built: run polyq-v1.0_amd/build_lib.py (or __graft_entry__.build())")
lib = ctypes.CDLL(_LIB_PATH)
try:
    # CPython fast paths (Simulator.amplitudes through the buffer protocol,
    # checked conversion of index lists); built beside the library
    from . import _pqfast
except ImportError:
    _pqfast = None
_U64 = np.dtype(np.uint64)

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_sim_p = ctypes.c_void_p

lib.pq_last_error.restype = ctypes.c_char_p
lib.pq_sim_create.argtypes = [ctypes.c_uint32, ctypes.c_size_t, _u8p, _u32p, _u32p, _u32p, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(_sim_p)]
lib.pq_sim_destroy.argtypes = [_sim_p]
lib.pq_sim_destroy.restype = None
lib.pq_sim_stats.argtypes = [_sim_p, _u64p]
lib.pq_sim_num_qubits.argtypes = [_sim_p]
lib.pq_sim_num_qubits.restype = ctypes.c_uint32
lib.pq_sim_amplitude_index.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _f64p]
lib.pq_sim_probability.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _f64p]
lib.pq_sim_marginal_probability.argtypes = [_sim_p, ctypes.c_uint64, _u32p, _u8p, ctypes.c_size_t, ctypes.c_int, _f64p]
lib.pq_sim_amplitude.argtypes = [_sim_p, _u8p, _u8p, ctypes.c_int, _f64p]
lib.pq_sim_probability_bytes.argtypes = [_sim_p, _u8p, _u8p, ctypes.c_int, _f64p]
lib.pq_sim_marginal_probability_bytes.argtypes = [_sim_p, _u8p, _u32p, _u8p, ctypes.c_size_t, ctypes.c_int, _f64p]
# hot paths take raw addresses (ndarray.ctypes.data) rather than typed pointers
lib.pq_sim_amplitudes.argtypes = [_sim_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
lib.pq_batch_amplitudes.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
lib.pq_sim_statevector.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
lib.pq_sim_amplitude_estimate.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int, _f64p]
lib.pq_sim_sample_many.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
lib.pq_last_timing.argtypes = [_f64p]
lib.pq_set_timing.argtypes = [ctypes.c_int]
lib.pq_get_timing.argtypes = []
lib.pq_set_walk_tuning.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
lib.pq_get_walk_tuning.argtypes = [ctypes.c_int, _u32p]
lib.pq_set_walk_shape.argtypes = [ctypes.c_int, ctypes.c_uint32]
lib.pq_get_walk_shape.argtypes = [ctypes.c_int, _u32p]
lib.pq_set_segment_bits.argtypes = [ctypes.c_uint32]
lib.pq_get_segment_bits.restype = ctypes.c_uint32
lib.pq_sim_seg_info.argtypes = [_sim_p, _u64p]
lib.pq_sim_prologue_selfcheck.argtypes = [_sim_p, ctypes.c_size_t, _u64p, _u64p, _u64p]
lib.pq_sim_seg_dump.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_size_t, _u8p, _u32p, _u32p, ctypes.POINTER(ctypes.c_size_t)]
lib.pq_sim_seg_lane_ref.argtypes = lib.pq_sim_seg_dump.argtypes
lib.pq_sim_seg_selfcheck.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _f64p]
lib.pq_graph_create.argtypes = [ctypes.c_size_t, ctypes.POINTER(_sim_p), _u64p, _u64p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
lib.pq_graph_run.argtypes = [ctypes.c_void_p, _f64p, _f64p]
lib.pq_graph_destroy.argtypes = [ctypes.c_void_p]
lib.pq_graph_destroy.restype = None
lib.pq_sim_branch_dump.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t, _u64p, _u8p, _u8p, _u8p, _u32p]
lib.pq_sim_walk_dump.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_size_t, _u8p, _f32p, _u64p, ctypes.POINTER(ctypes.c_size_t)]
lib.pq_sim_selfcheck.argtypes = [_sim_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, _u64p]
lib.pq_sim_z8_info.argtypes = [_sim_p, _u32p]
lib.pq_sim_marginal_stats.argtypes = [_sim_p, _u32p, ctypes.c_size_t, _u64p]
lib.pq_device_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
lib.pq_engine_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
lib.pq_engine_destroy.argtypes = [ctypes.c_void_p]
lib.pq_engine_destroy.restype = None
lib.pq_engine_stats.argtypes = [ctypes.c_void_p, _u64p]
lib.pq_engine_eval.argtypes = [ctypes.c_void_p, _u8p, ctypes.c_size_t, ctypes.c_int, _f64p]
lib.pq_engine_eval_batch.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
lib.pq_engine_sum_subspace.argtypes = [ctypes.c_void_p, _u64p, _u64p, _f64p, ctypes.c_int, _f64p]
lib.pq_engine_branch_dump.argtypes = [ctypes.c_void_p, _u8p, ctypes.c_size_t, _u64p, _u8p, _u8p, _u8p, _u32p]
lib.pq_best_pairs.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
lib.pq_valu_probe.argtypes = [_f64p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32]
lib.pq_valu_probe_name.argtypes = [ctypes.c_int]
lib.pq_valu_probe_name.restype = ctypes.c_char_p

FORMS = {"z8": 0, "branch_walk": 1, "plain_fp64": 2}
STRATEGIES = {"auto": 0, "cached": 1, "gpu_auto": 3}

# Gate codes (src/circuit.rs:25-42)
H, X, Y, Z, S, SDG, T, TDG, P, CZ, CX, SWAP, GPHASE = range(13)


class PolyqError(RuntimeError):
    pass


def _check(rc):
    if rc < 0:
        raise PolyqError(lib.pq_last_error().decode())
    return rc


def _ptr(a, t):
    return a.ctypes.data_as(t)


# Hot-path buffers: ctypes arrays over the caller's data (an ndarray's
# .ctypes.data costs about 1 us per access; these a fraction of that).
_U64_ARRAYS = {}
_F64_ARRAYS = {}


_NOT_INDEX = (bool, np.bool_)       # integers to Python, but not basis indices


def _index(v):
    """One basis index as a Python int in [0, 2^64): integers only (Python
    ints, numpy integer scalars); a negative or non-integer index raises
    instead of wrapping or truncating on its way to a C uint64."""
    if isinstance(v, _NOT_INDEX) or not isinstance(v, (int, np.integer)):
        raise TypeError("basis index must be an integer, not %s" % type(v).__name__)
    v = int(v)
    if v < 0 or v >> 64:
        raise OverflowError("basis index %d outside [0, 2^64)" % v)
    return v


def _wide_index(v, n):
    """A basis index of an n-qubit circuit as a Python int in [0, 2^n)."""
    if isinstance(v, _NOT_INDEX) or not isinstance(v, (int, np.integer)):
        raise TypeError("basis index must be an integer, not %s" % type(v).__name__)
    v = int(v)
    if v < 0 or v >> n:
        raise OverflowError("basis index %d outside [0, 2^%d)" % (v, n))
    return v


def _u64_array(v):
    """v (a list, tuple or ndarray of basis indices) as a C-contiguous uint64
    ndarray.  Integer arrays only: a float or bool array raises TypeError, and
    a signed array with a negative entry raises OverflowError (a plain cast
    would wrap -1 to 2^64 - 1)."""
    if isinstance(v, np.ndarray):
        if v.dtype == np.uint64:
            return np.ascontiguousarray(v)
        if v.dtype.kind not in "iu":
            raise TypeError("basis indices must be an integer array, not %s" % v.dtype)
        if v.dtype.kind == "i" and v.size and int(v.min()) < 0:
            raise OverflowError("negative basis index")
        return np.ascontiguousarray(v, dtype=np.uint64)
    if _pqfast is not None:
        # the same checks in C, about 10 ns per index
        a = np.empty(len(v), dtype=np.uint64)
        _pqfast.u64_fill(v, a, _NOT_INDEX)
        return a
    # one numpy conversion of the checked Python ints (about 3x faster than
    # ctypes' per-element one for thousands of them)
    return np.array([_index(x) for x in v], dtype=np.uint64)


def _u64_arg(v):
    """A ctypes uint64 array holding v (see _u64_array), sharing memory with a
    writable C-contiguous uint64 ndarray."""
    n = len(v)
    t = _U64_ARRAYS.get(n)
    if t is None:
        t = _U64_ARRAYS[n] = ctypes.c_uint64 * n
    if not isinstance(v, np.ndarray) or v.dtype != np.uint64 or not v.flags.c_contiguous or not v.flags.writeable:
        v = _u64_array(v)
        if not v.flags.writeable:
            v = v.copy()
    return t.from_buffer(v)


def _handle_array(sims, n):
    """The simulators' handles as one uint64 array (a C array of pointers)."""
    return np.fromiter((s._hv for s in sims), dtype=np.uint64, count=n)


def _complex_out(n):
    """An n-element complex128 result: (ctypes buffer for the call, ndarray view)."""
    t = _F64_ARRAYS.get(n)
    if t is None:
        t = _F64_ARRAYS[n] = ctypes.c_double * (2 * n)
    buf = t()
    return buf, np.frombuffer(buf, dtype=np.complex128)


class Phase:
    """Angle 2*pi*k/2^32 (src/phase.rs:22-52)."""

    ZERO, PI, HALF_PI, QUARTER_PI = 0, 1 << 31, 1 << 30, 1 << 29

    @staticmethod
    def from_radians(r):
        turns = r / (2.0 * math.pi)
        frac = turns - math.floor(turns)
        return int(round(frac * 4294967296.0)) & 0xFFFFFFFF

    @staticmethod
    def from_pi_fraction(num, log2_den):
        assert log2_den <= 31
        return (num << (31 - log2_den)) % (1 << 32)


class Circuit:
    """Gate list with the reference's builders (src/circuit.rs:97-216)."""

    def __init__(self, num_qubits=0):
        self.num_qubits = num_qubits
        self.ops, self.q0, self.q1, self.ph = [], [], [], []

    def _push(self, op, a, b=0, ph=0):
        self._arr = None
        if op != GPHASE:
            self.num_qubits = max(self.num_qubits, a + 1)
        if op in (CZ, CX, SWAP):
            self.num_qubits = max(self.num_qubits, b + 1)
        self.ops.append(op); self.q0.append(a); self.q1.append(b); self.ph.append(ph & 0xFFFFFFFF)
        return self

    def h(self, q): return self._push(H, q)
    def x(self, q): return self._push(X, q)
    def y(self, q): return self._push(Y, q)
    def z(self, q): return self._push(Z, q)
    def s(self, q): return self._push(S, q)
    def sdg(self, q): return self._push(SDG, q)
    def t(self, q): return self._push(T, q)
    def tdg(self, q): return self._push(TDG, q)
    def cz(self, a, b): return self._push(CZ, a, b)
    def cx(self, c, t): return self._push(CX, c, t)
    def swap(self, a, b): return self._push(SWAP, a, b)
    def gphase(self, phase): return self._push(GPHASE, 0, 0, phase)

    def p(self, q, theta):
        """diag(1, e^{i theta}), theta in radians rounded to the 2^32 grid (pybind.rs:83-86)."""
        return self._push(P, q, 0, Phase.from_radians(theta))

    def p_phase(self, q, phase):
        """diag(1, e^{i phase}) with an exact ring phase."""
        return self._push(P, q, 0, phase)

    def rz(self, q, theta):
        """RZ = P(theta) plus global phase e^{-i theta/2} (src/circuit.rs:136-139)."""
        self._push(P, q, 0, Phase.from_radians(theta))
        return self._push(GPHASE, 0, 0, Phase.from_radians(-theta / 2.0))

    def ccx(self, a, b, c):
        """Toffoli with 7 T gates (src/circuit.rs:163-179)."""
        self.h(c).cx(b, c).tdg(c).cx(a, c).t(c).cx(b, c).tdg(c).cx(a, c).t(b).t(c).h(c).cx(a, b).t(a).tdg(b)
        return self.cx(a, b)

    def ccz(self, a, b, c):
        return self.h(c).ccx(a, b, c).h(c)

    def t_count(self):
        n = 0
        for op, ph in zip(self.ops, self.ph):
            if op in (T, TDG) or (op == P and (ph & ((1 << 30) - 1)) != 0):
                n += 1
        return n

    def __len__(self):
        return len(self.ops)

    _EXTEND_OPS = {"H": (H, 1), "X": (X, 1), "Y": (Y, 1), "Z": (Z, 1), "S": (S, 1), "SDG": (SDG, 1), "T": (T, 1), "TDG": (TDG, 1),
                   "CX": (CX, 2), "CZ": (CZ, 2), "SWAP": (SWAP, 2)}

    def extend(self, ops):
        """Append gates from a flat list of (name, qubits) pairs, names H, X,
        Y, Z, S, SDG, T, TDG, CX, CZ, SWAP (pybind.rs:101-121); a bad name or
        qubit count raises ValueError and appends nothing of that entry on."""
        for name, q in ops:
            q = [int(v) for v in q]
            spec = self._EXTEND_OPS.get(name)
            if spec is None or len(q) != spec[1] or any(v < 0 for v in q):
                raise ValueError("bad op %s %s" % (name, q))
            self._push(spec[0], *q)
        return self

    def arrays(self):
        """(ops, q0, q1, phase) as uint8/uint32 arrays; cached until the next gate."""
        if getattr(self, "_arr", None) is None:
            self._arr = (np.asarray(self.ops, dtype=np.uint8), np.asarray(self.q0, dtype=np.uint32),
                         np.asarray(self.q1, dtype=np.uint32), np.asarray(self.ph, dtype=np.uint32))
        return self._arr


class Simulator:
    """Simulator::with_options (src/sim.rs:144-148) with the GPU branch evaluator.

    strategy: "auto" | "cached"; pairs=p selects Strategy::Pairs(p)
    (pybind.rs:124-131).  form: "z8" (default: segment tables + outer Gray
    walk), "branch_walk" (per-branch Z8 walk), "plain_fp64" (FP64 port of the
    reference walk, one query per launch).
    """

    def __init__(self, circuit, strategy="auto", pairs=None, form="z8"):
        if pairs is not None:
            st, pr = 2, int(pairs)
        elif strategy in STRATEGIES:
            st, pr = STRATEGIES[strategy], 0
        else:
            raise ValueError("strategy must be 'auto', 'cached' or 'gpu_auto'")
        if form not in FORMS:
            raise ValueError("form must be one of %s" % sorted(FORMS))
        self.form = form
        self.circuit = circuit
        ops, q0, q1, ph = circuit.arrays()
        h = _sim_p()
        _check(lib.pq_sim_create(circuit.num_qubits, len(ops), _ptr(ops, _u8p), _ptr(q0, _u32p), _ptr(q1, _u32p), _ptr(ph, _u32p), st, pr, ctypes.byref(h)))
        self._h = h
        self._hv = h.value
        self.num_qubits = lib.pq_sim_num_qubits(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.pq_sim_destroy(h)
            self._h = None
            self._hv = 0

    @property
    def handle(self):
        return self._h

    def _form(self, form):
        return FORMS[form or self.form]

    def _qubit_bytes(self, v):
        """Basis index v (bit q = qubit q) as one byte per qubit."""
        v = _wide_index(v, self.num_qubits)
        return np.frombuffer(bytes(((v >> q) & 1) for q in range(self.num_qubits)), dtype=np.uint8)

    def amplitude(self, input, output, form=None):
        out = (ctypes.c_double * 2)()
        if self.num_qubits > 64:
            x, y = self._qubit_bytes(input), self._qubit_bytes(output)
            _check(lib.pq_sim_amplitude(self._h, _ptr(x, _u8p), _ptr(y, _u8p), self._form(form), out))
        else:
            _check(lib.pq_sim_amplitude_index(self._h, _index(input), _index(output), self._form(form), out))
        return (out[0], out[1])

    def probability(self, input, output, form=None):
        out = ctypes.c_double()
        if self.num_qubits > 64:
            x, y = self._qubit_bytes(input), self._qubit_bytes(output)
            _check(lib.pq_sim_probability_bytes(self._h, _ptr(x, _u8p), _ptr(y, _u8p), self._form(form), ctypes.byref(out)))
        else:
            _check(lib.pq_sim_probability(self._h, _index(input), _index(output), self._form(form), ctypes.byref(out)))
        return out.value

    def marginal(self, input, measured, form=None):
        """Marginal probability of [(qubit, outcome), ...] (src/sim.rs:225-231)."""
        qs = np.asarray([m[0] for m in measured], dtype=np.uint32)
        os_ = np.asarray([1 if m[1] else 0 for m in measured], dtype=np.uint8)
        out = ctypes.c_double()
        if self.num_qubits > 64:
            x = self._qubit_bytes(input)
            _check(lib.pq_sim_marginal_probability_bytes(self._h, _ptr(x, _u8p), _ptr(qs, _u32p), _ptr(os_, _u8p), len(qs), self._form(form),
                                                         ctypes.byref(out)))
        else:
            _check(lib.pq_sim_marginal_probability(self._h, _index(input), _ptr(qs, _u32p), _ptr(os_, _u8p), len(qs), self._form(form),
                                                   ctypes.byref(out)))
        return out.value

    def amplitudes(self, inputs, outputs, form=None):
        n = len(inputs)
        if len(outputs) != n:
            raise ValueError("inputs and outputs differ in length")
        if (_pqfast is not None and type(inputs) is np.ndarray and type(outputs) is np.ndarray and inputs.dtype == _U64
                and outputs.dtype == _U64 and inputs.ndim == 1 and outputs.ndim == 1
                and inputs.flags.c_contiguous and outputs.flags.c_contiguous):
            out = np.empty(n, dtype=np.complex128)
            _check(_pqfast.amplitudes(self._hv, inputs, outputs, out, FORMS[form or self.form]))
            return out
        buf, out = _complex_out(n)
        _check(lib.pq_sim_amplitudes(self._h, n, _u64_arg(inputs), _u64_arg(outputs), self._form(form), buf))
        return out

    def amplitude_estimate(self, input, output, samples, seed, form=None):
        """Unbiased Monte-Carlo estimate of <output|C|input>: (re, im,
        standard_error) (sim.rs:187-193, pybind.rs:149-153).  The exact value
        with error 0 when the branch subspace has at most `samples` points."""
        o = np.zeros(3, dtype=np.float64)
        _check(lib.pq_sim_amplitude_estimate(self._h, _index(input), _index(output), int(samples), int(seed), self._form(form), _ptr(o, _f64p)))
        return float(o[0]), float(o[1]), float(o[2])

    def statevector_array(self, input, form=None):
        """All 2^n amplitudes <y|C|input> as a complex128 array indexed by y
        (sim.rs:201-213)."""
        n = self.num_qubits
        out = np.empty(1 << n, dtype=np.complex128)
        _check(lib.pq_sim_statevector(self._h, _index(input), self._form(form), out.ctypes.data))
        return out

    def statevector(self, input, form=None):
        """The statevector as raw little-endian complex128 bytes (re, im per
        amplitude, indexed by y), as the reference's binding returns it
        (pybind.rs:158-168); np.frombuffer(b, np.complex128) reads it back."""
        a = self.statevector_array(input, form)
        return a.astype("<c16", copy=False).tobytes()

    def sample_many(self, input, shots, seed, form=None):
        """shots measurement outcomes of all qubits, as basis indices
        (sim.rs:272-284, pybind.rs:169-172): the chain rule over qubit
        prefixes with the reference's per-shot SplitMix64 streams."""
        out = np.empty(int(shots), dtype=np.uint64)
        _check(lib.pq_sim_sample_many(self._h, _index(input), int(shots), int(seed), self._form(form), out.ctypes.data))
        return [int(v) for v in out]

    def stats(self):
        """(free_vars, planes, lines, radical, terms, branch_dim, pairs) (pybind.rs:177-180)."""
        o = np.zeros(7, dtype=np.uint64)
        _check(lib.pq_sim_stats(self._h, _ptr(o, _u64p)))
        return tuple(int(v) for v in o)

    def marginal_stats(self, qubits):
        qs = np.asarray(list(qubits), dtype=np.uint32)
        o = np.zeros(7, dtype=np.uint64)
        _check(lib.pq_sim_marginal_stats(self._h, _ptr(qs, _u32p), len(qs), _ptr(o, _u64p)))
        return tuple(int(v) for v in o)

    def z8_info(self):
        o = np.zeros(10, dtype=np.uint32)
        _check(lib.pq_sim_z8_info(self._h, _ptr(o, _u32p)))
        return dict(zip(("d", "inner", "slots", "inner_pivots", "rank_k", "imag_slot0", "ops_per_branch", "pivot_free_mid",
                         "low_slots_pivot_free", "inner_slots"), (int(v) for v in o)))

    def seg_info(self):
        o = np.zeros(8, dtype=np.uint64)
        _check(lib.pq_sim_seg_info(self._h, _ptr(o, _u64p)))
        return dict(zip(("d", "inner", "inner_pivots", "outer", "lane_low", "walkable", "table_bytes", "mid_bits"), (int(v) for v in o)))

    def _seg_call(self, fn, x, y, mid_bits):
        total = 1 << self.seg_info()["outer"]
        k = np.zeros(total, np.uint8); hs = np.zeros(total, np.uint32); pio = np.zeros(total, np.uint32)
        n = ctypes.c_size_t()
        rc = _check(fn(self._h, int(x), int(y), int(mid_bits), total, _ptr(k, _u8p), _ptr(hs, _u32p), _ptr(pio, _u32p), ctypes.byref(n)))
        if rc == 1:
            return None
        return k[: n.value], hs[: n.value], pio[: n.value]

    def seg_dump(self, x, y, mid_bits=0):
        """Device segment walk state (K mod 8, sign mask, inner pivots) at every outer point."""
        return self._seg_call(lib.pq_sim_seg_dump, x, y, mid_bits)

    def seg_lane_ref(self, x, y, mid_bits=0):
        """Host restatement of seg_dump."""
        return self._seg_call(lib.pq_sim_seg_lane_ref, x, y, mid_bits)

    def prologue_selfcheck(self, inputs, outputs):
        """Host-only: mismatches between the direct and the assignment prologue."""
        xs = _u64_array(inputs)
        ys = _u64_array(outputs)
        bad = np.zeros(1, dtype=np.uint64)
        _check(lib.pq_sim_prologue_selfcheck(self._h, len(xs), _ptr(xs, _u64p), _ptr(ys, _u64p), _ptr(bad, _u64p)))
        return int(bad[0])

    def seg_selfcheck(self, x, y, trials=8, seed=1):
        """Host-only: max |table segment - direct inner sum| / inner magnitude mass."""
        err = ctypes.c_double()
        rc = _check(lib.pq_sim_seg_selfcheck(self._h, int(x), int(y), trials, seed, ctypes.byref(err)))
        return None if rc == 1 else float(err.value)

    def selfcheck(self, x, y, trials=256, seed=1):
        bad = ctypes.c_uint64()
        rc = _check(lib.pq_sim_selfcheck(self._h, int(x), int(y), trials, seed, ctypes.byref(bad)))
        return None if rc == 1 else int(bad.value)

    def branch_dump(self, x, y, s_list):
        s = np.ascontiguousarray(s_list, dtype=np.uint64)
        n = len(s)
        zero = np.zeros(n, np.uint8); k8 = np.zeros(n, np.uint8); kt = np.zeros(n, np.uint8)
        r = ctypes.c_uint32()
        rc = _check(lib.pq_sim_branch_dump(self._h, int(x), int(y), n, _ptr(s, _u64p), _ptr(zero, _u8p), _ptr(k8, _u8p), _ptr(kt, _u8p), ctypes.byref(r)))
        if rc == 1:
            return None
        return zero, k8, kt, int(r.value)

    def walk_dump(self, x, y, mid_bits=0):
        info = self.z8_info()
        total = 1 << info["slots"]
        k = np.zeros(total, np.uint8); mag = np.zeros(total, np.float32); s = np.zeros(total, np.uint64)
        n = ctypes.c_size_t()
        rc = _check(lib.pq_sim_walk_dump(self._h, int(x), int(y), mid_bits, total, _ptr(k, _u8p), _ptr(mag, _f32p), _ptr(s, _u64p), ctypes.byref(n)))
        if rc == 1:
            return None
        return k[: n.value], mag[: n.value], s[: n.value]


class _EngineDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("nf", "planes", "lines", "radical", "terms", "branch_dim", "rank", "consistency_rows",
                                            "num_h", "global_phase")] + \
               [("zero", ctypes.c_int32), ("sigma", ctypes.c_int32), ("free_vars", ctypes.c_uint32)] + \
               [("ucols", _u64p), ("mcols", _u64p), ("diag", _u64p), ("o3", _u64p), ("fixed_v", _u8p), ("fixed_theta", _u32p),
                ("b_fixed", _u64p), ("theta", _u32p), ("pivot_col", _u32p), ("pivot_row", _u64p), ("consistency", _u64p),
                ("kernel", _u64p), ("kernel_u", _u64p)]


_DESC_ARRAYS = {"ucols": (np.uint64, _u64p), "mcols": (np.uint64, _u64p), "diag": (np.uint64, _u64p), "o3": (np.uint64, _u64p),
                "fixed_v": (np.uint8, _u8p), "fixed_theta": (np.uint32, _u32p), "b_fixed": (np.uint64, _u64p),
                "theta": (np.uint32, _u32p), "pivot_col": (np.uint32, _u32p), "pivot_row": (np.uint64, _u64p),
                "consistency": (np.uint64, _u64p), "kernel": (np.uint64, _u64p), "kernel_u": (np.uint64, _u64p)}


class Engine:
    """One path-sum Engine (src/engine.rs:273-284) handed to the GPU by its
    fields (pq_engine_create): the engine-level seam below the Simulator.
    `fields` holds the sizes and arrays of pq_engine_desc (include/polyq_amd.h)
    under the same names; kernel_u may be None (derived from mcols)."""

    def __init__(self, fields):
        d = _EngineDesc()
        for name, _ in _EngineDesc._fields_[:13]:
            setattr(d, name, int(fields[name]))
        self._keep = []
        for name, (dt, pt) in _DESC_ARRAYS.items():
            a = fields.get(name)
            if a is None:
                setattr(d, name, None)
                continue
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            setattr(d, name, a.ctypes.data_as(pt))
        h = ctypes.c_void_p()
        _check(lib.pq_engine_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self.nf = int(fields["nf"])
        w = lambda n: (int(n) + 63) // 64
        self._uwords = 2 * w(fields["planes"]) + w(fields["lines"]) + w(fields["radical"])
        self._swords = w(fields["terms"])
        self._keep = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.pq_engine_destroy(h)
            self._h = None

    def stats(self):
        o = np.zeros(6, np.uint64)
        _check(lib.pq_engine_stats(self._h, _ptr(o, _u64p)))
        return tuple(int(v) for v in o)

    def eval(self, values, form="z8"):
        """Engine::eval (src/engine.rs:551-561): values aligned with fixed()."""
        v = np.ascontiguousarray(values, dtype=np.uint8)
        out = (ctypes.c_double * 2)()
        _check(lib.pq_engine_eval(self._h, _ptr(v, _u8p), len(v), FORMS[form], out))
        return complex(out[0], out[1])

    def sum_subspace(self, u_start, s_start, c_start, form="z8"):
        """Engine::sum_subspace (src/engine.rs:669-693): the sum of
        weight(s) ROOT8[phase8(u(s))] over s_start + span(kernel), from the
        start shift u_start (one packed shift row) and branch vector s_start
        (packed terms bits, the reference's term order) with c_start its
        weight."""
        u = np.ascontiguousarray(u_start, dtype=np.uint64)
        st = np.ascontiguousarray(s_start, dtype=np.uint64)
        if u.size < self._uwords or st.size < self._swords:
            raise ValueError("u_start needs %d words and s_start %d" % (self._uwords, self._swords))
        c = np.array([complex(c_start).real, complex(c_start).imag], np.float64)
        out = (ctypes.c_double * 2)()
        _check(lib.pq_engine_sum_subspace(self._h, _ptr(u, _u64p), _ptr(st, _u64p), _ptr(c, _f64p), FORMS[form], out))
        return complex(out[0], out[1])

    def branch_dump(self, values, s_list):
        v = np.ascontiguousarray(values, dtype=np.uint8)
        s = np.ascontiguousarray(s_list, dtype=np.uint64)
        n = len(s)
        zero = np.zeros(n, np.uint8); k8 = np.zeros(n, np.uint8); kt = np.zeros(n, np.uint8); r = ctypes.c_uint32()
        _check(lib.pq_engine_branch_dump(self._h, _ptr(v, _u8p), n, _ptr(s, _u64p), _ptr(zero, _u8p), _ptr(k8, _u8p), _ptr(kt, _u8p), ctypes.byref(r)))
        return zero, k8, kt, r.value


def engine_eval_batch(engines, values, form="z8"):
    """Engine::eval of many (engine, values) queries in one launch
    (pq_engine_eval_batch); summing the branch engines of a RankEngine for the
    same values gives RankEngine::eval_assignment (src/rank.rs:184-190)."""
    n = len(engines)
    vs = [np.ascontiguousarray(v, dtype=np.uint8) for v in values]
    hs = (ctypes.c_void_p * n)(*[e._h.value for e in engines])
    ps = (ctypes.c_void_p * n)(*[v.ctypes.data for v in vs])
    buf, out = _complex_out(n)
    _check(lib.pq_engine_eval_batch(n, hs, ps, FORMS[form], buf))
    return out.copy()


def batch_amplitudes(sims, inputs, outputs, form="z8"):
    """One persistent launch over queries of many simulators."""
    n = len(sims)
    if len(inputs) != n or len(outputs) != n:
        raise ValueError("sims, inputs and outputs differ in length")
    hs = _handle_array(sims, n)
    buf, out = _complex_out(n)
    _check(lib.pq_batch_amplitudes(n, hs.ctypes.data, _u64_arg(inputs), _u64_arg(outputs), FORMS[form], buf))
    return out


class Graph:
    """A batch captured into a hipGraph and replayed."""

    def __init__(self, sims, inputs, outputs, form="z8"):
        n = len(sims)
        self._sims = list(sims)
        hs = _handle_array(self._sims, n)
        self._xs = _u64_array(inputs)
        self._ys = _u64_array(outputs)
        g = ctypes.c_void_p()
        _check(lib.pq_graph_create(n, hs.ctypes.data_as(ctypes.POINTER(_sim_p)), _ptr(self._xs, _u64p), _ptr(self._ys, _u64p), FORMS[form], ctypes.byref(g)))
        self._g = g
        self.n = n

    def run(self):
        """Replay: (amplitudes as complex128, device milliseconds of the replay)."""
        out = np.empty(self.n, dtype=np.complex128)       # re, im interleaved: the 2n doubles the call writes
        ms = ctypes.c_double()
        _check(lib.pq_graph_run(self._g, ctypes.cast(out.ctypes.data, _f64p), ctypes.byref(ms)))
        return out, ms.value

    def __del__(self):
        g = getattr(self, "_g", None)
        if g:
            lib.pq_graph_destroy(g)
            self._g = None


def set_walk_tuning(offset, max_mid, form="z8"):
    _check(lib.pq_set_walk_tuning(FORMS[form], int(offset), int(max_mid)))


def get_walk_tuning(form="z8"):
    o = np.zeros(2, dtype=np.uint32)
    _check(lib.pq_get_walk_tuning(FORMS[form], _ptr(o, _u32p)))
    return int(o[0]), int(o[1])


def set_walk_shape(waves, form="z8"):
    """Workgroup shape of a batched form: waves of 64 lanes per workgroup
    (the segment walk takes 1, 2 or 4; the branch walk only 4)."""
    _check(lib.pq_set_walk_shape(FORMS[form], int(waves)))


def get_walk_shape(form="z8"):
    o = np.zeros(1, dtype=np.uint32)
    _check(lib.pq_get_walk_shape(FORMS[form], _ptr(o, _u32p)))
    return int(o[0])


def set_segment_bits(bits):
    """Cap on the table bits I + npin of the segment tables of simulators
    created afterwards (default 16); an engine of branch dimension d gets
    min(cap, max(14, d - 14))."""
    _check(lib.pq_set_segment_bits(int(bits)))


def get_segment_bits():
    return int(lib.pq_get_segment_bits())


_TUNE_SHAPES = {"z8": (1, 2, 4), "branch_walk": (4,)}
_TUNE_GRID = {"z8": ((8, 10, 11, 12, 13, 14, 16), (0, 2, 4, 6, 8)),
              "branch_walk": ((23, 24, 25, 26, 27, 28, 29), (4, 6, 8))}


def branch_walk_mid_bits(d, offset, max_mid, inner_slots):
    """Mid coordinates a lane walks in the per-branch walk for an engine of
    branch dimension d at lane walk tuning (offset, max_mid), with
    inner_slots inner coordinates per outer point (z8_info()["inner_slots"]):
    2^(inner_slots + M) branches per lane (the library's own rule,
    z8_auto_mid_bits)."""
    if d <= offset:
        return 0
    return min(d - offset, max_mid, d - inner_slots if d > inner_slots else 0)


def autotune(sims, inputs, outputs, form="z8", offsets=None, max_mids=None, shapes=None, repeats=3, cache=True, objective="wall",
             rounds=3):
    """Pick the lane walk split (offset, max_mid) and the workgroup shape
    (waves per workgroup, `shapes`) of a batched form that minimise the time
    of this batch: its wall time per call (objective "wall", the default:
    host prologue, launch and wait included) or its walk kernel time from HIP
    events ("kernel").  Splits that give every engine of the batch the same
    walk length are one configuration, timed once per shape; the (walk,
    shape) trials are timed in `rounds` interleaved rounds of `repeats` calls
    (after one untimed call each), so slow drifts of the host do not favour
    one of them, and each is scored by its median.  Applies the best and, if
    cache, stores it per device and form in ~/.cache/polyq_amd/walk_tuning.json.
    Returns (offset, max_mid, {"offset/max_mid/waves": ms}); the chosen shape
    is get_walk_shape(form)."""
    import json
    import time
    offsets = offsets or _TUNE_GRID[form][0]
    max_mids = max_mids or _TUNE_GRID[form][1]
    shapes = shapes or _TUNE_SHAPES[form]
    xs = _u64_array(inputs)
    ys = _u64_array(outputs)
    distinct = list({id(s): s for s in sims}.values())
    configs = {}                      # walk lengths per engine -> [(offset, max_mid), ...]
    for off in offsets:
        for mx in max_mids:
            set_walk_tuning(off, mx, form)
            if form == "z8":
                key = tuple(s.seg_info()["mid_bits"] for s in distinct)
            elif all(s.stats()[6] == 0 for s in distinct):      # one branch engine per simulator
                key = tuple(branch_walk_mid_bits(s.z8_info()["d"], off, mx, s.z8_info()["inner_slots"]) for s in distinct)
            else:
                key = (off, mx)
            configs.setdefault(key, []).append((off, mx))
    trials = [(key, w) for key in configs for w in shapes]
    samples = {t: [] for t in trials}
    prev = get_timing()
    set_timing(objective == "kernel")
    try:
        for _ in range(rounds):
            for key, w in trials:
                set_walk_tuning(configs[key][0][0], configs[key][0][1], form)
                set_walk_shape(w, form)
                batch_amplitudes(sims, xs, ys, form=form)     # untimed
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    batch_amplitudes(sims, xs, ys, form=form)
                    wall = 1e3 * (time.perf_counter() - t0)
                    samples[(key, w)].append(last_timing()["kernel_ms"] if objective == "kernel" else wall)
    finally:
        set_timing(prev)
    table = {}
    best = None
    for key, w in trials:
        t = float(np.median(samples[(key, w)]))
        for off, mx in configs[key]:
            table["%d/%d/%d" % (off, mx, w)] = t
        if best is None or t < best[0]:
            best = (t, configs[key][0][0], configs[key][0][1], w)
    set_walk_tuning(best[1], best[2], form)
    set_walk_shape(best[3], form)
    if cache:
        try:
            path = os.path.join(os.path.expanduser("~"), ".cache", "polyq_amd", "walk_tuning.json")
            os.makedirs(os.path.dirname(path), exist_ok=True)
            data = {}
            if os.path.exists(path):
                with open(path) as f:
                    data = json.load(f)
            data["%s|%s" % (device_info(), form)] = {"offset": best[1], "max_mid": best[2], "waves": best[3], "table_ms": table}
            with open(path, "w") as f:
                json.dump(data, f)
        except OSError:
            pass
    return best[1], best[2], table


def load_tuning(form="z8"):
    """Apply a cached autotuner result for this device and form, if any."""
    import json
    path = os.path.join(os.path.expanduser("~"), ".cache", "polyq_amd", "walk_tuning.json")
    try:
        with open(path) as f:
            rec = json.load(f).get("%s|%s" % (device_info(), form))
    except (OSError, ValueError):
        return None
    if rec:
        set_walk_tuning(rec["offset"], rec["max_mid"], form)
        set_walk_shape(rec.get("waves", 4), form)
        return rec["offset"], rec["max_mid"]
    return None


def set_timing(on):
    """Record HIP timing events around eager batches (off by default: each
    event is a queue barrier that adds microseconds to a short batch)."""
    lib.pq_set_timing(1 if on else 0)


def get_timing():
    return bool(lib.pq_get_timing())


class timed:
    """Context manager: timing events on inside, previous setting after."""

    def __enter__(self):
        self._prev = get_timing()
        set_timing(True)
        return self

    def __exit__(self, *exc):
        set_timing(self._prev)
        return False


def last_timing():
    """kernel_ms and total_ms (HIP events; NaN unless timing is on), branches,
    items and grid of the last evaluation on this thread."""
    o = np.zeros(5, dtype=np.float64)
    lib.pq_last_timing(_ptr(o, _f64p))
    return dict(zip(("kernel_ms", "total_ms", "branches", "items", "grid"), o.tolist()))


def device_info():
    buf = ctypes.create_string_buffer(256)
    _check(lib.pq_device_info(buf, 256))
    return buf.value.decode()


def best_pairs(stats, available_pairs, t_class_pairs, gpu_model=False):
    """Pairs Strategy::Auto decomposes for an engine of the given stats
    (free_vars, planes, lines, radical, terms, branch_dim): the cost model of
    src/rank.rs:205-222, or its GPU refit."""
    st = np.asarray(list(stats)[:6], dtype=np.uint64)
    return _check(lib.pq_best_pairs(_ptr(st, _u64p), int(available_pairs), int(t_class_pairs), 1 if gpu_model else 0))


def valu_probe(blocks_per_cu=8, iters=16384):
    """VALU issue rates of the device (pq_valu_probe): per instruction,
    lane-ops/s, lanes per clock per CU at the in-kernel clock, that clock (GHz)
    and the kernel time."""
    n = 3 + 4 * 16
    o = np.zeros(n, dtype=np.float64)
    k = _check(lib.pq_valu_probe(_ptr(o, _f64p), n, blocks_per_cu, iters))
    ops = {}
    for i in range(k):
        r = o[3 + 4 * i: 7 + 4 * i]
        ops[lib.pq_valu_probe_name(i).decode()] = {"lane_ops_per_s": float(r[0]), "lanes_per_clk_per_cu": float(r[1]),
                                                   "clock_ghz": float(r[2]), "ms": float(r[3])}
    return {"cus": int(o[0]), "wall_clock_mhz": float(o[1]), "blocks_per_cu": int(o[2]), "ops": ops}


def rng_u64(seed):
    """First SplitMix64 output of Rng::new(seed) (src/rng.rs:13-23, pybind.rs:221-224)."""
    from .workloads import Rng
    return Rng(seed).next_u64()


lib.pq_server_create.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
lib.pq_server_destroy.argtypes = [ctypes.c_void_p]
lib.pq_server_destroy.restype = None
lib.pq_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_size_t, _u8p, _u32p, _u32p, _u32p, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_uint64, ctypes.c_uint64, _u32p, _u8p, ctypes.c_size_t, _u64p]
lib.pq_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, _f64p, _f64p]
lib.pq_server_stats.argtypes = [ctypes.c_void_p, _f64p]
lib.pq_server_set_host_budget.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
lib.pq_server_host_stats.argtypes = [ctypes.c_void_p, _f64p]
lib.pq_server_phase_stats.argtypes = [ctypes.c_void_p, _f64p]

KINDS = {"amplitude": 0, "probability": 1, "marginal": 2}


class Server:
    """Request loop with resident engines, HBM admission and continuous batching."""

    def __init__(self, hbm_budget_bytes=64 << 30, workers=8, form="z8", strategy="cached", host_budget_bytes=None):
        h = ctypes.c_void_p()
        _check(lib.pq_server_create(int(hbm_budget_bytes), workers, FORMS[form], ctypes.byref(h)))
        self._h = h
        self.strategy = STRATEGIES[strategy]
        if host_budget_bytes is not None:
            self.set_host_budget(host_budget_bytes)

    def set_host_budget(self, nbytes):
        """Host bytes of built engine images kept for idle circuits (default 8 GiB)."""
        _check(lib.pq_server_set_host_budget(self._h, int(nbytes)))

    def host_stats(self):
        o = np.zeros(3, dtype=np.float64)
        _check(lib.pq_server_host_stats(self._h, _ptr(o, _f64p)))
        return dict(zip(("host_bytes", "host_budget", "host_drops"), o.tolist()))

    PHASES = ("queue", "build", "upload", "prologue", "launch_wait", "device")

    def phase_stats(self):
        """Mean ms per request phase over the last 65536 completed requests
        ("all") and over the slowest 1% of them ("slowest_1pct")."""
        o = np.zeros(13, dtype=np.float64)
        _check(lib.pq_server_phase_stats(self._h, _ptr(o, _f64p)))
        return {"requests": int(o[0]), "all": dict(zip(self.PHASES, o[1:7].tolist())),
                "slowest_1pct": dict(zip(self.PHASES, o[7:13].tolist()))}

    def submit(self, circuit, kind, x=0, y=0, measured=()):
        ops, q0, q1, ph = circuit.arrays()
        mq = np.asarray([m[0] for m in measured], dtype=np.uint32)
        mo = np.asarray([1 if m[1] else 0 for m in measured], dtype=np.uint8)
        t = ctypes.c_uint64()
        _check(lib.pq_server_submit(self._h, circuit.num_qubits, len(ops), _ptr(ops, _u8p), _ptr(q0, _u32p), _ptr(q1, _u32p), _ptr(ph, _u32p),
                                    self.strategy, KINDS[kind], _index(x), _index(y), _ptr(mq, _u32p), _ptr(mo, _u8p), len(mq), ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """(value, latency_ms); value is complex for amplitudes, float otherwise."""
        out = (ctypes.c_double * 2)()
        lat = ctypes.c_double()
        rc = lib.pq_server_wait(self._h, ticket, timeout, out, ctypes.byref(lat))
        if rc < 0:
            raise PolyqError(lib.pq_last_error().decode())
        if rc == 1:
            raise TimeoutError("request %d not finished" % ticket)
        return complex(out[0], out[1]), lat.value

    def stats(self):
        o = np.zeros(10, dtype=np.float64)
        _check(lib.pq_server_stats(self._h, _ptr(o, _f64p)))
        keys = ("completed", "failed", "batches", "evictions", "resident_bytes", "builds", "max_batch", "overlapped_batches", "kernel_ms", "resident_circuits")
        return dict(zip(keys, o.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            lib.pq_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
