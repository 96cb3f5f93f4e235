"""Seeded workload generators (SplitMix64, src/rng.rs:7-102) and the benchmark
configurations of the branch evaluator.

All generators draw from one SplitMix64 stream in a fixed, documented order,
so a (config, seed) pair names a unique circuit set.
"""
import math

from . import Circuit, Phase

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


class Rng:
    """SplitMix64 exactly as src/rng.rs:9-32."""

    def __init__(self, seed):
        self.s = (seed + GAMMA) & M64

    def next_u64(self):
        self.s = (self.s + GAMMA) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def next_f64(self):
        return (self.next_u64() >> 11) / float(1 << 53)

    def below(self, n):
        return self.next_u64() % n

    def pair(self, n):
        a = self.below(n)
        b = self.below(n)
        while b == a:
            b = self.below(n)
        return a, b


def random_clifford_t(n, depth, t_count, generic, rng):
    """The reference's test generator (src/rng.rs:36-74), same draw order."""
    c = Circuit(n)
    t_left = t_count
    total = max(depth, t_count)
    for step in range(total):
        if t_left > 0 and (total - step <= t_left or rng.next_f64() < 0.3):
            q = rng.below(n)
            k = rng.below(4 if generic else 2)
            if k == 0:
                c.t(q)
            elif k == 1:
                c.tdg(q)
            elif k == 2:
                c.p_phase(q, Phase.from_pi_fraction(1 + 2 * rng.below(7), 4))
            else:
                c.p_phase(q, Phase.from_pi_fraction(1 + 2 * rng.below(3), 3))
            t_left -= 1
            continue
        q = rng.below(n)
        r = rng.below(9 if n > 1 else 5)
        if r in (0, 1):
            c.h(q)
        elif r == 2:
            c.s(q)
        elif r == 3:
            c.z(q)
        elif r == 4:
            c.x(q)
        else:
            q2 = rng.below(n)
            while q2 == q:
                q2 = rng.below(n)
            if r in (5, 6):
                c.cz(q, q2)
            elif r == 7:
                c.cx(q, q2)
            else:
                c.swap(q, q2)
    return c


def random_layered(n, rounds, rng):
    """src/rng.rs:79-102."""
    c = Circuit(n)
    for _ in range(rounds):
        for q in range(n):
            c.h(q)
        for q in range(n):
            k = rng.below(3)
            if k == 0:
                c.t(q)
            elif k == 1:
                c.tdg(q)
            else:
                c.s(q)
        if n > 1:
            for _ in range(n):
                a, b = rng.below(n), rng.below(n)
                if a != b:
                    c.cz(a, b)
    return c


def uniform_gates(n, n_gates, rng):
    """Workload 'plain CPU reference': gates i.i.d. uniform over {H,S,CX,CZ,T} on
    uniform qubits (distinct pair for CX/CZ).  Draws per gate: kind, then qubit(s)."""
    c = Circuit(n)
    for _ in range(n_gates):
        k = rng.below(5)
        if k == 0:
            c.h(rng.below(n))
        elif k == 1:
            c.s(rng.below(n))
        elif k == 2:
            a, b = rng.pair(n)
            c.cx(a, b)
        elif k == 3:
            a, b = rng.pair(n)
            c.cz(a, b)
        else:
            c.t(rng.below(n))
    return c


def sandwich_layer(n, n_nonclifford, n_twoq, n_s, rng, angle=None):
    """H on every qubit; a seeded random interleaving of n_nonclifford T (or
    Rz(theta) when angle='uniform'), n_twoq two-qubit gates (CX or CZ with
    probability 1/2 on a uniform distinct pair) and n_s S gates on uniform
    qubits; then H on every qubit.

    Draw order: Fisher-Yates shuffle of the token list (T..., 2q..., S...)
    with below(i+1) for i = len-1 .. 1, then per token in order: T -> qubit
    [, theta = 2*pi*next_f64()], 2q -> below(2), pair, S -> qubit."""
    toks = ["t"] * n_nonclifford + ["2"] * n_twoq + ["s"] * n_s
    for i in range(len(toks) - 1, 0, -1):
        j = rng.below(i + 1)
        toks[i], toks[j] = toks[j], toks[i]
    c = Circuit(n)
    for q in range(n):
        c.h(q)
    for tk in toks:
        if tk == "t":
            q = rng.below(n)
            if angle == "uniform":
                c.rz(q, 2.0 * math.pi * rng.next_f64())
            else:
                c.t(q)
        elif tk == "2":
            kind = rng.below(2)
            a, b = rng.pair(n)
            if kind == 0:
                c.cx(a, b)
            else:
                c.cz(a, b)
        else:
            c.s(rng.below(n))
    for q in range(n):
        c.h(q)
    return c


def config1(seed=1):
    """n=12, 150 uniform {H,S,CX,CZ,T} gates; amplitudes <0|C|y> for 32 random
    inputs y.  Returns (circuit, inputs, outputs)."""
    rng = Rng(seed)
    c = uniform_gates(12, 150, rng)
    ys = [rng.below(1 << 12) for _ in range(32)]
    return c, ys, [0] * 32


def config2(seed=1):
    """n=40: H^40; 36 T + 300 CX/CZ + 100 S interleaved; H^40; amplitudes
    <0|C|y> for 16 random inputs y.  Returns (circuit, inputs, outputs)."""
    rng = Rng(seed)
    c = sandwich_layer(40, 36, 300, 100, rng)
    ys = [rng.below(1 << 40) for _ in range(16)]
    return c, ys, [0] * 16


def config3(seed=1, n_amps=1024):
    """n=48, the config-2 generator (300 CX/CZ and 100 S interleaved) with 40
    Rz(theta), theta ~ U[0, 2pi), in place of the T gates; amplitudes <0|C|y>
    for random inputs y.  Returns (circuit, inputs, outputs)."""
    rng = Rng(seed)
    c = sandwich_layer(48, 40, 300, 100, rng, angle="uniform")
    ys = [rng.below(1 << 48) for _ in range(n_amps)]
    return c, ys, [0] * n_amps


def config4_circuit(rng):
    """One batched-workload circuit: n ~ U[20,56], T ~ U[8,32], 7.5 n two-qubit
    gates and 2.5 n S gates (the config-2 densities per qubit)."""
    n = 20 + rng.below(37)
    t = 8 + rng.below(25)
    return sandwich_layer(n, t, (15 * n) // 2, (5 * n) // 2, rng)


def config4(seed=1, n_circuits=256, per_circuit=16):
    """256 circuits x 16 random (x, y) pairs."""
    rng = Rng(seed)
    out = []
    for _ in range(n_circuits):
        c = config4_circuit(rng)
        pairs = [(rng.below(1 << c.num_qubits), rng.below(1 << c.num_qubits)) for _ in range(per_circuit)]
        out.append((c, pairs))
    return out
