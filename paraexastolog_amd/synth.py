"""Seeded synthetic random Boolean networks and their `.bnet` text.

Generator version 1.  Random stream: SplitMix64 seeds xoshiro256**; a uniform
double is (x >> 11) * 2^-53 and an integer below m is floor(u * m).  Draw order:

1. for node i = 0..N-1: its arity K_i (fixed K, or 2/3 with probability 1/2
   each when K = 0), then K_i distinct inputs from {0..N-1} minus {i} by a
   partial Fisher-Yates shuffle, then its 2^K_i truth-table bits ~ Bernoulli(1/2)
   (table index bit k = value of the k-th drawn input);
2. the fixed nodes of the initial state (partial Fisher-Yates) and their values;
3. the rate sets: per node (up, down) = exp(U(ln 0.1, ln 10)).

`.bnet` text: header `targets, factors`, one line `v<i>, <formula>` per node
with the formula an OR of minterms such as `(v3 & !v7 & v9)`; a constant-false
rule is written `(vA & !vA)` and a constant-true one `(vA | !vA)` with vA one
of its inputs; exactly one trailing newline, no blank lines (both reference
readers accept this form: cuda/src/model.cpp:80-125, python/exastolog.py:17-26).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GENERATOR_VERSION = 1
_M64 = (1 << 64) - 1


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & _M64


class Xoshiro256ss:
    def __init__(self, seed: int):
        s = seed & _M64
        st = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & _M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            st.append(z ^ (z >> 31))
        self.s = st

    def next(self) -> int:
        s = self.s
        result = (_rotl((s[1] * 5) & _M64, 7) * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def below(self, m: int) -> int:
        return min(int(self.uniform() * m), m - 1)


@dataclass
class Network:
    N: int
    inputs: list = field(default_factory=list)  # per node: drawn input order
    tables: list = field(default_factory=list)  # per node: list of 0/1, length 2^K
    fixed: list = field(default_factory=list)  # node indices
    fixed_vals: list = field(default_factory=list)
    seed: int = 0
    cycles: list = field(default_factory=list)  # spliced (a, b, c) triples

    @property
    def nodes(self):
        return [f"v{i}" for i in range(self.N)]

    def rule_text(self, i: int) -> str:
        ins = self.inputs[i]
        tab = self.tables[i]
        K = len(ins)
        terms = []
        for idx in range(1 << K):
            if tab[idx]:
                lits = [(f"v{ins[k]}" if (idx >> k) & 1 else f"!v{ins[k]}") for k in range(K)]
                terms.append("(" + " & ".join(lits) + ")")
        a = f"v{ins[0]}"
        if not terms:
            return f"({a} & !{a})"
        if len(terms) == 1 << K:
            return f"({a} | !{a})"
        return " | ".join(terms)

    def bnet(self) -> str:
        lines = ["targets, factors"] + [f"v{i}, {self.rule_text(i)}" for i in range(self.N)]
        return "\n".join(lines) + "\n"

    def fixed_spec(self):
        return [f"v{i}" for i in self.fixed], [bool(v) for v in self.fixed_vals]


def generate(N: int, K: int, seed: int, n_fixed: int = 0, rng: Xoshiro256ss | None = None) -> tuple:
    """Returns (network, rng) with the rng positioned after the fixed-node draws,
    ready for rate_sets()."""
    if N < 2:
        raise ValueError("N must be >= 2")
    rng = rng or Xoshiro256ss(seed)
    net = Network(N=N, seed=seed)
    for i in range(N):
        k = K if K > 0 else (2 if rng.uniform() < 0.5 else 3)
        k = min(k, N - 1)
        pool = [j for j in range(N) if j != i]
        for a in range(k):
            b = a + rng.below(len(pool) - a)
            pool[a], pool[b] = pool[b], pool[a]
        net.inputs.append(pool[:k])
        net.tables.append([1 if rng.uniform() < 0.5 else 0 for _ in range(1 << k)])
    pool = list(range(N))
    for a in range(n_fixed):
        b = a + rng.below(N - a)
        pool[a], pool[b] = pool[b], pool[a]
    net.fixed = pool[:n_fixed]
    net.fixed_vals = [1 if rng.uniform() < 0.5 else 0 for _ in range(n_fixed)]
    return net, rng


def rate_sets(rng: Xoshiro256ss, N: int, count: int, lo: float = 0.1, hi: float = 10.0) -> np.ndarray:
    """[count][2N] log-uniform rates; [2i] = up, [2i+1] = down."""
    la, lb = np.log(lo), np.log(hi)
    out = np.empty((count, 2 * N))
    for c in range(count):
        for j in range(2 * N):
            out[c, j] = np.exp(la + (lb - la) * rng.uniform())
    return out


def splice_cycles(net: Network, rng: Xoshiro256ss, count: int) -> list:
    """Splices `count` negative-feedback 3-cycles a -> b -> c -| a into the rules:
    3*count distinct nodes are drawn by a partial Fisher-Yates shuffle (after the
    fixed-node draws), taken three at a time as (a, b, c), and their rules are
    replaced by f_b = a, f_c = b, f_a = !c.  Each cycle alone has no fixed point
    and a 6-state cyclic attractor on its three bits; nodes that read the cycle
    nodes spread it, which is what makes the terminal SCCs large.  Returns the
    triples."""
    N = net.N
    pool = list(range(N))
    for a in range(3 * count):
        b = a + rng.below(N - a)
        pool[a], pool[b] = pool[b], pool[a]
    triples = [tuple(pool[3 * t:3 * t + 3]) for t in range(count)]
    for a, b, c in triples:
        net.inputs[b], net.tables[b] = [a], [0, 1]
        net.inputs[c], net.tables[c] = [b], [0, 1]
        net.inputs[a], net.tables[a] = [c], [1, 0]
    return triples


# The workloads the benchmarks and tests use.
def config(name: str, seed: int = 0):
    """(network, rng) for a named workload:
    c0  N=10, K=2, nodes v0..v2 fixed to 0 (128 initial states), rates 1.0
    c1  N=20, K in {2,3}, 5 seeded fixed nodes, log-uniform rates
    c2  N=16, K=2, single topology for rate-set batches
    c3  N=22, K=3, 5 seeded fixed nodes, then 3 spliced negative-feedback
        3-cycles (splice_cycles), log-uniform rates
    """
    if name == "c0":
        net, rng = generate(10, 2, seed, 0)
        net.fixed, net.fixed_vals = [0, 1, 2], [0, 0, 0]
        return net, rng
    if name == "c1":
        return generate(20, 0, seed, 5)
    if name == "c2":
        return generate(16, 2, seed, 0)
    if name == "c3":
        net, rng = generate(22, 3, seed, 5)
        net.cycles = splice_cycles(net, rng, 3)
        return net, rng
    raise KeyError(name)


UNEVEN_SIZES = (12, 13, 14, 15, 16, 17, 18, 18)


def uneven(seed: int = 0):
    """Eight distinct networks of N = 12, 13, ..., 18, 18 (K = 2, no fixed
    nodes): network k is generate(N_k, 2, 1000 * seed + k).  Returns a list of
    (network, rng)."""
    return [generate(n, 2, 1000 * seed + k, 0) for k, n in enumerate(UNEVEN_SIZES)]
