"""Seeded synthetic workloads for the Pauli-propagation path.

The reference has no circuit or layer object: a layer is whatever generator
slice the caller hands to ``evolve!`` (spv_evolve.jl:312-387).  This module
pins the generator order and angles of each workload so that every consumer
(GPU path, CPU oracle, fixtures, bench) sees identical inputs.

Angles come from this module's own PRNG: SplitMix64 seeding xoshiro256**
(Blackman & Vigna 2018).  A uniform double in [0, 1) is ``(next() >> 11) * 2^-53``.

Gate convention (spv_kernels.jl:131-141): ``evolve!(O, G, θ)`` maps
O -> exp(iθ/2 G) O exp(-iθ/2 G), so a gate exp(-iθ/2 G) is generator G with
angle θ; generators apply left to right (evolve.jl:103-109).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

from .basis import Pauli, PauliBasis
from .truncation import CoeffTruncation, CompositeTruncation, TruncationStrategy, WeightTruncation

M64 = (1 << 64) - 1


class Xoshiro256ss:
    """xoshiro256** seeded through SplitMix64."""

    def __init__(self, seed: int):
        s = seed & M64
        st = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            st.append(z ^ (z >> 31))
        self.s = st

    @staticmethod
    def _rotl(x: int, k: int) -> int:
        return ((x << k) | (x >> (64 - k))) & M64

    def next(self) -> int:
        s = self.s
        result = (self._rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (inclusive) by rejection."""
        span = hi - lo + 1
        limit = (1 << 64) - ((1 << 64) % span)
        while True:
            r = self.next()
            if r < limit:
                return lo + r % span


@dataclass
class Layer:
    generators: List[PauliBasis]
    angles: List[float]


@dataclass
class Workload:
    name: str
    N: int
    observable: List[Tuple[PauliBasis, float]]
    layers: List[Layer]
    truncation: TruncationStrategy
    channel: Optional[Tuple[str, float]] = None  # applied after every layer
    window: int = 1
    notes: str = ""
    meta: dict = field(default_factory=dict)

    def flat(self) -> Tuple[List[PauliBasis], List[float]]:
        gens, angs = [], []
        for L in self.layers:
            gens.extend(L.generators)
            angs.extend(L.angles)
        return gens, angs

    def n_rotations(self) -> int:
        return sum(len(L.generators) for L in self.layers)


def chain_bonds(N: int) -> List[Tuple[int, int]]:
    """Open 1D chain bonds (j, j+1), j = 1..N-1 ascending."""
    return [(j, j + 1) for j in range(1, N)]


def lattice_bonds(rows: int, cols: int) -> List[Tuple[int, int]]:
    """Open square-lattice nearest-neighbour bonds, sites row-major q = cols*r + c + 1,
    ordered lexicographically by (min site, max site)."""
    bonds = []
    for r in range(rows):
        for c in range(cols):
            q = cols * r + c + 1
            if c + 1 < cols:
                bonds.append((q, q + 1))
            if r + 1 < rows:
                bonds.append((q, q + cols))
    bonds.sort()
    return bonds


def kicked_ising_16() -> Workload:
    """16-qubit open chain, 8 layers of ZZ(0.3) on the 15 bonds then X(0.7) on
    the 16 sites; observable Z_8; CoeffTruncation(1e-5); no noise."""
    N = 16
    layers = []
    for _ in range(8):
        g = [Pauli(N, Z=[a, b]) for a, b in chain_bonds(N)] + [Pauli(N, X=[j]) for j in range(1, N + 1)]
        a = [0.3] * (N - 1) + [0.7] * N
        layers.append(Layer(g, a))
    return Workload("kicked_ising_chain16", N, [(Pauli(N, Z=[8]), 1.0)], layers,
                    CoeffTruncation(1e-5))


def chain64(n_layers: int = 30, seed: int = 1234) -> Workload:
    """64-qubit open chain; each layer RZZ on the 63 bonds, RX then RZ on all
    sites, angles i.i.d. U[0, π) in that draw order; depolarizing p=1e-3 after
    each layer; |c| <= 1e-6 and weight <= 16 truncation; observable Z_32."""
    N = 64
    rng = Xoshiro256ss(seed)
    layers = []
    for _ in range(n_layers):
        g = ([Pauli(N, Z=[a, b]) for a, b in chain_bonds(N)] + [Pauli(N, X=[j]) for j in range(1, N + 1)]
             + [Pauli(N, Z=[j]) for j in range(1, N + 1)])
        a = [rng.uniform() * math.pi for _ in g]
        layers.append(Layer(g, a))
    return Workload("chain64_depol", N, [(Pauli(N, Z=[32]), 1.0)], layers,
                    CompositeTruncation(CoeffTruncation(1e-6), WeightTruncation(16)),
                    channel=("depolarizing", 1e-3), meta={"seed": seed})


def theta_h_values() -> List[float]:
    return [k * (math.pi / 2) / 15 for k in range(16)]


def lattice121(theta_h: float, n_layers: int = 20) -> Workload:
    """121-qubit 11x11 open lattice; each layer ZZ(θ_J = -π/2) on the 220 bonds
    then X(θ_h) on all sites; dephasing p=5e-4 after each layer; |c| <= 1e-4
    and weight <= 12 truncation; observable mean of Z_i (coefficient 1/121)."""
    N = 121
    bonds = lattice_bonds(11, 11)
    layers = []
    for _ in range(n_layers):
        g = [Pauli(N, Z=[a, b]) for a, b in bonds] + [Pauli(N, X=[j]) for j in range(1, N + 1)]
        a = [-math.pi / 2] * len(bonds) + [theta_h] * N
        layers.append(Layer(g, a))
    obs = [(Pauli(N, Z=[j]), 1.0 / 121) for j in range(1, N + 1)]
    return Workload("lattice121_dephase", N, obs, layers,
                    CompositeTruncation(CoeffTruncation(1e-4), WeightTruncation(12)),
                    channel=("dephasing", 5e-4), meta={"theta_h": theta_h})


def surrogate_chain36_layers(theta: List[float]) -> Workload:
    """36-qubit chain, 10 layers; layer l uses RZZ angle theta[2l] on all bonds
    and RX angle theta[2l+1] on all sites; observable Z_18; |c| <= 1e-6."""
    N = 36
    layers = []
    for l in range(10):
        g = [Pauli(N, Z=[a, b]) for a, b in chain_bonds(N)] + [Pauli(N, X=[j]) for j in range(1, N + 1)]
        a = [theta[2 * l]] * (N - 1) + [theta[2 * l + 1]] * N
        layers.append(Layer(g, a))
    return Workload("chain36_surrogate", N, [(Pauli(N, Z=[18]), 1.0)], layers, CoeffTruncation(1e-6))


def surrogate_thetas(count: int, seed: int = 99) -> List[List[float]]:
    rng = Xoshiro256ss(seed)
    return [[-math.pi + 2 * math.pi * rng.uniform() for _ in range(20)] for _ in range(count)]


def grid_bonds(N: int) -> Tuple[int, int, List[Tuple[int, int]]]:
    """Near-square open grid holding N sites: rows = floor(sqrt N), cols =
    ceil(N / rows), sites row-major q = cols*r + c + 1 <= N (the last row may be
    partial); nearest-neighbour bonds between present sites, ordered
    lexicographically by (min site, max site)."""
    rows = max(1, int(math.isqrt(N)))
    cols = -(-N // rows)
    bonds = []
    for q in range(1, N + 1):
        r, c = divmod(q - 1, cols)
        if c + 1 < cols and q + 1 <= N:
            bonds.append((q, q + 1))
        if q + cols <= N:
            bonds.append((q, q + cols))
    bonds.sort()
    return rows, cols, bonds


def batched_instances(count: int = 256, seed: int = 7, max_depth: Optional[int] = None) -> List[Workload]:
    """The batched workload: `count` independent circuits from one seed-7
    stream, drawn instance by instance in this order: N ~ U{24..128},
    depth ~ U{5..40}, topology bit (0: open chain, 1: near-square grid,
    grid_bonds), observable site ~ U{1..N}; then per layer RZZ on every bond,
    RX on every site, RZ on every site, each angle i.i.d. U[0, π) in that
    order.  Observable Z_site; CoeffTruncation(1e-5); FP64.  max_depth (for
    tests) caps the number of layers kept without changing the draws of
    later instances."""
    rng = Xoshiro256ss(seed)
    out = []
    for i in range(count):
        N = rng.randint(24, 128)
        depth = rng.randint(5, 40)
        grid = rng.next() >> 63
        site = rng.randint(1, N)
        bonds = grid_bonds(N)[2] if grid else chain_bonds(N)
        gens = ([Pauli(N, Z=[a, b]) for a, b in bonds] + [Pauli(N, X=[j]) for j in range(1, N + 1)]
                + [Pauli(N, Z=[j]) for j in range(1, N + 1)])
        layers = []
        for _ in range(depth):
            layers.append(Layer(gens, [rng.uniform() * math.pi for _ in gens]))
        keep = layers if max_depth is None else layers[:max_depth]
        out.append(Workload(f"batched_{i}", N, [(Pauli(N, Z=[site]), 1.0)], keep, CoeffTruncation(1e-5),
                            meta={"index": i, "depth": depth, "topology": "grid" if grid else "chain",
                                  "site": site, "bonds": len(bonds)}))
    return out
