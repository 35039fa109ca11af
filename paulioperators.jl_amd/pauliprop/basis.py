"""Host-side Pauli algebra: PauliBasis, Pauli construction, Ket, PauliSum.

Mirrors the bit encoding of PauliOperators.jl (src/type_PauliBasis.jl:15-64,
docs/src/representation.md:12-59): site j (1-based) is bit j-1 of the
integers ``z`` and ``x``; letters are I=(0,0), X=(0,1), Z=(1,0), Y=(1,1);
the first character of a string is site 1.  These objects are construction
and conversion helpers only; nothing here is on the evolution path.
"""
from __future__ import annotations

import random
from typing import Dict, Iterable, Optional


def word_count(N: int) -> int:
    """Packed 64-bit words per z (or x) bitstring.

    type_SparsePauliVector.jl:23-26 picks UInt64 for N <= 64 and UInt128 for
    N <= 128; here that is 1 or 2 little-endian uint64 words.
    """
    if N < 1:
        raise ValueError("N must be >= 1")
    if N <= 64:
        return 1
    if N <= 128:
        return 2
    raise ValueError("N > 128 requires a multi-word type (at most 128 qubits are supported)")


class PauliBasis:
    """Hermitian Pauli string on N qubits (type_PauliBasis.jl:15-23)."""

    __slots__ = ("N", "z", "x")

    def __init__(self, N: int, z: int = 0, x: int = 0):
        mask = (1 << N) - 1
        if z & ~mask or x & ~mask:
            raise ValueError("bits set beyond qubit N")
        self.N = int(N)
        self.z = int(z)
        self.x = int(x)

    @classmethod
    def from_string(cls, s: str) -> "PauliBasis":
        """type_PauliBasis.jl:38-64 — first character is site 1 (the LSB)."""
        z = x = 0
        for j, ch in enumerate(s):
            if ch not in "IXYZ":
                raise ValueError("Bad string: " + s)
            if ch in "XY":
                x |= 1 << j
            if ch in "ZY":
                z |= 1 << j
        return cls(len(s), z, x)

    def __str__(self) -> str:
        out = []
        for j in range(self.N):
            zb = (self.z >> j) & 1
            xb = (self.x >> j) & 1
            out.append("IXZY"[zb * 2 + xb])
        return "".join(out)

    def __repr__(self) -> str:
        return f"PauliBasis({self})"

    def __eq__(self, other) -> bool:
        return isinstance(other, PauliBasis) and (self.N, self.z, self.x) == (other.N, other.z, other.x)

    def __hash__(self) -> int:
        return hash((self.N, self.z, self.x))

    def key(self):
        """Sort key of the reference's (z, x) z-major order (spv_kernels.jl:14-16)."""
        return (self.z, self.x)


def Pauli(N: int, X: Iterable[int] = (), Y: Iterable[int] = (), Z: Iterable[int] = ()) -> PauliBasis:
    """``PauliBasis(Pauli(N; X=..., Y=..., Z=...))`` (type_Pauli.jl:128-149), sites 1-based."""
    z = x = 0
    for q in X:
        x |= 1 << (q - 1)
    for q in Y:
        x |= 1 << (q - 1)
        z |= 1 << (q - 1)
    for q in Z:
        z |= 1 << (q - 1)
    return PauliBasis(N, z, x)


def weight(p: PauliBasis) -> int:
    """clip.jl:6-8 — number of non-identity sites."""
    return bin(p.z | p.x).count("1")


def x_weight(p: PauliBasis) -> int:
    """clip.jl:53-55 — number of X/Y sites."""
    return bin(p.x).count("1")


def majorana_weight(p: PauliBasis) -> int:
    """clip.jl:88-117 — Jordan-Wigner Majorana weight via suffix parity on 128 bits."""
    m128 = (1 << 128) - 1
    xb = p.x & m128
    zb = p.z & ~xb & m128
    S = xb
    for sh in (1, 2, 4, 8, 16, 32, 64):
        S ^= S >> sh
    ctrl = ~(S ^ xb) & m128
    return (bin(xb).count("1") + 2 * bin(zb & ctrl).count("1")
            + 2 * bin(~(p.z | xb) & ~ctrl & m128).count("1"))


def commute(a: PauliBasis, b: PauliBasis) -> bool:
    """type_PauliBasis.jl:152 — symplectic parity test."""
    return (bin(a.x & b.z).count("1") - bin(a.z & b.x).count("1")) % 2 == 0


class Ket:
    """Computational-basis ket |v> (type_Ket.jl:4-6); site 1 is bit 0 of v."""

    __slots__ = ("N", "v")

    def __init__(self, N: int, v: int = 0):
        self.N = int(N)
        self.v = int(v)


class PauliSum(dict):
    """Dict-backed sum of PauliBasis -> coefficient (type_PauliSum.jl:8-11)."""

    def __init__(self, N: int, items: Optional[Dict[PauliBasis, complex]] = None):
        super().__init__()
        self.N = int(N)
        if items:
            for k, v in items.items():
                self[k] = v

    def copy(self) -> "PauliSum":
        out = PauliSum(self.N)
        out.update(self)
        return out


def rand_paulibasis(N: int, rng: random.Random) -> PauliBasis:
    """Uniform random Pauli string (type_PauliBasis.jl:122-125), from a Python RNG."""
    return PauliBasis(N, rng.getrandbits(N), rng.getrandbits(N))
