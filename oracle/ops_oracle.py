"""CPU restatement of the SparsePauliVector utilities around the evolution path.

TEST INFRASTRUCTURE ONLY: the checker for pp_otimes, pp_largest,
pp_select_diag and pp_matrix.  Terms are (z, x, c) triples with Python-int
bitstrings, in the reference's key order (z-major, spv_kernels.jl:14-16);
c is a float (Float64) or a complex (ComplexF64; Python's complex product is
Julia's (a re b re - a im b im, a re b im + a im b re)).
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

Term = Tuple[int, int, complex]


def key_sorted(terms: Sequence[Term]) -> List[Term]:
    return sorted(terms, key=lambda t: (t[0], t[1]))


def otimes(N1: int, t1: Sequence[Term], N2: int, t2: Sequence[Term]) -> List[Term]:
    """spv_ops.jl:342-358: every pair (op1 ⊗ op2, c1 * c2) with op1 ⊗ op2 =
    (z1 | z2 << N1, x1 | x2 << N1) (type_PauliBasis.jl:130-132), sorted by key
    (keys are distinct, so _triples_to_live! copies)."""
    out = [(z1 | (z2 << N1), x1 | (x2 << N1), c1 * c2) for z1, x1, c1 in t1 for z2, x2, c2 in t2]
    return key_sorted(out)


def add(t1: Sequence[Term], t2: Sequence[Term], zero) -> List[Term]:
    """v1 + v2 (_add_spv, spv_ops.jl:131-163): a two-pointer walk, zero(T) + c1 + c2."""
    out, i, j = [], 0, 0
    while i < len(t1) or j < len(t2):
        take1 = j >= len(t2) or (i < len(t1) and not (t2[j][:2] < t1[i][:2]))
        take2 = i >= len(t1) or (j < len(t2) and not (t1[i][:2] < t2[j][:2]))
        acc = zero
        if take1:
            k = t1[i][:2]
            acc = acc + t1[i][2]
            i += 1
        if take2:
            k = t2[j][:2]
            acc = acc + t2[j][2]
            j += 1
        out.append((k[0], k[1], acc))
    return out


def osum(N1: int, t1: Sequence[Term], N2: int, t2: Sequence[Term], one) -> List[Term]:
    """spv_ops.jl:360-366: v1 ⊗ I_M + I_N ⊗ v2."""
    return add(otimes(N1, t1, N2, [(0, 0, one)]), otimes(N1, [(0, 0, one)], N2, t2), one * 0)


def largest_index(terms: Sequence[Term], diag_only: bool = False) -> int:
    """spv_ops.jl:520-553: best = -1.0, the first term with abs(c) > best
    (x == 0 terms only for largest_diag); -1 when none (largest of a
    non-empty vector of NaNs returns its first term, bi = 1)."""
    best, bi = -1.0, -1
    for i, (z, x, c) in enumerate(terms):
        if diag_only and x != 0:
            continue
        a = abs(c)
        if a > best:
            best, bi = a, i
    return bi


def select_diag(terms: Sequence[Term], offdiag: bool) -> List[Term]:
    """diag (spv_ops.jl:571-587) keeps x == 0, offdiag (:555-569) x != 0."""
    return [t for t in terms if (t[1] == 0) != offdiag]


_P = {"I": np.array([[1, 0], [0, 1]]), "X": np.array([[0, 1], [1, 0]]),
      "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.array([[1, 0], [0, -1]])}


def pauli_basis_matrix(N: int, z: int, x: int) -> np.ndarray:
    """Matrix(::PauliBasis) (type_PauliBasis.jl:71-95): mat = kron(P_i, mat)
    for sites i = 1..N, so site 1 is the last factor."""
    mat = np.ones((1, 1), dtype=np.int64)
    for i in range(N):
        zi, xi = (z >> i) & 1, (x >> i) & 1
        letter = "I" if not (zi or xi) else ("X" if xi and not zi else ("Z" if zi and not xi else "Y"))
        mat = np.kron(_P[letter], mat)
    return mat


def matrix(N: int, terms: Sequence[Term], complex_coeffs: bool) -> np.ndarray:
    """Matrix(v) (spv_ops.jl:109-115): out .+= Matrix(op) .* c over the terms
    in key order.  A Float64 out cannot take an imaginary entry (InexactError)."""
    D = 1 << N
    out = np.zeros((D, D), dtype=np.complex128 if complex_coeffs else np.float64)
    for z, x, c in terms:
        m = pauli_basis_matrix(N, z, x) * c
        if not complex_coeffs:
            if np.iscomplexobj(m) and np.any(m.imag != 0):
                raise ValueError("InexactError: an imaginary entry in a Float64 matrix")
            m = m.real if np.iscomplexobj(m) else m
        out = out + m
    return out
