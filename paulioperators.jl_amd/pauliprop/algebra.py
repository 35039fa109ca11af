"""Pauli-sum products on the device (spv_ops.jl:255-484).

``multiply(N, A, B; conj_a, conj_b)`` is ``_mul_spv`` (``A * B``, ``A' * B``,
``A * B'``, ``A' * B'`` and a single Pauli as a one-term operand),
``commutator`` is ``commutator``/``commutator!`` (``[A, B] = AB - BA``) and
``anticommutator`` is ``anticommutator``/``anticommutator!``.  Operands are
SparsePauliVector objects (Float64 coefficients) or ``(z, x, c)`` arrays: z, x
of shape (n, W) uint64 (any order) and c float64 or complex128 (the
reference's ComplexF64).  The result is ``(z, x, c)`` in key order with
equal keys summed (zero sums kept, as the reference's ``_triples_to_live!``);
for two SparsePauliVector operands a SparsePauliVector.

Float64 coefficients follow the reference's conversions: a product with a
nonzero imaginary part raises (the reference's InexactError), and the
commutator raises ("commutator coefficients are imaginary").
"""
from __future__ import annotations

import ctypes
from typing import Dict, Tuple

import numpy as np

from . import _lib
from .basis import word_count
from ._lib import check
from .spv import SparsePauliVector

_handles: Dict[int, ctypes.c_void_p] = {}


def _handle(W: int) -> ctypes.c_void_p:
    """A device context of word count W (its live set is never used)."""
    h = _handles.get(W)
    if h is None:
        h = ctypes.c_void_p()
        check(_lib.lib().pp_create(64 * W, 16, 16, ctypes.byref(h)))
        _handles[W] = h
    return h


def _arrays(N: int, t):
    if isinstance(t, SparsePauliVector):
        if t.N != N:
            raise _lib.ArgumentError("qubit count mismatch")
        return t.arrays()
    z, x, c = t
    W = word_count(N)
    z = np.ascontiguousarray(np.asarray(z, dtype=np.uint64).reshape(-1, W))
    x = np.ascontiguousarray(np.asarray(x, dtype=np.uint64).reshape(-1, W))
    c = np.asarray(c)
    if len(z) != len(c) or len(x) != len(c):
        raise _lib.DimensionMismatch("z, x and c must have the same length")
    return z, x, c


def product(N: int, A, B, op: int, conj_a: bool = False, conj_b: bool = False) -> Tuple[np.ndarray, ...]:
    """pp_algebra: op = PP_ALG_MUL / PP_ALG_COMMUTATOR / PP_ALG_ANTICOMMUTATOR."""
    W = word_count(N)
    az, ax, ac = _arrays(N, A)
    bz, bx, bc = _arrays(N, B)
    cplx = np.iscomplexobj(ac) or np.iscomplexobj(bc)
    keep = []

    def terms(z, x, c, conj):
        re = np.ascontiguousarray(np.real(c), dtype=np.float64)
        im = np.ascontiguousarray(np.imag(c), dtype=np.float64) if cplx else None
        keep.extend([z, x, re, im])
        P = ctypes.POINTER
        return _lib.PPTerms(z.ctypes.data_as(P(ctypes.c_uint64)), x.ctypes.data_as(P(ctypes.c_uint64)),
                            re.ctypes.data_as(P(ctypes.c_double)),
                            im.ctypes.data_as(P(ctypes.c_double)) if im is not None else None,
                            len(c), 1 if conj else 0)

    ta = terms(az, ax, ac, conj_a)
    tb = terms(bz, bx, bc, conj_b)
    L = _lib.lib()
    h = _handle(W)
    n = ctypes.c_int64()
    u64 = ctypes.POINTER(ctypes.c_uint64)
    dbl = ctypes.POINTER(ctypes.c_double)
    # first try with room for every pair up to 4M terms; a larger result is
    # reported with its size and computed again into arrays of that size
    cap = max(1, min(len(ac) * len(bc), 1 << 22))
    for _ in range(2):
        oz = np.zeros((cap, W), dtype=np.uint64)
        ox = np.zeros((cap, W), dtype=np.uint64)
        ore = np.zeros(cap, dtype=np.float64)
        oim = np.zeros(cap, dtype=np.float64)
        rc = L.pp_algebra(h, op, 1 if cplx else 0, ctypes.byref(ta), ctypes.byref(tb), oz.ctypes.data_as(u64),
                          ox.ctypes.data_as(u64), ore.ctypes.data_as(dbl), oim.ctypes.data_as(dbl), cap,
                          ctypes.byref(n))
        if rc != _lib.PP_EOVERFLOW:
            break
        cap = n.value
    check(rc)
    k = n.value
    if cplx:
        c = np.empty(k, dtype=np.complex128)
        c.real = ore[:k]
        c.imag = oim[:k]
    else:
        c = ore[:k].copy()
    return oz[:k], ox[:k], c


def _wrap(N, A, B, out):
    if isinstance(A, SparsePauliVector) and isinstance(B, SparsePauliVector):
        return SparsePauliVector(N, out)
    return out


def multiply(N: int, A, B, conj_a: bool = False, conj_b: bool = False):
    """``_mul_spv(A, conj_a, B, conj_b)`` (spv_ops.jl:257-298)."""
    return _wrap(N, A, B, product(N, A, B, _lib.PP_ALG_MUL, conj_a, conj_b))


def commutator(N: int, A, B):
    """``commutator(A, B) = AB - BA`` (spv_ops.jl:418-460, 474-479)."""
    return product(N, A, B, _lib.PP_ALG_COMMUTATOR)


def anticommutator(N: int, A, B):
    """``anticommutator(A, B) = AB + BA`` (spv_ops.jl:462-472, 481-484)."""
    return _wrap(N, A, B, product(N, A, B, _lib.PP_ALG_ANTICOMMUTATOR))
