"""GPU parity of the Pauli-sum products (pp_algebra) against the oracle
(oracle/algebra_oracle.py) in its stable summation order: keys and the bit
patterns of the coefficients (real and imaginary parts) must be equal.

Inputs are seeded random sums whose keys live on a few qubits of a wider
space, so that many pairs share a product key (long runs to sum); both key
widths (N <= 64 and N > 64) and every operation, including the adjoint
products, the reference's real-coefficient rules and empty operands.
"""
import numpy as np
import pytest

import algebra_oracle as ao
from pauliprop import ArgumentError, SparsePauliVector, anticommutator, commutator, multiply
from pauliprop import _lib
from pauliprop.algebra import product

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def random_sum(N, n, rng, cplx=True, support_bits=None, dup_free=False):
    W = 1 if N <= 64 else 2
    sb = support_bits or N
    z = np.zeros((n, W), np.uint64)
    x = np.zeros((n, W), np.uint64)
    # support on bits spread over the whole width (both words when W = 2)
    pos = np.linspace(0, N - 1, sb).astype(int)
    for k in range(n):
        zv = int(sum(1 << int(p) for p in pos[rng.random(sb) < 0.5]))
        xv = int(sum(1 << int(p) for p in pos[rng.random(sb) < 0.5]))
        for w in range(W):
            z[k, w] = (zv >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
            x[k, w] = (xv >> (64 * w)) & 0xFFFFFFFFFFFFFFFF
    c = rng.standard_normal(n)
    if cplx:
        c = c + 1j * rng.standard_normal(n)
    return z, x, c


def same_bits(got, want):
    gz, gx, gc = got
    wz, wx, wc = want
    assert gz.shape == wz.shape, (gz.shape, wz.shape)
    assert np.array_equal(gz, wz) and np.array_equal(gx, wx), "keys"
    assert np.array_equal(np.real(gc).view(np.uint64), np.real(wc).view(np.uint64)), "real parts"
    if np.iscomplexobj(wc):
        assert np.array_equal(np.imag(gc).view(np.uint64), np.imag(wc).view(np.uint64)), "imaginary parts"


@pytest.mark.parametrize("N", [40, 100])
@pytest.mark.parametrize("op", [ao.MUL, ao.COMMUTATOR, ao.ANTICOMMUTATOR])
def test_complex_products_bitexact(N, op):
    rng = np.random.default_rng(100 * N + op)
    A = random_sum(N, 300, rng, support_bits=6)
    B = random_sum(N, 250, rng, support_bits=6)
    log(f"pp_algebra op={op} N={N} pairs={300 * 250}")
    same_bits(product(N, A, B, op), ao.product(A, B, op))


@pytest.mark.parametrize("conj", [(True, False), (False, True), (True, True)])
def test_adjoint_products_bitexact(conj):
    N = 70
    rng = np.random.default_rng(7)
    A = random_sum(N, 120, rng, support_bits=5)
    B = random_sum(N, 90, rng, support_bits=5)
    log(f"pp_algebra adjoint {conj}")
    same_bits(multiply(N, A, B, *conj), ao.product(A, B, ao.MUL, *conj))


def test_large_product_bitexact():
    """2.56e6 pairs onto ~4e3 keys (runs of ~600 terms), N = 48."""
    N = 48
    rng = np.random.default_rng(3)
    A = random_sum(N, 1600, rng, support_bits=6)
    B = random_sum(N, 1600, rng, support_bits=6)
    log("pp_algebra large product")
    same_bits(product(N, A, B, ao.MUL), ao.product(A, B, ao.MUL))


def test_real_coefficients_follow_the_reference():
    """Float64 operands: anticommutator and real-phase products are bit-exact
    SparsePauliVectors; a product with an imaginary phase and the commutator
    raise, as the reference's convert and commutator! do."""
    N = 30
    rng = np.random.default_rng(5)
    Az, Ax, Ac = random_sum(N, 200, rng, cplx=False, support_bits=8)
    Bz, Bx, Bc = random_sum(N, 150, rng, cplx=False, support_bits=8)
    A = (Az, Ax, Ac)
    B = (Bz, Bx, Bc)
    log("pp_algebra real anticommutator")
    same_bits(anticommutator(N, A, B), ao.product(A, B, ao.ANTICOMMUTATOR))
    # Z-only sums commute term by term: every product phase is real
    Zs = (Az, np.zeros_like(Ax), Ac)
    Zt = (Bz, np.zeros_like(Bx), Bc)
    same_bits(multiply(N, Zs, Zt), ao.product(Zs, Zt, ao.MUL))
    with pytest.raises(ArgumentError, match="InexactError"):
        multiply(N, A, B)
    with pytest.raises(ArgumentError, match="imaginary"):
        commutator(N, A, B)
    # SparsePauliVector operands give a SparsePauliVector
    va = SparsePauliVector(N, {(int(z[0]), int(x[0])): float(c) for z, x, c in zip(Az[:50], Ax[:50], Ac[:50])})
    vb = SparsePauliVector(N, {(int(z[0]), int(x[0])): float(c) for z, x, c in zip(Bz[:40], Bx[:40], Bc[:40])})
    out = anticommutator(N, va, vb)
    assert isinstance(out, SparsePauliVector)
    same_bits(out.arrays(), ao.product(va.arrays(), vb.arrays(), ao.ANTICOMMUTATOR))


def test_empty_and_capacity():
    """Empty operands give an empty result; a capacity below the result size
    reports the size (PP_EOVERFLOW) without writing."""
    import ctypes
    N = 20
    rng = np.random.default_rng(9)
    A = random_sum(N, 30, rng, support_bits=4)
    E = (np.zeros((0, 1), np.uint64), np.zeros((0, 1), np.uint64), np.zeros(0, np.complex128))
    for op in (ao.MUL, ao.COMMUTATOR, ao.ANTICOMMUTATOR):
        z, x, c = product(N, A, E, op)
        assert len(c) == 0
    want = ao.product(A, A, ao.MUL)
    re = np.ascontiguousarray(np.real(A[2]))
    im = np.ascontiguousarray(np.imag(A[2]))
    P = ctypes.POINTER
    t = _lib.PPTerms(A[0].ctypes.data_as(P(ctypes.c_uint64)), A[1].ctypes.data_as(P(ctypes.c_uint64)),
                     re.ctypes.data_as(P(ctypes.c_double)), im.ctypes.data_as(P(ctypes.c_double)), len(re), 0)
    from pauliprop.algebra import _handle
    oz = np.zeros((1, 1), np.uint64)
    ox = np.zeros((1, 1), np.uint64)
    ore = np.full(1, 7.0)
    oim = np.full(1, 7.0)
    n = ctypes.c_int64()
    rc = _lib.lib().pp_algebra(_handle(1), _lib.PP_ALG_MUL, 1, ctypes.byref(t), ctypes.byref(t),
                               oz.ctypes.data_as(P(ctypes.c_uint64)), ox.ctypes.data_as(P(ctypes.c_uint64)),
                               ore.ctypes.data_as(P(ctypes.c_double)), oim.ctypes.data_as(P(ctypes.c_double)), 1,
                               ctypes.byref(n))
    assert rc == _lib.PP_EOVERFLOW and n.value == len(want[2]) > 1
    assert ore[0] == 7.0 and oim[0] == 7.0


def test_single_run_of_all_pairs():
    """Every pair lands on one key (A and B hold one basis string each, with
    repeats): one run of nA x nB terms summed left to right in generation
    order, as the reference's scan over its sorted workspace does."""
    N = 50
    rng = np.random.default_rng(21)
    na, nb = 400, 400
    A = (np.full((na, 1), 0b1011, np.uint64), np.full((na, 1), 0b0110, np.uint64),
         rng.standard_normal(na) + 1j * rng.standard_normal(na))
    B = (np.full((nb, 1), 0b0011, np.uint64), np.full((nb, 1), 0b0101, np.uint64),
         rng.standard_normal(nb) * 1e8 + 1j * rng.standard_normal(nb))
    log("pp_algebra one run of 160000 pairs")
    got = product(N, A, B, ao.MUL)
    assert len(got[2]) == 1
    same_bits(got, ao.product(A, B, ao.MUL))


def test_commuting_operands_and_full_width_keys():
    """Z-only sums commute: an empty commutator; keys that use the top bits
    of N = 128 (both words) give the same products as the oracle."""
    N = 128
    rng = np.random.default_rng(23)
    n = 200
    z = rng.integers(0, 2 ** 63, size=(n, 2), dtype=np.uint64) | np.uint64(1 << 63)
    A = (z, np.zeros_like(z), rng.standard_normal(n) + 1j * rng.standard_normal(n))
    B = (z[::-1].copy(), np.zeros_like(z), rng.standard_normal(n) + 0j)
    log("pp_algebra commuting Z sums, N = 128")
    assert len(product(N, A, B, ao.COMMUTATOR)[2]) == 0
    x = rng.integers(0, 2 ** 63, size=(n, 2), dtype=np.uint64) | np.uint64(1 << 63)
    C = (z, x, rng.standard_normal(n) + 1j * rng.standard_normal(n))
    for op in (ao.MUL, ao.COMMUTATOR, ao.ANTICOMMUTATOR):
        same_bits(product(N, A, C, op), ao.product(A, C, op))


def test_buffers_reused_across_calls():
    """The device buffers of a product stay on the handle between calls:
    a large call, a smaller one, then a larger one than the first (each
    buffer regrows) all equal the oracle."""
    N = 90
    rng = np.random.default_rng(31)
    for na, nb, sb in ((900, 800, 7), (40, 30, 4), (1400, 1300, 8), (60, 50, 5)):
        A = random_sum(N, na, rng, support_bits=sb)
        B = random_sum(N, nb, rng, support_bits=sb)
        log(f"pp_algebra reuse {na}x{nb}")
        same_bits(product(N, A, B, ao.COMMUTATOR), ao.product(A, B, ao.COMMUTATOR))
