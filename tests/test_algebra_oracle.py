"""Pins of the Pauli-sum product oracle (oracle/algebra_oracle.py), CPU only.

* the reference's own known answers for commutators and anticommutators
  (test/test_phase1.jl:208-260) and its product/commutator error rules
  (test/test_spv_equivalence.jl:72-87);
* brute force: A·B, A†·B, A·B†, A†·B†, [A, B] and {A, B} against dense 2^N
  matrices of random complex sums on 3 and 4 qubits;
* the fused-phase identity against dense single-term products on 5 qubits;
* summation order: the stable order (the device's) against the reference's
  quicksort order restated, on inputs with long equal-key runs: the same keys
  and sums within 1e-12 of the largest |c|.
"""
import numpy as np
import pytest

import algebra_oracle as ao

I2 = np.eye(2, dtype=complex)
X2 = np.array([[0, 1], [1, 0]], dtype=complex)
Y2 = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z2 = np.array([[1, 0], [0, -1]], dtype=complex)


def dense_term(N, z, x):
    """Matrix of the basis Pauli with bit j of z, x for site j + 1 (I, X: x only,
    Z: z only, Y: both); site 1 is the first Kronecker factor."""
    m = np.ones((1, 1), dtype=complex)
    for j in range(N):
        zb, xb = (int(z) >> j) & 1, (int(x) >> j) & 1
        s = {(0, 0): I2, (0, 1): X2, (1, 0): Z2, (1, 1): Y2}[(zb, xb)]
        m = np.kron(m, s)
    return m


def dense(N, A):
    z, x, c = A
    out = np.zeros((2 ** N, 2 ** N), dtype=complex)
    for k in range(len(c)):
        out += c[k] * dense_term(N, z[k, 0], x[k, 0])
    return out


def random_sum(N, n, rng, cplx=True, support=None):
    bits = support if support is not None else N
    z = rng.integers(0, 2 ** bits, size=(n, 1), dtype=np.uint64)
    x = rng.integers(0, 2 ** bits, size=(n, 1), dtype=np.uint64)
    c = rng.standard_normal(n)
    if cplx:
        c = c + 1j * rng.standard_normal(n)
    return z, x, c


def single(z, x, c):
    return (np.array([[z]], np.uint64), np.array([[x]], np.uint64), np.array([c]))


# site 1 is bit 0: "XI" = X on site 1
XI, YI, ZI, IZ, IX = (0, 1), (1, 1), (1, 0), (2, 0), (0, 2)


def as_dict(out):
    z, x, c = out
    return {(int(z[k, 0]), int(x[k, 0])): c[k] for k in range(len(c))}


def test_reference_commutator_answers():
    """test_phase1.jl:208-235."""
    A = single(*XI, 1 + 0j)
    B = single(*YI, 1 + 0j)
    assert as_dict(ao.product(A, B, ao.COMMUTATOR)) == {ZI: 2j}
    assert len(ao.product(single(*ZI, 1 + 0j), single(*IZ, 1 + 0j), ao.COMMUTATOR)[2]) == 0
    A2 = (np.array([[XI[0]], [IZ[0]]], np.uint64), np.array([[XI[1]], [IZ[1]]], np.uint64), np.array([1 + 0j, 1]))
    B2 = (np.array([[YI[0]], [IX[0]]], np.uint64), np.array([[YI[1]], [IX[1]]], np.uint64), np.array([1 + 0j, 1]))
    a, b = dense(2, A2), dense(2, B2)
    assert np.allclose(dense(2, ao.product(A2, B2, ao.COMMUTATOR)), a @ b - b @ a, atol=1e-14)


def test_reference_anticommutator_answers():
    """test_phase1.jl:237-260."""
    X = single(*XI, 1 + 0j)
    assert len(ao.product(X, single(*YI, 1 + 0j), ao.ANTICOMMUTATOR)[2]) == 0
    assert as_dict(ao.product(X, X, ao.ANTICOMMUTATOR)) == {(0, 0): 2 + 0j}
    A2 = (np.array([[XI[0]], [IZ[0]]], np.uint64), np.array([[XI[1]], [IZ[1]]], np.uint64), np.array([1 + 0j, 1]))
    B2 = (np.array([[YI[0]], [IX[0]]], np.uint64), np.array([[YI[1]], [IX[1]]], np.uint64), np.array([1 + 0j, 1]))
    a, b = dense(2, A2), dense(2, B2)
    assert np.allclose(dense(2, ao.product(A2, B2, ao.ANTICOMMUTATOR)), a @ b + b @ a, atol=1e-14)


def test_reference_error_rules():
    """A real-typed commutator is an error (test_spv_equivalence.jl:83-86);
    a real product whose phase is imaginary fails the reference's convert."""
    r = single(1, 0, 1.0)
    with pytest.raises(ValueError):
        ao.product(r, r, ao.COMMUTATOR)
    with pytest.raises(ValueError, match="InexactError"):
        ao.product(single(*XI, 1.0), single(*YI, 1.0), ao.MUL)  # XY = iZ
    assert as_dict(ao.product(single(*XI, 2.0), single(*XI, 3.0), ao.MUL)) == {(0, 0): 6.0}
    assert as_dict(ao.product(single(*XI, 2.0), single(*XI, 3.0), ao.ANTICOMMUTATOR)) == {(0, 0): 12.0}


@pytest.mark.parametrize("N,seed", [(3, 1), (4, 2), (4, 3)])
def test_dense_brute_force(N, seed):
    rng = np.random.default_rng(seed)
    A = random_sum(N, 18, rng)
    B = random_sum(N, 14, rng)
    a, b = dense(N, A), dense(N, B)
    ah, bh = a.conj().T, b.conj().T
    for (ca, cb), want in {(False, False): a @ b, (True, False): ah @ b, (False, True): a @ bh,
                           (True, True): ah @ bh}.items():
        got = dense(N, ao.product(A, B, ao.MUL, ca, cb))
        assert np.allclose(got, want, atol=1e-12), (ca, cb)
    assert np.allclose(dense(N, ao.product(A, B, ao.COMMUTATOR)), a @ b - b @ a, atol=1e-12)
    assert np.allclose(dense(N, ao.product(A, B, ao.ANTICOMMUTATOR)), a @ b + b @ a, atol=1e-12)


def test_fused_phase_identity_against_dense_products():
    """Every pair of 5-qubit basis terms from a random sample: the oracle's
    key and phase equal the dense product (a transposed operand or a wrong
    popcount term fails here)."""
    N = 5
    rng = np.random.default_rng(7)
    for _ in range(300):
        za, xa, zb, xb = (int(v) for v in rng.integers(0, 32, size=4))
        out = ao.product(single(za, xa, 1 + 0j), single(zb, xb, 1 + 0j), ao.MUL)
        assert len(out[2]) == 1
        want = dense_term(N, za, xa) @ dense_term(N, zb, xb)
        got = out[2][0] * dense_term(N, out[0][0, 0], out[1][0, 0])
        assert np.allclose(got, want, atol=1e-14)


@pytest.mark.parametrize("op", [ao.MUL, ao.COMMUTATOR, ao.ANTICOMMUTATOR])
def test_reference_order_agrees_with_stable_order(op):
    """Keys on 3 qubits of a 10-qubit space: ~60 pairs per output key, so the
    two summation orders differ; the sums agree to 1e-12 of the scale."""
    N = 10
    rng = np.random.default_rng(11 + op)
    A = random_sum(N, 70, rng, support=3)
    B = random_sum(N, 60, rng, support=3)
    s = ao.product(A, B, op, order="stable")
    r = ao.product(A, B, op, order="reference")
    assert np.array_equal(s[0], r[0]) and np.array_equal(s[1], r[1])
    scale = np.max(np.abs(s[2]))
    assert np.max(np.abs(s[2] - r[2])) <= 1e-12 * scale
    # the keys are strictly increasing in the reference's (z, x) order
    keys = [(int(z), int(x)) for z, x in zip(s[0][:, 0], s[1][:, 0])]
    assert keys == sorted(set(keys))


def test_run_sums_are_left_to_right():
    """Three equal keys summed as ((a + b) + c), not in another order: values
    chosen so that the orders round differently."""
    z = np.zeros((3, 1), np.uint64)
    x = np.zeros((3, 1), np.uint64)
    re = np.array([1.0, 1e16, -1e16])
    im = np.zeros(3)
    _, _, c = ao.reduce_runs(z, x, re, im, False)
    assert c[0] == (1.0 + 1e16) + -1e16 == 0.0
    assert 1.0 + (1e16 + -1e16) == 1.0
