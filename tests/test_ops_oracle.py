"""Pins oracle/ops_oracle.py (otimes, osum, largest, largest_diag, diag,
offdiag, Matrix: spv_ops.jl:109-115,342-366,520-587) by identities a
restatement mistake breaks: Matrix(v1 ⊗ v2) = kron(Matrix(v2), Matrix(v1))
(v2 on the high sites), Matrix(osum) = kron(I, M1) + kron(M2, I), the
restated Matrix against the independent kron construction of
tests/dictpath.py and tr(Matrix(v)) = 2^N c_I, largest against numpy's first
argmax, and diag + offdiag = v with Matrix(diag(v)) diagonal."""
import random

import numpy as np
import pytest

import dictpath as dp
import ops_oracle as oo


def rand_terms(rng, N, n, cplx):
    n = min(n, 4 ** N)
    d = {}
    while len(d) < n:
        k = (rng.getrandbits(N), rng.getrandbits(N))
        d[k] = complex(rng.gauss(0, 1), rng.gauss(0, 1)) if cplx else rng.gauss(0, 1)
    return oo.key_sorted([(z, x, c) for (z, x), c in d.items()])


def M(N, terms, cplx=True):
    return oo.matrix(N, terms, cplx)


@pytest.mark.parametrize("cplx", [False, True])
def test_otimes_is_kron_with_v2_high(cplx):
    rng = random.Random(1)
    for N1, N2 in ((1, 2), (2, 3), (3, 1)):
        a, b = rand_terms(rng, N1, 5, cplx), rand_terms(rng, N2, 6, cplx)
        out = oo.otimes(N1, a, N2, b)
        assert len(out) == len(a) * len(b) and out == oo.key_sorted(out)
        assert np.allclose(M(N1 + N2, out), np.kron(M(N2, b), M(N1, a)), atol=1e-13)


def test_osum_is_kron_sum():
    rng = random.Random(2)
    N1, N2 = 2, 3
    a, b = rand_terms(rng, N1, 6, True), rand_terms(rng, N2, 7, True)
    ref = np.kron(np.eye(2 ** N2), M(N1, a)) + np.kron(M(N2, b), np.eye(2 ** N1))
    assert np.allclose(M(N1 + N2, oo.osum(N1, a, N2, b, 1 + 0j)), ref, atol=1e-13)


def test_matrix_restatement_against_independent_kron():
    rng = random.Random(3)
    for N in (1, 3, 5):
        t = rand_terms(rng, N, min(20, 4 ** N), True)
        ref = dp.sum_matrix(N, {(z, x): c for z, x, c in t})
        assert np.allclose(M(N, t), ref, atol=1e-13)
        ident = [c for z, x, c in t if z == 0 and x == 0]
        assert abs(np.trace(M(N, t)) - (2 ** N) * (ident[0] if ident else 0)) < 1e-12
    # Float64: a term with one Y gives an imaginary entry (InexactError)
    with pytest.raises(ValueError):
        oo.matrix(2, [(1, 1, 0.5)], False)
    # two Y: real, fine
    assert np.allclose(oo.matrix(2, [(3, 3, 0.5)], False), 0.5 * np.kron(dp._Y, dp._Y).real)


def test_largest_first_maximum():
    rng = random.Random(4)
    for cplx in (False, True):
        t = rand_terms(rng, 6, 200, cplx)
        a = np.array([abs(c) for _, _, c in t])
        assert oo.largest_index(t) == int(np.argmax(a))
        dia = [i for i, (z, x, c) in enumerate(t) if x == 0]
        want = dia[int(np.argmax(a[dia]))] if dia else -1
        assert oo.largest_index(t, diag_only=True) == want
    tie = [(1, 0, 0.5), (2, 0, -0.5), (3, 1, 0.5)]
    assert oo.largest_index(tie) == 0 and oo.largest_index(tie, True) == 0
    assert oo.largest_index([]) == -1 and oo.largest_index([(1, 1, 2.0)], True) == -1


def test_diag_offdiag_partition():
    rng = random.Random(5)
    N = 4
    t = rand_terms(rng, N, 60, True)
    d, o = oo.select_diag(t, False), oo.select_diag(t, True)
    assert len(d) + len(o) == len(t) and all(x == 0 for _, x, _ in d) and all(x != 0 for _, x, _ in o)
    md, mo = M(N, d), M(N, o)
    assert np.allclose(md, np.diag(np.diag(M(N, t))), atol=1e-13)
    assert np.allclose(np.diag(mo), 0) and np.allclose(md + mo, M(N, t), atol=1e-13)
