"""GPU parity of the SparsePauliVector utilities around the evolution path
against oracle/ops_oracle.py: otimes/osum (spv_ops.jl:342-366) bit for bit
(keys, order, coefficient bit patterns) across word-count boundaries,
largest/largest_diag (:520-553) including ties and the empty cases,
diag/offdiag (:555-587), Matrix(v) (:109-115) for Float64 and ComplexF64, and
copy (type_SparsePauliVector.jl:249-254)."""
import random

import numpy as np
import pytest

import ops_oracle as oo
from pauliprop import (ArgumentError, PauliBasis, SparsePauliVector, diag, largest, largest_diag, offdiag,
                       osum, otimes, Matrix)

pytestmark = pytest.mark.gpu


def rand_terms(rng, N, n, cplx):
    n = min(n, 4 ** N)
    d = {}
    while len(d) < n:
        k = (rng.getrandbits(N), rng.getrandbits(N))
        d[k] = complex(rng.gauss(0, 1), rng.gauss(0, 1)) if cplx else rng.gauss(0, 1)
    return oo.key_sorted([(z, x, c) for (z, x), c in d.items()])


def spv(N, terms, cplx):
    return SparsePauliVector(N, [((z, x), c) for z, x, c in terms], T=complex if cplx else float)


def assert_terms(v, terms, what=""):
    got = v.terms()
    assert [(z, x) for z, x, _ in got] == [(z, x) for z, x, _ in terms], what
    dt = np.complex128 if v.complex else np.float64
    a = np.array([c for _, _, c in got], dtype=dt)
    b = np.array([c for _, _, c in terms], dtype=dt)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), what


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("N1,N2", [(3, 5), (40, 24), (40, 30), (64, 64), (70, 50), (100, 28), (1, 127)])
def test_otimes(N1, N2, cplx):
    rng = random.Random(N1 * 131 + N2)
    a, b = rand_terms(rng, N1, 300, cplx), rand_terms(rng, N2, 200, cplx)
    out = otimes(spv(N1, a, cplx), spv(N2, b, cplx))
    assert out.N == N1 + N2 and out.check()
    assert_terms(out, oo.otimes(N1, a, N2, b), f"otimes {N1}x{N2}")
    assert out.capacities() == (max(16, len(a) * len(b)), 16)


@pytest.mark.parametrize("cplx", [False, True])
def test_osum(cplx):
    rng = random.Random(9)
    N1, N2 = 30, 50
    a, b = rand_terms(rng, N1, 100, cplx), rand_terms(rng, N2, 80, cplx)
    a.insert(0, (0, 0, 0.75 + (0.5j if cplx else 0)))
    b.insert(0, (0, 0, -0.25))
    one = 1 + 0j if cplx else 1.0
    out = osum(spv(N1, a, cplx), spv(N2, b, cplx))
    assert_terms(out, oo.osum(N1, a, N2, b, one), "osum")


@pytest.mark.parametrize("cplx", [False, True])
def test_largest_and_diag(cplx):
    rng = random.Random(10)
    for N in (6, 64, 100):
        t = rand_terms(rng, N, 5000, cplx)
        # planted diagonal terms and a tie at the maximum
        extra = {}
        while len(extra) < 40:
            extra[(rng.getrandbits(N), 0)] = complex(rng.gauss(0, 1), rng.gauss(0, 1)) if cplx else rng.gauss(0, 1)
        t = oo.key_sorted(list({(z, x): c for z, x, c in t + [(k[0], k[1], c) for k, c in extra.items()]}.items()))
        t = [(k[0], k[1], c) for k, c in t]
        top = max(abs(c) for _, _, c in t) * 2
        i1, i2 = 17, len(t) - 5
        t[i1] = (t[i1][0], t[i1][1], top)
        t[i2] = (t[i2][0], t[i2][1], -top)
        v = spv(N, t, cplx)
        i = oo.largest_index(t)
        assert i == i1
        lv = largest(v)
        assert lv.terms() == [(t[i][0], t[i][1], t[i][2])]
        j = oo.largest_index(t, diag_only=True)
        p, c = largest_diag(v)
        assert (p.z, p.x, c) == t[j] and p.x == 0
        assert_terms(diag(v), oo.select_diag(t, False), "diag")
        assert_terms(offdiag(v), oo.select_diag(t, True), "offdiag")
        assert_terms(v, t, "input untouched")
    with pytest.raises(ArgumentError):
        largest(SparsePauliVector(8, []))
    with pytest.raises(ArgumentError):
        largest_diag(SparsePauliVector(8, [((0, 1), 1.0)]))


@pytest.mark.parametrize("cplx", [False, True])
def test_matrix(cplx):
    rng = random.Random(12)
    for N in (1, 4, 8):
        t = rand_terms(rng, N, 300, cplx)
        if not cplx:  # Float64: only terms with an even number of Y
            t = [(z, x, c) for z, x, c in t if bin(z & x).count("1") % 2 == 0]
        v = spv(N, t, cplx)
        got = Matrix(v)
        ref = oo.matrix(N, t, cplx)
        assert got.dtype == ref.dtype and np.array_equal(got, ref), N
    with pytest.raises(ArgumentError):
        Matrix(SparsePauliVector(3, [((1, 1), 0.5)]))
    assert np.array_equal(Matrix(SparsePauliVector(3, [((1, 1), 0.0)])), np.zeros((8, 8)))


def test_copy():
    rng = random.Random(13)
    for cplx in (False, True):
        t = rand_terms(rng, 90, 3000, cplx)
        v = spv(90, t, cplx)
        v.coeff_clip_(0.5)
        w = v.copy()
        assert w.capacities() == v.capacities()
        assert_terms(w, [tt for tt in t if abs(tt[2]) > 0.5], "copy")
        w.mul_(2.0)
        assert_terms(v, [tt for tt in t if abs(tt[2]) > 0.5], "original untouched")
