"""Pins the oracle's ComplexF64 instantiation (oracle/spv_oracle.c, _w1c/_w2c).

The reference's default element type is ComplexF64
(type_SparsePauliVector.jl:138,167); its arithmetic differs from Float64 in
three places: the child (T(k-2) sinθ) c is a Complex * Complex product
(spv_kernels.jl:163), every filter measures |c| with Julia's hypot, and sums
and real scalings act per component.  Checks, each chosen so a plausible
restatement mistake fails it:
  * the hypot restatement is correctly rounded (decided in exact rational
    arithmetic) on random, scaled, nearly-tied and extreme inputs, and meets
    the IEEE special cases (in the band where Julia's correction step
    underflows, 2^-511 <= min < max < 2^-484, it is faithful, as Julia's is);
  * ComplexF64 with zero imaginary parts reproduces the pinned Float64 oracle
    bit for bit (keys, real parts) under every deterministic strategy;
  * linearity: without coefficient truncation, evolving a + ib gives
    evolve(a) + i evolve(b) exactly;
  * phase covariance: evolving i·c gives i·evolve(c) exactly under a
    coefficient threshold (|i c| = |c|, so the same terms survive);
  * the independent Dict-path restatement (tests/dictpath.py) with complex
    coefficients and a correctly rounded |c| agrees bit for bit;
  * dense conjugation and a state vector with complex operators.
"""
import math
import random

import numpy as np
import pytest

import dictpath as dp
import pyoracle as po
from pauliprop import (AdaptiveTruncation, CoeffTruncation, CompositeTruncation, MajoranaWeightTruncation,
                       NoTruncation, Pauli, WeightDampedTruncation, WeightTruncation, XWeightDampedTruncation,
                       XWeightTruncation, compile_filter)
from test_oracle import heisenberg_trotter


def _hypot_cases(rng, n):
    out = [(3.0, 4.0), (0.0, 0.0), (-0.0, 5.0), (1e-310, 0.0), (1.0, 1e-9), (1.0, 1e-8),
           (1e300, 1e300), (1.7e308, 1.7e308), (1e-300, 1e-300), (5e-324, 5e-324), (2.0 ** -540, 2.0 ** -540),
           (1.0, 1.0), (0.1, 0.2), (6.0, 8.0), (2.0 ** 511, 2.0 ** 511), (2.0 ** -511, 2.0 ** -513)]
    for _ in range(n):
        kind = rng.randrange(5)
        if kind == 0:
            x, y = rng.gauss(0, 1), rng.gauss(0, 1)
        elif kind == 1:  # widely scaled
            x = rng.gauss(0, 1) * 2.0 ** rng.randint(-1000, 1000)
            y = x * rng.gauss(0, 1) * 2.0 ** rng.randint(-40, 40)
        elif kind == 2:  # near the widely-varying cutoff sqrt(eps/2)
            x = rng.uniform(0.5, 2)
            y = x * 2.0 ** -26.5 * rng.uniform(0.99, 1.01)
        elif kind == 3:  # Pythagorean-like integers scaled (exact results)
            a, b = rng.randint(1, 2 ** 20), rng.randint(1, 2 ** 20)
            s = 2.0 ** rng.randint(-60, 60)
            x, y = (a * a - b * b) * s, 2 * a * b * s
        else:  # near overflow / underflow scaling thresholds
            e = rng.choice([510, 511, 512, -510, -511, -512, -1020, 1020])
            x, y = rng.uniform(0.5, 1) * 2.0 ** e, rng.uniform(0.5, 1) * 2.0 ** e
        out.append((x, y))
    return out


def _residuals_underflow(x, y):
    """Julia's algorithm leaves operands with min(|x|, |y|) >= sqrt(floatmin)
    = 2^-511 unscaled; if then max(|x|, |y|) < 2^-484 the correction's
    residuals (about ulp(max)^2) fall below the subnormal spacing 2^-1074 and
    the result is only faithful.  The restatement follows Julia's operations
    there too; the rounding claim is checked outside that band."""
    ax, ay = max(abs(x), abs(y)), min(abs(x), abs(y))
    return ay >= 2.0 ** -511 and ax < 2.0 ** -484


def test_hypot_restatement_is_correctly_rounded():
    rng = random.Random(5)
    band = 0
    for x, y in _hypot_cases(rng, 6000):
        got = po.julia_hypot(x, y)
        ref = dp.exact_hypot(x, y)
        if _residuals_underflow(x, y):
            band += 1
            assert abs(got - ref) <= math.ulp(ref), (x, y, got)
        else:
            assert got == ref, (x, y, got)
        assert got == po.julia_hypot(y, x) == po.julia_hypot(-x, y)
    assert band > 50  # the band was exercised


def test_hypot_special_values():
    assert po.julia_hypot(math.inf, math.nan) == math.inf
    assert po.julia_hypot(math.nan, -math.inf) == math.inf
    assert math.isnan(po.julia_hypot(math.nan, 1.0))
    assert po.julia_hypot(3.0, 4.0) == 5.0
    assert po.julia_hypot(0.0, -7.5) == 7.5
    assert po.julia_hypot(1.7976931348623157e308, 1.0) == 1.7976931348623157e308


STRATS = [NoTruncation(), CoeffTruncation(1e-4), WeightTruncation(3), XWeightTruncation(2),
          MajoranaWeightTruncation(6), WeightDampedTruncation(0.5, 1e-4),
          XWeightDampedTruncation(0.5, 1e-4), CompositeTruncation(WeightTruncation(4), CoeffTruncation(1e-6))]


def _evolve(N, terms, gens, angs, T, strat=None, window=1, adaptive=None):
    v = po.OracleSPV(N, terms, T=T)
    v.evolve(gens, angs, window=window, strict=compile_filter(strat) if strat is not None else None,
             adaptive=adaptive)
    assert v.check()
    return v.arrays()


def _random_terms(rng, N, n, T):
    seen, out = set(), []
    while len(out) < n:
        k = (rng.getrandbits(N), rng.getrandbits(N))
        if k in seen:
            continue
        seen.add(k)
        c = complex(rng.gauss(0, 1), rng.gauss(0, 1)) if T is complex else rng.gauss(0, 1)
        out.append((k[0], k[1], c))
    return out


def _random_gens(rng, N, L):
    return [(rng.getrandbits(N) & rng.getrandbits(N), rng.getrandbits(N) & rng.getrandbits(N)) for _ in range(L)], \
        [rng.uniform(-3, 3) for _ in range(L)]


@pytest.mark.parametrize("strat", STRATS, ids=lambda s: type(s).__name__)
@pytest.mark.parametrize("N", [8, 70])
def test_zero_imaginary_reproduces_float64(strat, N):
    rng = random.Random(N)
    terms = _random_terms(rng, N, 60, float)
    gens, angs = _random_gens(rng, N, 12)  # at most 60 * 2^12 terms without truncation
    rz, rx, rc = _evolve(N, terms, gens, angs, float, strat)
    cz, cx, cc = _evolve(N, [(z, x, complex(c, 0.0)) for z, x, c in terms], gens, angs, complex, strat)
    assert np.array_equal(rz, cz) and np.array_equal(rx, cx)
    assert np.array_equal(rc.view(np.uint64), cc.real.copy().view(np.uint64))
    assert np.all(cc.imag == 0)


@pytest.mark.parametrize("window", [1, 3])
@pytest.mark.parametrize("N", [8, 100])
def test_linearity_without_coefficient_truncation(window, N):
    """The term set depends only on the keys under a weight filter, and every
    operation is per component, so evolve(a + ib) = evolve(a) + i evolve(b)."""
    po.set_sort_mode(True)  # the same sort on all three runs: equal keys summed in one order
    try:
        rng = random.Random(N + window)
        base = _random_terms(rng, N, 50, float)
        im = [rng.gauss(0, 1) for _ in base]
        gens, angs = _random_gens(rng, N, 12)
        strat = WeightTruncation(N // 2)
        az, ax, ac = _evolve(N, base, gens, angs, float, strat, window)
        bz, bx, bc = _evolve(N, [(z, x, b) for (z, x, _), b in zip(base, im)], gens, angs, float, strat, window)
        cz, cx, cc = _evolve(N, [(z, x, complex(a, b)) for (z, x, a), b in zip(base, im)], gens, angs, complex,
                             strat, window)
    finally:
        po.set_sort_mode(False)
    assert np.array_equal(az, cz) and np.array_equal(bz, cz) and np.array_equal(ax, cx)
    assert np.array_equal(cc.real, ac) and np.array_equal(cc.imag, bc)


@pytest.mark.parametrize("N", [10, 128])
def test_phase_covariance_under_coefficient_threshold(N):
    rng = random.Random(7 * N)
    terms = _random_terms(rng, N, 60, complex)
    gens, angs = _random_gens(rng, N, 14)
    strat = CoeffTruncation(0.05)
    z1, x1, c1 = _evolve(N, terms, gens, angs, complex, strat)
    z2, x2, c2 = _evolve(N, [(z, x, 1j * c) for z, x, c in terms], gens, angs, complex, strat)
    assert len(c1) > 10
    assert np.array_equal(z1, z2) and np.array_equal(x1, x2)
    assert np.array_equal(c2.real, -c1.imag) and np.array_equal(c2.imag, c1.real)


@pytest.mark.parametrize("strat", STRATS + [AdaptiveTruncation(40, 1e-12)], ids=lambda s: type(s).__name__)
def test_sequence_bitexact_vs_dictpath_complex(strat):
    N = 8
    gens, angs = heisenberg_trotter(N)
    p2 = Pauli(N, X=[2], Y=[3])
    O0 = {(Pauli(N, Z=[4]).z, 0): complex(0.8, 0.6), (p2.z, p2.x): complex(0.3, -0.2)}
    ref = dp.dict_evolve(N, O0, gens, angs, strat)
    v = po.OracleSPV(N, [(k[0], k[1], c) for k, c in O0.items()], T=complex)
    if isinstance(strat, AdaptiveTruncation):
        v.evolve(gens, angs, adaptive=(strat.max_terms, strat.min_thresh))
    else:
        v.evolve(gens, angs, strict=compile_filter(strat))
    got = {(z, x): c for z, x, c in v.terms()}
    assert set(got) == set(ref)
    assert all(got[k] == ref[k] for k in ref)


def test_random_rotations_dense_complex():
    rng = random.Random(9)
    N = 4
    for _ in range(25):
        d = {(rng.getrandbits(N), rng.getrandbits(N)): complex(rng.gauss(0, 1), rng.gauss(0, 1)) for _ in range(12)}
        G = (rng.getrandbits(N), rng.getrandbits(N))
        theta = rng.uniform(-3, 3)
        v = po.OracleSPV(N, [(k[0], k[1], c) for k, c in d.items()], T=complex)
        v.evolve_single(G[0], G[1], theta)
        got = dp.sum_matrix(N, {(z, x): c for z, x, c in v.terms()})
        ref = dp.dense_heisenberg(N, dp.sum_matrix(N, d), G, theta)
        assert np.linalg.norm(got - ref) < 1e-12


def test_statevector_expectation_complex():
    """<0|U† O U|0> for a non-Hermitian O: the complex expectation."""
    rng = random.Random(11)
    N = 8
    gens, angs = _random_gens(rng, N, 12)
    O0 = {(rng.getrandbits(N), rng.getrandbits(N)): complex(rng.gauss(0, 1), rng.gauss(0, 1)) for _ in range(20)}
    v = po.OracleSPV(N, [(k[0], k[1], c) for k, c in O0.items()], T=complex)
    v.evolve(gens, angs)
    psi = np.zeros(2 ** N, dtype=complex)
    psi[0] = 1.0
    for g, a in reversed(list(zip(gens, angs))):
        psi = math.cos(a / 2) * psi - 1j * math.sin(a / 2) * dp.apply_pauli_state(N, psi, g)
    ref = sum(c * np.vdot(psi, dp.apply_pauli_state(N, psi, k)) for k, c in O0.items())
    assert abs(v.expectation(0) - ref) < 1e-12


def test_channels_per_component():
    """c *= λ^k on both components: the channel of a + ib equals channel(a) + i channel(b)."""
    rng = random.Random(13)
    N = 70
    base = _random_terms(rng, N, 80, float)
    im = [rng.gauss(0, 1) for _ in base]
    mask = rng.getrandbits(N)
    for kind, kw in [("depolarizing", dict(p=0.03)), ("dephasing", dict(p=0.07)), ("bit_flip", dict(p=0.02)),
                     ("bit_phase_flip", dict(p=0.05)), ("pauli", dict(pX=0.01, pY=0.02, pZ=0.03))]:
        a = po.OracleSPV(N, base)
        b = po.OracleSPV(N, [(z, x, q) for (z, x, _), q in zip(base, im)])
        c = po.OracleSPV(N, [(z, x, complex(p, q)) for (z, x, p), q in zip(base, im)], T=complex)
        for v in (a, b, c):
            v.channel(kind, mask=mask, **kw)
        assert np.array_equal(c.arrays()[2].real, a.arrays()[2])
        assert np.array_equal(c.arrays()[2].imag, b.arrays()[2])


def test_variance_complex_dense():
    """variance(O, ψ) = ||O|ψ>||^2 - <ψ|O|ψ>^2 (statistics.jl:19-21: real(e2 - e1 e1)) by state vectors."""
    rng = random.Random(17)
    for N in (3, 5):
        d = {(rng.getrandbits(N), rng.getrandbits(N)): complex(rng.gauss(0, 1), rng.gauss(0, 1)) for _ in range(10)}
        v = po.OracleSPV(N, [(k[0], k[1], c) for k, c in d.items()], T=complex)
        for ket in (0, 5 % 2 ** N):
            psi = np.zeros(2 ** N, dtype=complex)
            psi[ket] = 1.0
            opsi = sum(c * dp.apply_pauli_state(N, psi, k) for k, c in d.items())
            e1 = opsi[ket]
            ref = (np.vdot(opsi, opsi) - e1 * e1).real
            assert abs(v.variance(ket) - ref) < 1e-12
