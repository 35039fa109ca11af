"""variance(O, ψ) on the device (statistics.jl:6-22) and the
EnergyVarianceCorrection boundary (spv_evolve.jl:246-270,
truncation.jl:273-310), against the oracle.

* variance: the device groups terms by x (the sorted order puts a group's
  terms side by side), sums each group's phased coefficients in z order and
  adds the groups' |a_x|^2 in double-double; the oracle sums the groups in the
  same order and the squares sequentially.  Agreement within 1e-13 of
  (sum |c|)^2.  The reference's known answers Var(ZI, |00>) = 0 and
  Var(XI, |00>) = 1 (test_phase1.jl:131-150) hold exactly.
* evolve! with EnergyVarianceCorrection: the term set equals the stable-sort
  oracle's bit for bit (keys, order, coefficient bits, counters); both
  accumulators agree within 1e-13.
* truncate! with EnergyVarianceCorrection on test_truncation.jl:305-323's
  operator (clip at 0.1: variance 0.0901 -> 0.09, energy unchanged).
"""
import random

import numpy as np
import pytest

import pyoracle as po
from pauliprop import (AdaptiveTruncation, CoeffTruncation, CompositeTruncation, EnergyCorrection,
                       EnergyVarianceCorrection, Pauli, PPCounters, SparsePauliVector, WeightTruncation,
                       compile_filter, variance)

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


class stable_oracle:
    def __enter__(self):
        po.set_sort_mode(True)

    def __exit__(self, *a):
        po.set_sort_mode(False)


def assert_same(gpu, orc, what=""):
    gz, gx, gc = gpu.arrays()
    oz, ox, oc = orc.arrays()
    assert gz.shape == oz.shape, f"{what}: n gpu={len(gc)} oracle={len(oc)}"
    assert np.array_equal(gz, oz), f"{what}: z keys differ"
    assert np.array_equal(gx, ox), f"{what}: x keys differ"
    assert np.array_equal(gc.view(np.uint64), oc.view(np.uint64)), f"{what}: coefficients differ"


def random_terms(N, n, xbits, seed):
    """n distinct strings whose x parts come from few values (large groups)."""
    rng = random.Random(seed)
    xs = [rng.getrandbits(N) for _ in range(xbits)] + [0]
    terms = {}
    while len(terms) < n:
        terms[(rng.getrandbits(N), rng.choice(xs))] = rng.gauss(0, 1)
    return [(z, x, c) for (z, x), c in terms.items()]


def test_variance_known_answers():
    v = SparsePauliVector(2, [((Pauli(2, Z=[1]).z, 0), 1.0)])
    assert abs(v.variance(0)) < 1e-15
    v = SparsePauliVector(2, [((0, Pauli(2, X=[1]).x), 1.0)])
    assert v.variance(0) == 1.0
    assert variance(v, 0) == 1.0
    assert SparsePauliVector(3, []).variance(5) == 0.0


@pytest.mark.parametrize("N,n,groups", [(6, 200, 7), (40, 60000, 40), (40, 30000, 30000), (40, 200000, 0),
                                        (121, 50000, 100),
                                        (121, 1, 1)])
def test_variance_vs_oracle(N, n, groups):
    O0 = random_terms(N, min(n, 4 ** N), groups, seed=N + n)
    log(f"variance N={N} n={len(O0)} groups<={groups + 1}")
    g = SparsePauliVector(N, [((z, x), c) for z, x, c in O0])
    o = po.OracleSPV(N, O0)
    rng = random.Random(n)
    scale = sum(abs(c) for _, _, c in O0) ** 2
    for _ in range(4):
        k = rng.getrandbits(N)
        got, ref = g.variance(k), o.variance(k)
        assert abs(got - ref) <= 1e-13 * scale, (k, got, ref)


def chain_gens(N, reps=2):
    gens = []
    for _ in range(reps):
        gens += [Pauli(N, Z=[i, i + 1]) for i in range(1, N)]
        gens += [Pauli(N, X=[i]) for i in range(1, N + 1)]
    return gens


@pytest.mark.parametrize("N,window,adaptive", [(8, 1, None), (8, 4, None), (12, 7, (30, 1e-6)),
                                               (70, 3, None), (121, 5, (25, 0.0))])
def test_evolve_variance_correction_vs_oracle(N, window, adaptive):
    lo = max(1, N - 10)
    gens = [p for p in chain_gens(N) if (p.z | p.x) >> (lo - 1)]
    rng = random.Random(N * 10 + window)
    angs = [rng.uniform(-1.2, 1.2) for _ in gens]
    O0 = [(Pauli(N, Z=[lo + 3]).z, 0, 1.0), (0, Pauli(N, X=[lo + 5]).x, 0.4)]
    ket = rng.getrandbits(N)
    strat = CompositeTruncation(CoeffTruncation(2e-3), WeightTruncation(5))
    log(f"evolve EVC N={N} window={window} adaptive={adaptive} L={len(gens)}")
    g = SparsePauliVector(N, [((z, x), c) for z, x, c in O0])
    corr = EnergyVarianceCorrection(ket)
    cnt = PPCounters()
    g.evolve_(gens, angs, window=window, truncation=AdaptiveTruncation(*adaptive) if adaptive else strat,
              correction=corr, counters=cnt)
    o = po.OracleSPV(N, O0)
    e, var = [ket], []
    with stable_oracle():
        oc = o.evolve([(p.z, p.x) for p in gens], angs, window=window,
                      strict=None if adaptive else compile_filter(strat), adaptive=adaptive, energy=e,
                      variance=var)
    assert_same(g, o, f"EVC N={N} window={window}")
    for f in ("rotations", "terms_swept", "terms_created", "merge_in", "merge_out", "early_merges"):
        assert getattr(cnt, f) == getattr(oc, f), f
    assert abs(var[0]) > 1e-8
    assert abs(corr.accumulated_variance - var[0]) < 1e-13, (corr.accumulated_variance, var[0])
    assert abs(corr.accumulated_energy - e[1]) < 1e-13, (corr.accumulated_energy, e[1])
    # the energy part equals EnergyCorrection's (the same truncations are measured)
    g2 = SparsePauliVector(N, [((z, x), c) for z, x, c in O0])
    c2 = EnergyCorrection(ket)
    g2.evolve_(gens, angs, window=window, truncation=AdaptiveTruncation(*adaptive) if adaptive else strat,
               correction=c2)
    assert_same(g2, o, "EnergyCorrection run")
    assert abs(c2.accumulated_energy - corr.accumulated_energy) < 1e-13


def test_truncate_variance_known_answer():
    N = 4
    terms = [((0, 0), 1.0), ((Pauli(N, Z=[1]).z, 0), 0.5), ((0, Pauli(N, X=[1]).x), 0.3),
             ((0, Pauli(N, X=[3]).x), 0.01)]
    g = SparsePauliVector(N, terms)
    assert abs(g.variance(0) - 0.0901) < 1e-15
    corr = EnergyVarianceCorrection(0)
    g.truncate_(CoeffTruncation(0.1), corr)
    assert len(g) == 3
    assert abs(g.variance(0) - 0.09) < 1e-15
    assert abs(corr.accumulated_variance - (-0.0001)) < 1e-15
    assert corr.accumulated_energy == 0.0
    # accumulation continues from the stored values
    g.truncate_(CoeffTruncation(0.4), corr)
    assert abs(corr.accumulated_variance - (-0.0001 - 0.09)) < 1e-15
