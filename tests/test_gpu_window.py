"""GPU parity of the windowed path (window > 1), boundary strategies that do
not compile into one filter (AdaptiveTruncation, StochasticCoeffTruncation,
composites of them), EnergyCorrection and truncate!/stochastic_clip!.

The device sorts pending appends with a stable radix sort, so equal keys are
summed in append order.  The oracle in its stable-sort mode uses the same
order, and the comparison with it is bit for bit (keys, order, coefficient
bit patterns, counters, capacities).  Against the reference's own quicksort
order the coefficients agree to 1e-12 relative and the string sets agree
away from the cut (terms within 1e-12 relative of the threshold excluded).
"""
import math
import random

import numpy as np
import pytest

import pyoracle as po
from pauliprop import (AdaptiveTruncation, CoeffTruncation, CompositeTruncation, EnergyCorrection,
                       NoTruncation, Pauli, PauliBasis, PPCounters, SparsePauliVector,
                       StochasticCoeffTruncation, WeightTruncation, compile_filter)
from pauliprop import workloads as wl

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def assert_same(gpu, orc, what=""):
    gz, gx, gc = gpu.arrays()
    oz, ox, oc = orc.arrays()
    assert gz.shape == oz.shape, f"{what}: n gpu={len(gc)} oracle={len(oc)}"
    assert np.array_equal(gz, oz), f"{what}: z keys differ"
    assert np.array_equal(gx, ox), f"{what}: x keys differ"
    assert np.array_equal(gc.view(np.uint64), oc.view(np.uint64)), f"{what}: coefficients differ"


def heisenberg_gens(N, dt=0.05, nsteps=2, lo=1, span=8):
    terms = []
    hi = min(N, lo + span - 1)
    for i in range(lo, hi):
        terms += [(Pauli(N, X=[i, i + 1]), 1.0), (Pauli(N, Y=[i, i + 1]), 0.7), (Pauli(N, Z=[i, i + 1]), 1.1)]
    terms += [(Pauli(N, Z=[i]), 0.3) for i in range(lo, hi + 1)]
    gens, angs = [], []
    for _ in range(nsteps):
        half = [(p, dt * c) for p, c in terms]
        for p, a in half + half[::-1]:
            gens.append(p)
            angs.append(a)
    return gens, angs


class stable_oracle:
    def __enter__(self):
        po.set_sort_mode(True)

    def __exit__(self, *a):
        po.set_sort_mode(False)


def run_pair(N, O0, gens, angs, window, strat=NoTruncation(), local=NoTruncation(), adaptive=None,
             stable=True, energy_ket=None, **kw):
    g = SparsePauliVector(N, [((z, x), c) for z, x, c in O0], **kw)
    o = po.OracleSPV(N, O0, **kw)
    cnt = PPCounters()
    corr = EnergyCorrection(energy_ket) if energy_ket is not None else None
    gstrat = AdaptiveTruncation(*adaptive) if adaptive else strat
    extra = {"correction": corr} if corr is not None else {}
    g.evolve_(gens, angs, window=window, truncation=gstrat, local_truncation=local, counters=cnt, **extra)
    e = [energy_ket] if energy_ket is not None else None
    if stable:
        with stable_oracle():
            oc = o.evolve([(p.z, p.x) for p in gens], angs, window=window,
                          strict=None if adaptive else compile_filter(strat),
                          adaptive=adaptive, local=compile_filter(local), energy=e)
    else:
        oc = o.evolve([(p.z, p.x) for p in gens], angs, window=window,
                      strict=None if adaptive else compile_filter(strat),
                      adaptive=adaptive, local=compile_filter(local), energy=e)
    return g, o, cnt, oc, corr, e


def assert_counters(cnt, oc):
    for f in ("rotations", "terms_swept", "terms_created", "merge_in", "merge_out", "early_merges"):
        assert getattr(cnt, f) == getattr(oc, f), f


@pytest.mark.parametrize("N,window", [(8, 2), (8, 7), (12, 64), (70, 5), (121, 3), (40, 1000)])
def test_window_bitexact_vs_stable_oracle(N, window):
    lo = max(1, N - 9)
    gens, angs = heisenberg_gens(N, dt=0.2, nsteps=2, lo=lo, span=9)
    O0 = [(Pauli(N, Z=[lo + 4]).z, 0, 1.0), (Pauli(N, X=[lo + 1]).z, Pauli(N, X=[lo + 1]).x, -0.25)]
    log(f"window N={N} window={window} L={len(gens)}")
    g, o, cnt, oc, _, _ = run_pair(N, O0, gens, angs, window,
                                   strat=CompositeTruncation(CoeffTruncation(1e-7), WeightTruncation(7)),
                                   local=WeightTruncation(8))
    assert_same(g, o, f"N={N} window={window}")
    assert_counters(cnt, oc)
    assert g.capacities() == o.capacities()
    assert oc.merge_in > oc.merge_out


def test_window_early_merges_tight_append_capacity():
    N = 10
    gens, angs = heisenberg_gens(N, dt=0.3, nsteps=2)
    O0 = [(Pauli(N, Z=[5]).z, 0, 1.0)]
    g, o, cnt, oc, _, _ = run_pair(N, O0, gens, angs, 16, strat=CoeffTruncation(1e-6),
                                   append_factor=0.01, min_capacity=4)
    assert_same(g, o, "tight append capacity")
    assert_counters(cnt, oc)
    assert oc.early_merges > 0
    assert g.capacities() == o.capacities()


def test_window_vs_reference_quicksort_order():
    """Against the reference's quicksort tie order (which fixes the order in
    which duplicates are summed): every coefficient within 1e-12 of the
    operator's scale (max |c|), the operator within isapprox at rtol 1e-12
    (spv_ops.jl:88-107), and the string sets equal except for terms whose |c|
    lies within that tolerance of the cut.  Coefficients formed by
    cancellation carry absolute, not relative, rounding differences."""
    N = 12
    eps = 1e-6
    gens, angs = heisenberg_gens(N, dt=0.3, nsteps=3, span=12)
    O0 = [(Pauli(N, Z=[6]).z, 0, 1.0)]
    g, o, _, _, _, _ = run_pair(N, O0, gens, angs, 32, strat=CoeffTruncation(eps), stable=False,
                                append_factor=8)
    a = {(z, x): c for z, x, c in g.terms()}
    b = {(z, x): c for z, x, c in o.terms()}
    assert len(a) > 1000
    scale = max(abs(c) for c in b.values())
    tol = 1e-12 * scale
    diff2 = 0.0
    for k in set(a) | set(b):
        if k in a and k in b:
            assert abs(a[k] - b[k]) <= max(1e-12 * abs(b[k]), tol)
            diff2 += (a[k] - b[k]) ** 2
        else:
            c = a.get(k, b.get(k))
            assert abs(abs(c) - eps) <= tol
    norm = math.sqrt(sum(c * c for c in b.values()))
    assert math.sqrt(diff2) <= 1e-12 * norm


def test_window_independence_at_size_without_truncation():
    """Size-independent property at a larger size (~5e5 terms): without
    truncation every window gives the same operator (test_spv_evolution.jl:
    112-119): the same key set, coefficients within 1e-10 in L2."""
    N = 64
    gens, angs = heisenberg_gens(N, dt=0.3, nsteps=2, lo=28, span=10)
    obs = [((Pauli(N, Z=[32]).z, 0), 1.0)]
    res = []
    for window in (1, 9, len(gens)):
        v = SparsePauliVector(N, obs)
        cnt = PPCounters()
        v.evolve_(gens, angs, window=window, counters=cnt)
        log(f"heisenberg N=64 window={window}: n={len(v)} early merges={cnt.early_merges}")
        assert v.check()
        res.append(v.arrays())
    assert len(res[0][2]) > 500000
    for z, x, c in res[1:]:
        assert np.array_equal(z, res[0][0]) and np.array_equal(x, res[0][1])
        assert float(np.sqrt(np.sum((c - res[0][2]) ** 2))) < 1e-10


@pytest.mark.parametrize("window", [1, 6])
def test_adaptive_truncation_bitexact(window):
    N = 10
    gens, angs = heisenberg_gens(N, dt=0.3, nsteps=2)
    O0 = [(Pauli(N, Z=[5]).z, 0, 1.0)]
    g, o, cnt, oc, _, _ = run_pair(N, O0, gens, angs, window, adaptive=(200, 1e-9))
    assert_same(g, o, f"adaptive window={window}")
    assert_counters(cnt, oc)
    assert 0 < len(g) <= 200


def test_energy_correction_vs_oracle():
    N = 10
    ket = 0b1011001101
    gens, angs = heisenberg_gens(N, dt=0.3, nsteps=2)
    O0 = [(Pauli(N, Z=[5]).z, 0, 1.0)]
    for window, strat in ((1, CoeffTruncation(1e-3)), (5, CoeffTruncation(1e-3)),
                          (3, CompositeTruncation(WeightTruncation(4), CoeffTruncation(1e-4)))):
        g, o, _, _, corr, e = run_pair(N, O0, gens, angs, window, strat=strat, energy_ket=ket)
        assert_same(g, o, f"energy window={window}")
        # device sums in double-double, the oracle sequentially: agree within
        # the oracle's rounding (1e-14 absolute here, where sum |c| < 10)
        assert abs(e[1]) > 1e-6
        assert abs(corr.accumulated_energy - e[1]) < 1e-14, (window, corr.accumulated_energy, e[1])


def test_truncate_adaptive_and_composite():
    rng = random.Random(8)
    N = 70
    terms = {}
    while len(terms) < 20000:
        terms[(rng.getrandbits(N), rng.getrandbits(N))] = rng.gauss(0, 1) * 10 ** rng.uniform(-8, 0)
    O0 = [(z, x, c) for (z, x), c in terms.items()]
    for strat, steps in ((AdaptiveTruncation(5000, 1e-12), [("a", 5000, 1e-12)]),
                         (AdaptiveTruncation(50000, 1e-4), [("a", 50000, 1e-4)]),
                         (CompositeTruncation(WeightTruncation(60), AdaptiveTruncation(3000, 0.0)),
                          [("f", WeightTruncation(60)), ("a", 3000, 0.0)])):
        g = SparsePauliVector(N, [((z, x), c) for z, x, c in O0])
        o = po.OracleSPV(N, O0)
        g.truncate_(strat)
        for st in steps:
            if st[0] == "a":
                o.adaptive(st[1], st[2])
            else:
                o.compact(compile_filter(st[1]))
        assert_same(g, o, f"truncate {strat}")
    corr = EnergyCorrection(0)
    g = SparsePauliVector(N, [((z, x), c) for z, x, c in O0])
    before = g.expectation_value(0)
    g.truncate_(AdaptiveTruncation(100, 0.0), corr)
    assert abs(corr.accumulated_energy - (g.expectation_value(0) - before)) < 1e-15


def test_stochastic_clip_unbiased_and_deterministic():
    """stochastic_clip! (spv_evolve.jl:35-52): terms with |c| >= ε untouched;
    smaller ones kept with probability |c|/ε as ε·sign(c), so E[c'] = c.  The
    generator is counter-based: the same seed gives the same result.  The
    Julia RNG stream is not reproduced (parity of individual draws unpinned)."""
    N = 40
    eps = 1e-3
    rng = random.Random(4)
    keys = set()
    while len(keys) < 60000:
        keys.add((rng.getrandbits(N), rng.getrandbits(N)))
    keys = sorted(keys)
    fr = [0.05, 0.3, 0.75, 1.5]
    O0 = []
    for i, (z, x) in enumerate(keys):
        f = fr[i % 4]
        O0.append(((z, x), (1 if i % 3 else -1) * f * eps))
    g1 = SparsePauliVector(N, O0)
    g1.stochastic_clip_(eps, rng=1234)
    g2 = SparsePauliVector(N, O0)
    g2.stochastic_clip_(eps, rng=1234)
    t1 = g1.terms()
    assert t1 == g2.terms()
    got = {(z, x): c for z, x, c in t1}
    for i, ((z, x), c) in enumerate(O0):
        f = fr[i % 4]
        if f >= 1:
            assert got[(z, x)] == c
        elif (z, x) in got:
            assert got[(z, x)] == math.copysign(eps, c)
    for k, f in enumerate(fr[:3]):
        idx = [i for i in range(len(O0)) if i % 4 == k]
        kept = sum(1 for i in idx if O0[i][0] in got)
        p = kept / len(idx)
        sd = math.sqrt(f * (1 - f) / len(idx))
        assert abs(p - f) < 5 * sd, (f, p)
    g3 = SparsePauliVector(N, O0)
    g3.stochastic_clip_(eps, rng=99)
    assert g3.terms() != t1
    assert g1.check()


def test_stochastic_strategy_in_evolve_runs_and_keeps_order():
    N = 10
    gens, angs = heisenberg_gens(N, dt=0.3, nsteps=1)
    v = SparsePauliVector(N, [((Pauli(N, Z=[5]).z, 0), 1.0)])
    v.evolve_(gens, angs, window=4,
              truncation=CompositeTruncation(CoeffTruncation(1e-9), StochasticCoeffTruncation(1e-4, 5)))
    assert v.check()
    cs = [abs(c) for _, _, c in v.terms()]
    assert min(cs) >= 1e-4 * (1 - 1e-15)


def test_stochastic_sampling_truncation():
    """StochasticSamplingTruncation (spv_evolve.jl:69-99): exactly n_keep terms
    survive, in key order, each with its coefficient times one common factor
    that restores Σ|c|^2; large-|c| terms survive far more often than small
    ones (keys u^(1/|c|^2)); below n_keep nothing changes; same seed, same
    result.  Julia's RNG stream is not reproduced (individual draws unpinned)."""
    from pauliprop import StochasticSamplingTruncation
    rng = random.Random(5)
    N = 50
    d = {}
    while len(d) < 20000:
        d[(rng.getrandbits(N), rng.getrandbits(N))] = rng.choice([1.0, 1e-3]) * rng.choice([1, -1])
    items = sorted(d.items())
    v = SparsePauliVector(N, items)
    v.truncate_(StochasticSamplingTruncation(5000, 42))
    t = v.terms()
    assert len(t) == 5000 and v.check()
    norm0 = math.fsum(c * c for _, c in items)
    scale = math.sqrt(norm0 / math.fsum(d[(z, x)] ** 2 for z, x, _ in t))
    for z, x, c in t:
        assert abs(c - d[(z, x)] * scale) <= 1e-12 * abs(c)
    big = sum(1 for z, x, _ in t if abs(d[(z, x)]) == 1.0)
    nbig = sum(1 for _, c in items if abs(c) == 1.0)
    assert big == min(nbig, 5000) or big > 0.99 * min(nbig, 5000)
    v2 = SparsePauliVector(N, items)
    v2.truncate_(StochasticSamplingTruncation(5000, 42))
    assert v2.terms() == t
    v3 = SparsePauliVector(N, items)
    v3.truncate_(StochasticSamplingTruncation(30000, 1))
    assert len(v3) == 20000 and v3.terms() == SparsePauliVector(N, items).terms()
