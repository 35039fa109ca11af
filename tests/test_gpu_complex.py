"""GPU parity for ComplexF64 coefficients (the reference's default element
type, type_SparsePauliVector.jl:138,167) against the oracle's ComplexF64
instantiation, bit for bit: the surviving key set, key order and both
components of every coefficient as bit patterns (signed zeros included).

The device computes |c| with its own restatement of Julia's hypot; a
single-term vector's norm(v, 1) returns that value unrounded, which is
compared with the oracle's restatement on hard cases.  Sums (expectation,
inner product, norms) are double-double on the device: compared with the
exactly rounded sum (math.fsum) to a few ulps.
"""
import math
import random

import numpy as np
import pytest

import dictpath as dp
import pyoracle as po
from pauliprop import (AdaptiveTruncation, ArgumentError, CoeffTruncation, CompositeTruncation,
                       EnergyCorrection, EnergyVarianceCorrection, MajoranaWeightTruncation, NoTruncation, Pauli,
                       PauliBasis, PPCounters, SparsePauliVector, WeightDampedTruncation, WeightTruncation,
                       XWeightDampedTruncation, XWeightTruncation, compile_filter, inner_product, isapprox)
from pauliprop import workloads as wl
from test_gpu_parity import heisenberg_gens, random_terms, spread_support
from test_gpu_window import stable_oracle

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def assert_same(gpu, orc, what=""):
    gz, gx, gc = gpu.arrays()
    oz, ox, oc = orc.arrays()
    assert gc.dtype == np.complex128 and oc.dtype == np.complex128
    assert gz.shape == oz.shape, f"{what}: n gpu={len(gc)} oracle={len(oc)}"
    assert np.array_equal(gz, oz), f"{what}: z keys differ"
    assert np.array_equal(gx, ox), f"{what}: x keys differ"
    assert np.array_equal(gc.view(np.uint64), oc.view(np.uint64)), f"{what}: coefficients differ"


def cterms(N, n, rng, support=None):
    return [(z, x, complex(c, rng.gauss(0, 1))) for z, x, c in random_terms(N, n, rng, support)]


def pair(N, terms, **kw):
    g = SparsePauliVector(N, [((z, x), c) for z, x, c in terms], T=complex, **kw)
    o = po.OracleSPV(N, terms, T=complex, **kw)
    return g, o


def test_device_hypot_bitexact():
    """|c| on the device equals the oracle's hypot bit for bit (random, scaled,
    near the sqrt(eps/2) cutoff, the rescaling thresholds, subnormals)."""
    from test_complex_oracle import _hypot_cases
    rng = random.Random(21)
    cases = _hypot_cases(rng, 1500)
    v = SparsePauliVector(4, [((1, 0), 1.0)], T=complex)
    p = PauliBasis(4, 1, 0)
    bad = []
    for x, y in cases:
        if math.isinf(po.julia_hypot(x, y)):
            continue  # a single inf does not pass through the double-double sum unchanged
        v[p] = complex(x, y)
        got = v.norm(1)
        ref = po.julia_hypot(x, y)
        if not (got == ref or (math.isnan(got) and math.isnan(ref))):
            bad.append((x, y, got, ref))
    assert not bad, bad[:5]


@pytest.mark.parametrize("N", [6, 16, 64, 65, 128])
def test_single_rotation_random_generators(N):
    rng = random.Random(2000 + N)
    for trial in range(4):
        bits = min(N, 7)
        sup = spread_support(N, bits, rng)
        n = min(rng.choice([1, 64, 65, 2049, 5000]), 2 ** (2 * bits) - 1)
        g, o = pair(N, cterms(N, n, rng, sup))
        for k in range(4):
            gz, gx = (rng.getrandbits(N) & sup, rng.getrandbits(N) & sup) if k % 2 == 0 else \
                (rng.getrandbits(N), rng.getrandbits(N))
            theta = rng.choice([0.0, 0.3, math.pi / 2, math.pi, -1.234, rng.uniform(-3, 3)])
            log(f"complex N={N} trial={trial} n={n} step={k} θ={theta}")
            g.evolve_(PauliBasis(N, gz, gx), theta)
            o.evolve_single(gz, gx, theta)
            assert_same(g, o, f"N={N} trial={trial} step={k}")
            g.check()


STRATS = [NoTruncation(), CoeffTruncation(1e-4), WeightTruncation(3), XWeightTruncation(2),
          MajoranaWeightTruncation(6), WeightDampedTruncation(0.5, 1e-4), XWeightDampedTruncation(0.5, 1e-4),
          CompositeTruncation(WeightTruncation(4), CoeffTruncation(1e-6))]


@pytest.mark.parametrize("strat", STRATS, ids=lambda s: type(s).__name__)
def test_sequence_strategies(strat):
    for N in (8, 70):
        gens, angs = heisenberg_gens(8)
        gens = [PauliBasis(N, p.z, p.x) for p in gens]
        p2 = Pauli(N, X=[2], Y=[3])
        O0 = [(Pauli(N, Z=[4]).z, 0, complex(0.8, 0.6)), (p2.z, p2.x, complex(0.3, -0.2))]
        g, o = pair(N, O0)
        g.evolve_(gens, angs, truncation=strat)
        o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(strat))
        assert_same(g, o, f"{type(strat).__name__} N={N}")


def test_local_truncation_and_counters():
    N = 10
    gens, angs = heisenberg_gens(N)
    O0 = [(Pauli(N, Z=[5]).z, 0, complex(0.6, -0.8))]
    g, o = pair(N, O0)
    cnt = PPCounters()
    st, loc = CoeffTruncation(1e-5), WeightTruncation(4)
    g.evolve_(gens, angs, truncation=st, local_truncation=loc, counters=cnt)
    oc = o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(st), local=compile_filter(loc))
    assert_same(g, o, "local truncation")
    for f in ("rotations", "terms_swept", "terms_created", "merge_out", "early_merges"):
        assert getattr(cnt, f) == getattr(oc, f), f


def test_kicked_ising_chain16_complex_matches_float64():
    """The 16-qubit kicked-Ising workload as ComplexF64 with zero imaginary
    parts: the same term set and real parts as the Float64 run, bit for bit,
    and the ComplexF64 oracle's early-merge cadence and capacities."""
    w = wl.kicked_ising_16()
    gens, angs = w.flat()
    obs = [(p.z, p.x, c) for p, c in w.observable]
    gr = SparsePauliVector(w.N, [((z, x), c) for z, x, c in obs])
    gr.evolve_(gens, angs, truncation=w.truncation)
    g, o = pair(w.N, [(z, x, complex(c, 0.0)) for z, x, c in obs])
    g.evolve_(gens, angs, truncation=w.truncation)
    oc = o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(w.truncation))
    assert_same(g, o, "kicked ising 16 complex")
    assert oc.early_merges > 0 and g.capacities() == o.capacities()
    rz, rx, rc = gr.arrays()
    cz, cx, cc = g.arrays()
    assert np.array_equal(rz, cz) and np.array_equal(rx, cx)
    assert np.array_equal(rc.view(np.uint64), cc.real.copy().view(np.uint64)) and np.all(cc.imag == 0)
    ev = g.expectation_value(0)
    assert isinstance(ev, complex) and ev.imag == 0 and ev.real == gr.expectation_value(0)


def test_lattice121_first_layers_complex_phase():
    """W = 2 keys with a coefficient threshold: the 121-qubit lattice observable
    times e^{iφ}, two layers with dephasing, bit for bit."""
    w = wl.lattice121(wl.theta_h_values()[6], n_layers=2)
    ph = complex(math.cos(0.7), math.sin(0.7))
    obs = [(p.z, p.x, c * ph) for p, c in w.observable]
    g, o = pair(w.N, obs)
    for li in range(2):
        L = w.layers[li]
        g.evolve_(L.generators, L.angles, truncation=w.truncation)
        o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=compile_filter(w.truncation))
        kind, p = w.channel
        getattr(g, kind + "_channel_")(p)
        o.channel(kind, p)
        log(f"lattice121 complex layer {li}: n={len(g)}")
        assert_same(g, o, f"lattice121 complex layer {li}")
    ev, oe = g.expectation_value(0), o.expectation(0)
    assert abs(ev - oe) <= 1e-12


@pytest.mark.parametrize("kind", ["depolarizing", "dephasing", "bit_flip", "bit_phase_flip", "pauli"])
def test_channels_bitexact(kind):
    rng = random.Random(31)
    for N in (5, 64, 100):
        bits = min(N, 8)
        g, o = pair(N, cterms(N, min(3000, 4 ** bits // 2), rng, spread_support(N, bits, rng)))
        qs = sorted(rng.sample(range(1, N + 1), 3))
        mask = sum(1 << (q - 1) for q in qs)
        if kind == "pauli":
            g.pauli_channel_(0.05, 0.1, 0.15, qubits=qs)
            o.channel("pauli", pX=0.05, pY=0.1, pZ=0.15, mask=mask)
        else:
            getattr(g, kind + "_channel_")(0.13, qubits=qs)
            o.channel(kind, 0.13, mask=mask)
        assert_same(g, o, f"{kind} N={N}")


def test_clips_bitexact():
    rng = random.Random(32)
    for N in (8, 90):
        terms = [(z, x, c * 10 ** rng.uniform(-6, 0)) for z, x, c in cterms(N, 5000, rng, spread_support(N, 8, rng))]
        for strat in STRATS[1:]:
            g, o = pair(N, terms)
            g.truncate_(strat)
            o.compact(compile_filter(strat))
            assert_same(g, o, f"{type(strat).__name__} N={N}")


@pytest.mark.parametrize("N,window", [(8, 2), (12, 7), (70, 5), (121, 3)])
def test_window_bitexact_vs_stable_oracle(N, window):
    lo = max(1, N - 9)
    gens, angs = heisenberg_gens(N if N <= 9 else 9, dt=0.2)
    gens = [PauliBasis(N, p.z << (lo - 1), p.x << (lo - 1)) for p in gens]
    p1 = Pauli(N, X=[lo + 1])
    O0 = [(Pauli(N, Z=[lo + 4]).z, 0, complex(1.0, 0.5)), (p1.z, p1.x, complex(-0.25, 0.75))]
    strat = CompositeTruncation(CoeffTruncation(1e-7), WeightTruncation(7))
    g, o = pair(N, O0)
    cnt = PPCounters()
    g.evolve_(gens, angs, window=window, truncation=strat, local_truncation=WeightTruncation(8), counters=cnt)
    with stable_oracle():
        oc = o.evolve([(p.z, p.x) for p in gens], angs, window=window, strict=compile_filter(strat),
                      local=compile_filter(WeightTruncation(8)))
    assert_same(g, o, f"window N={N} w={window}")
    assert cnt.merge_out == oc.merge_out and cnt.early_merges == oc.early_merges
    assert g.capacities() == o.capacities()


@pytest.mark.parametrize("window", [1, 6])
def test_adaptive_truncation_bitexact(window):
    """The radix select over |c| = hypot bit patterns picks the reference's
    coeffs[end - max_terms] (spv_evolve.jl:103-114)."""
    N = 10
    gens, angs = heisenberg_gens(N, dt=0.3)
    O0 = [(Pauli(N, Z=[5]).z, 0, complex(0.6, 0.8))]
    g, o = pair(N, O0)
    g.evolve_(gens, angs, window=window, truncation=AdaptiveTruncation(200, 1e-9))
    with stable_oracle():
        o.evolve([(p.z, p.x) for p in gens], angs, window=window, adaptive=(200, 1e-9))
    assert_same(g, o, f"adaptive window={window}")
    assert 0 < len(g) <= 200


def test_energy_and_variance_corrections():
    """EnergyCorrection / EnergyVarianceCorrection on ComplexF64: real(⟨ψ|O|ψ⟩)
    and real(e2 - e1 e1) (truncation.jl:288-293), against the oracle."""
    N = 10
    ket = 0b1011001101
    gens, angs = heisenberg_gens(N, dt=0.3)
    p2 = Pauli(N, Z=[5], X=[6])
    O0 = [(Pauli(N, Z=[5]).z, 0, complex(1.0, 0.3)), (p2.z, p2.x, complex(-0.4, 0.2))]
    for window, corr_t in ((1, EnergyCorrection), (5, EnergyCorrection), (1, EnergyVarianceCorrection),
                           (4, EnergyVarianceCorrection)):
        g, o = pair(N, O0)
        corr = corr_t(ket)
        strat = CoeffTruncation(1e-3)
        g.evolve_(gens, angs, window=window, truncation=strat, correction=corr)
        e = [ket]
        var = [] if corr_t is EnergyVarianceCorrection else None
        with stable_oracle():
            o.evolve([(p.z, p.x) for p in gens], angs, window=window, strict=compile_filter(strat), energy=e,
                     variance=var)
        assert_same(g, o, f"{corr_t.__name__} window={window}")
        assert abs(e[1]) > 1e-6
        assert abs(corr.accumulated_energy - e[1]) < 1e-13
        if var is not None:
            assert abs(corr.accumulated_variance - var[0]) < 1e-13
    v = SparsePauliVector(N, [((z, x), c) for z, x, c in O0], T=complex)
    ov = po.OracleSPV(N, O0, T=complex)
    assert abs(v.variance(ket) - ov.variance(ket)) < 1e-13


def test_ops_complex():
    rng = random.Random(41)
    N = 70
    d = {}
    while len(d) < 20000:
        d[(rng.getrandbits(N), rng.getrandbits(N))] = complex(rng.gauss(0, 1), rng.gauss(0, 1)) * 10 ** rng.uniform(-6, 0)
    v = SparsePauliVector(N, list(d.items()), T=complex)
    keys = list(d)
    vals, found = v.lookup(keys)
    assert found.all() and all(vals[i] == d[k] for i, k in enumerate(keys))
    p = PauliBasis(N, *keys[3])
    assert v[p] == d[keys[3]] and isinstance(v[p], complex)
    # norms: Σ abs2, Σ hypot, max hypot, Σ hypot^3
    cs = [d[k] for k in sorted(d)]
    abs_ = [dp.exact_hypot(c.real, c.imag) for c in cs]
    assert abs(v.norm(2) - math.sqrt(math.fsum(c.real * c.real + c.imag * c.imag for c in cs))) <= \
        4 * np.spacing(v.norm(2))
    assert abs(v.norm(1) - math.fsum(abs_)) <= 4 * np.spacing(v.norm(1))
    assert v.norm(math.inf) == max(abs_)
    ref3 = math.fsum(a ** 3 for a in abs_) ** (1 / 3)
    assert abs(v.norm(3) - ref3) <= 1e-13 * ref3
    # expectation: Σ over diagonal terms of ±c, both components exactly rounded
    ket = rng.getrandbits(N)
    diag = [(1 - 2 * (bin(z & ket).count("1") & 1)) * c for (z, x), c in d.items() if x == 0]
    d[(5, 0)] = complex(0.25, -0.5)
    v[PauliBasis(N, 5, 0)] = complex(0.25, -0.5)
    diag.append((1 - 2 * (bin(5 & ket).count("1") & 1)) * complex(0.25, -0.5))
    ev = v.expectation_value(ket)
    assert abs(ev.real - math.fsum(c.real for c in diag)) <= 2 * np.spacing(1.0)
    assert abs(ev.imag - math.fsum(c.imag for c in diag)) <= 2 * np.spacing(1.0)
    # inner product Σ conj(a) b and isapprox
    b = {k: c + complex(rng.gauss(0, 1e-3), 0) for k, c in d.items() if rng.random() < 0.5}
    vb = SparsePauliVector(N, list(b.items()), T=complex)
    ip = [d[k].conjugate() * b[k] for k in b]
    got = inner_product(v, vb)
    assert abs(got.real - math.fsum(c.real for c in ip)) <= 16 * np.spacing(abs(got))
    assert abs(got.imag - math.fsum(c.imag for c in ip)) <= 16 * np.spacing(abs(got))
    dist = math.sqrt(math.fsum(abs(d.get(k, 0) - b.get(k, 0)) ** 2 for k in set(d) | set(b)))
    assert isapprox(v, vb, atol=dist * (1 + 1e-9)) and not isapprox(v, vb, atol=dist * (1 - 1e-9))
    # mul!: a real scalar per component, a complex scalar as Complex * Complex
    w = v.copy().mul_(0.5)
    assert np.array_equal(w.arrays()[2], v.arrays()[2] * 0.5)
    s = complex(0.3, -1.7)
    w = v.copy().mul_(s)
    c0 = v.arrays()[2]
    ref = np.array([complex(c.real * s.real - c.imag * s.imag, c.real * s.imag + c.imag * s.real) for c in c0])
    assert np.array_equal(w.arrays()[2].view(np.uint64), ref.view(np.uint64))
    # setindex!/delete!/sum!
    u = SparsePauliVector(N, list(b.items()), T=complex)
    u.sum_(v)
    su = dict(b)
    for k, c in d.items():
        su[k] = su[k] + c if k in su else c
    uz, ux, uc = u.arrays()
    assert len(uc) == len(su)
    got = {(int(z[0]) | (int(z[1]) << 64) if len(z) > 1 else int(z[0]),
            int(x[0]) | (int(x[1]) << 64) if len(x) > 1 else int(x[0])): c for z, x, c in zip(uz, ux, uc)}
    assert all(got[k] == su[k] for k in su)
    del u[PauliBasis(N, *keys[0])]
    assert len(u) == len(su) - 1 and not u.haskey(PauliBasis(N, *keys[0]))


def test_type_errors():
    N = 8
    vr = SparsePauliVector(N, [((1, 0), 1.0)])
    vc = SparsePauliVector(N, [((1, 0), 1j)], T=complex)
    with pytest.raises(ArgumentError):
        vr.mul_(1j)
    with pytest.raises(ArgumentError):
        vr.sum_(vc)
    with pytest.raises(ArgumentError):
        inner_product(vr, vc)
    with pytest.raises(ValueError):
        SparsePauliVector(N, [((1, 0), 1j)])  # imaginary part into Float64 storage
    vr.mul_(complex(2.0, 0.0))
    assert vr[PauliBasis(N, 1, 0)] == 2.0
