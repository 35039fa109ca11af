"""Pins the CPU oracle (oracle/spv_oracle.c) before it is trusted as a checker.

Each check is chosen so that a plausible restatement mistake fails it:
  * known-answer tests restated from the reference's own tests
    (test/test_truncation.jl, test/test_phase1.jl, test/test_channels.jl);
  * brute force: dense-matrix conjugation and Kraus sums (N <= 4), state-vector
    expectation values (N <= 10);
  * an independent pure-Python Dict-path restatement (evolve.jl:32-46) with a
    site-by-site phase table, compared bit for bit at window = 1;
  * invariants: sortedness, window independence without truncation, early
    merges under tight capacity.
"""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

import dictpath as dp
import pyoracle as po
from pauliprop import (AdaptiveTruncation, CoeffTruncation, CompositeTruncation,
                       MajoranaWeightTruncation, NoTruncation, Pauli, PauliBasis,
                       WeightDampedTruncation, WeightTruncation, XWeightDampedTruncation,
                       XWeightTruncation, compile_filter, majorana_weight)


def PB(s):
    return PauliBasis.from_string(s)


def drops(strategy, s, c):
    p = PB(s)
    return po.should_drop(p.N, compile_filter(strategy), p.z, p.x, abs(c))


# --- known-answer tests: truncation criteria (test_truncation.jl) -------------

def test_coeff_truncation_kat():
    # test_truncation.jl:16-28
    keep = {"IIII": 1.0, "XIII": 0.5}
    gone = {"IXII": 1e-4, "IIXI": 1e-8}
    for s, c in keep.items():
        assert not drops(CoeffTruncation(1e-3), s, c)
    for s, c in gone.items():
        assert drops(CoeffTruncation(1e-3), s, c)
    # inclusive boundary: |c| <= thresh drops (clip.jl:15-17)
    assert drops(CoeffTruncation(0.5), "XIII", 0.5)
    assert not drops(CoeffTruncation(0.5), "XIII", math.nextafter(0.5, 1))


def test_weight_truncation_kat():
    # test_truncation.jl:30-44
    for s, kept in [("IIII", True), ("XIII", True), ("XXII", True), ("XXXI", False), ("XXXX", False)]:
        assert drops(WeightTruncation(2), s, 1.0) == (not kept)


def test_xweight_truncation_kat():
    # test_truncation.jl:46-67
    for s, kept in [("IIII", True), ("ZZZZ", True), ("XIII", True), ("YZII", True),
                    ("XYII", False), ("XYZX", False)]:
        assert drops(XWeightTruncation(1), s, 1.0) == (not kept)


def test_xweight_damped_kat():
    # test_truncation.jl:82-93: remove iff |c| exp(-alpha xw) <= thresh
    st = XWeightDampedTruncation(0.5, 0.05)
    table = [("ZZZZ", 1.0, True), ("XIII", 0.01, False), ("XYII", 0.5, True),
             ("XYXI", 0.1, False), ("XYXY", 1.0, True)]
    for s, c, kept in table:
        assert drops(st, s, c) == (not kept)


def test_weight_damped_kat():
    # test_truncation.jl:153-166
    st = WeightDampedTruncation(0.5, 0.05)
    table = [("IIII", 1.0, True), ("XIII", 0.01, False), ("XXII", 0.5, True),
             ("XXXI", 0.1, False), ("XXXX", 1.0, True)]
    for s, c, kept in table:
        assert drops(st, s, c) == (not kept)


def test_majorana_kats():
    # test_phase1.jl:172-185 and test_truncation.jl:124-138
    assert po.majorana_weight_bits(2, PB("II").z, PB("II").x) == 0
    assert po.majorana_weight_bits(2, PB("XI").z, PB("XI").x) == 1
    assert po.majorana_weight_bits(2, PB("IZ").z, PB("IZ").x) == 2
    assert po.majorana_weight_bits(2, PB("XX").z, PB("XX").x) == 2
    assert drops(MajoranaWeightTruncation(0), "XIII", 0.5)
    assert drops(MajoranaWeightTruncation(0), "XXXX", 0.1)
    assert not drops(MajoranaWeightTruncation(0), "IIII", 1.0)


@pytest.mark.parametrize("N", [4, 40, 64, 65, 100, 128])
def test_majorana_word_level_matches_site_scan(N):
    # the branchless word-level formula (spv_kernels.jl:51-62) equals the
    # per-site control-flag scan it was derived from (clip.jl:78-87)
    rng = random.Random(N)
    for _ in range(200):
        z = rng.getrandbits(N)
        x = rng.getrandbits(N)
        expect = dp.majorana_weight((z, x))
        assert po.majorana_weight_bits(N, z, x) == expect
        assert majorana_weight(PauliBasis(N, z, x)) == expect


def test_composite_equals_sequential():
    # test_truncation.jl:250-275 (composite = sequential application)
    terms = {"IIII": 1.0, "XIII": 0.5, "XXXI": 0.3, "IIXI": 1e-8}
    st = CompositeTruncation(CoeffTruncation(1e-3), WeightTruncation(2))
    kept = {s for s, c in terms.items() if not drops(st, s, c)}
    assert kept == {"IIII", "XIII"}


# --- channel closed forms, Kraus brute force, λ^k --------------------------------

def _oracle_from(N, d):
    return po.OracleSPV(N, [(k[0], k[1], c) for k, c in d.items()])


def test_channel_closed_forms():
    # test_channels.jl:120-180
    N = 3
    p = 0.1
    ld = 1 - 4 * p / 3
    v = _oracle_from(N, {(PB("XIZ").z, PB("XIZ").x): 1.0, (PB("YYY").z, PB("YYY").x): 2.0,
                         (0, 0): 7.0})
    v.channel("depolarizing", p)
    got = {(z, x): c for z, x, c in v.terms()}
    assert got[(PB("XIZ").z, PB("XIZ").x)] == pytest.approx(ld ** 2, rel=1e-15)
    assert got[(PB("YYY").z, PB("YYY").x)] == pytest.approx(2 * ld ** 3, rel=1e-15)
    assert got[(0, 0)] == 7.0
    for kind, p, s, k in [("dephasing", 0.2, "XIZ", 1), ("dephasing", 0.2, "YYZ", 2),
                          ("dephasing", 0.2, "ZIZ", 0), ("bit_flip", 0.3, "XIZ", 1),
                          ("bit_flip", 0.3, "YYZ", 3), ("bit_flip", 0.3, "XIX", 0),
                          ("bit_phase_flip", 0.25, "XIZ", 2), ("bit_phase_flip", 0.25, "YYY", 0),
                          ("bit_phase_flip", 0.25, "XYZ", 2)]:
        q = PB(s)
        v = _oracle_from(N, {(q.z, q.x): 1.0})
        v.channel(kind, p)
        assert v.terms()[0][2] == pytest.approx((1 - 2 * p) ** k, rel=1e-15)
    pX, pY, pZ = 0.05, 0.10, 0.15
    lX, lY, lZ = 1 - 2 * (pY + pZ), 1 - 2 * (pX + pZ), 1 - 2 * (pX + pY)
    for s, expect in [("XYZ", lX * lY * lZ), ("XIX", lX * lX)]:
        q = PB(s)
        v = _oracle_from(N, {(q.z, q.x): 1.0})
        v.channel("pauli", pX=pX, pY=pY, pZ=pZ)
        assert v.terms()[0][2] == pytest.approx(expect, rel=1e-15)


def _kraus(N, pI, pX, pY, pZ, qubits):
    mats = (dp._I2, dp._X, dp._Y, dp._Z)
    probs = (pI, pX, pY, pZ)
    ks = []
    act = [q for q in range(1, N + 1) if q in qubits]
    for idx in range(4 ** len(act)):
        coef = 1.0
        m = np.ones((1, 1), dtype=complex)
        a = 0
        for i in range(1, N + 1):
            if i in qubits:
                ch = (idx // 4 ** a) % 4
                coef *= math.sqrt(probs[ch])
                m = np.kron(mats[ch], m)
                a += 1
            else:
                m = np.kron(dp._I2, m)
        ks.append(coef * m)
    return ks


@pytest.mark.parametrize("N", [1, 2, 3])
def test_channels_kraus_bruteforce(N):
    # test_channels.jl:185-241
    rng = random.Random(3 + N)
    cases = [("depolarizing", lambda p: (1 - p, p / 3, p / 3, p / 3)),
             ("dephasing", lambda p: (1 - p, 0.0, 0.0, p)),
             ("bit_flip", lambda p: (1 - p, p, 0.0, 0.0)),
             ("bit_phase_flip", lambda p: (1 - p, 0.0, p, 0.0))]
    for kind, wts in cases:
        for p in (0.0, 0.07, 0.3, 0.55, 0.91, 1.0):
            d = {}
            for _ in range(2 ** N):
                d[(rng.getrandbits(N), rng.getrandbits(N))] = rng.gauss(0, 1)
            ks = _kraus(N, *wts(p), qubits=set(range(1, N + 1)))
            M = dp.sum_matrix(N, d)
            ref = sum(K.conj().T @ M @ K for K in ks)
            v = _oracle_from(N, d)
            v.channel(kind, p)
            got = dp.sum_matrix(N, {(z, x): c for z, x, c in v.terms()})
            assert np.linalg.norm(got - ref) < 1e-11
    for (pX, pY, pZ) in ((0.1, 0.2, 0.3), (0.4, 0.0, 0.4), (0.0, 0.5, 0.0)):
        for qubits in ([1], [N], list(range(1, N + 1))):
            d = {(rng.getrandbits(N), rng.getrandbits(N)): rng.gauss(0, 1) for _ in range(2 ** N)}
            ks = _kraus(N, 1 - pX - pY - pZ, pX, pY, pZ, qubits=set(qubits))
            M = dp.sum_matrix(N, d)
            ref = sum(K.conj().T @ M @ K for K in ks)
            mask = sum(1 << (q - 1) for q in qubits)
            v = _oracle_from(N, d)
            v.channel("pauli", pX=pX, pY=pY, pZ=pZ, mask=mask)
            got = dp.sum_matrix(N, {(z, x): c for z, x, c in v.terms()})
            assert np.linalg.norm(got - ref) < 1e-11


def test_julia_powi_is_correctly_rounded_on_channel_lambdas():
    # λ^k from the compensated power-by-squaring vs the exact rational power,
    # rounded once: must agree to the last bit on these channel values
    # (Julia special-cases n == 3 as the literal x*x*x, base/math.jl pow_body)
    for lam in (1 - 4 * 1e-3 / 3, 1 - 2 * 5e-4, 0.8, 1 - 4 * 0.1 / 3, 0.4, 0.9999):
        exact = Fraction(lam)
        for k in range(0, 130):
            got = po.julia_powi(lam, k)
            if k == 3:
                assert got == lam * lam * lam
                continue
            ref = float(exact ** k)  # Fraction -> float rounds to nearest
            assert got == ref, (lam, k, got, ref)


# --- rotations vs dense conjugation ---------------------------------------------

def test_single_rotation_all_pairs_dense_n2():
    # every (G, P) pair on 2 qubits, several angles: O -> e^{iθG/2} O e^{-iθG/2}
    N = 2
    for gz in range(4):
        for gx in range(4):
            for pz in range(4):
                for px in range(4):
                    for theta in (0.3, math.pi / 2, -1.234):
                        v = po.OracleSPV(N, [(pz, px, 1.0)])
                        v.evolve_single(gz, gx, theta)
                        got = dp.sum_matrix(N, {(z, x): c for z, x, c in v.terms()})
                        ref = dp.dense_heisenberg(N, dp.pauli_matrix(N, (pz, px)), (gz, gx), theta)
                        assert np.linalg.norm(got - ref) < 1e-13


def test_random_rotations_dense_n4():
    rng = random.Random(31)
    N = 4
    for _ in range(30):
        d = {(rng.getrandbits(N), rng.getrandbits(N)): rng.gauss(0, 1) for _ in range(12)}
        G = (rng.getrandbits(N), rng.getrandbits(N))
        theta = rng.uniform(-3, 3)
        v = _oracle_from(N, d)
        v.evolve_single(G[0], G[1], theta)
        assert v.check()
        got = dp.sum_matrix(N, {(z, x): c for z, x, c in v.terms()})
        ref = dp.dense_heisenberg(N, dp.sum_matrix(N, d), G, theta)
        assert np.linalg.norm(got - ref) < 1e-12


# --- bit-exact parity with the independent Dict-path restatement ----------------

def heisenberg_trotter(N, dt=0.05, nsteps=2):
    """A fixed 2nd-order-like generator sequence on an open XX+YY+ZZ+Z chain."""
    terms = []
    for i in range(1, N):
        terms.append((Pauli(N, X=[i, i + 1]), 1.0))
        terms.append((Pauli(N, Y=[i, i + 1]), 0.7))
        terms.append((Pauli(N, Z=[i, i + 1]), 1.1))
    for i in range(1, N + 1):
        terms.append((Pauli(N, Z=[i]), 0.3))
    gens, angs = [], []
    for _ in range(nsteps):
        half = [(p, 2 * (dt / 2) * c) for p, c in terms]
        seq = half + list(reversed(half))
        for p, a in seq:
            gens.append((p.z, p.x))
            angs.append(a)
    return gens, angs


STRATS = [NoTruncation(), CoeffTruncation(1e-4), WeightTruncation(3), XWeightTruncation(2),
          MajoranaWeightTruncation(6), WeightDampedTruncation(0.5, 1e-4),
          XWeightDampedTruncation(0.5, 1e-4), CompositeTruncation(WeightTruncation(4), CoeffTruncation(1e-6))]


@pytest.mark.parametrize("strat", STRATS, ids=lambda s: type(s).__name__)
def test_sequence_bitexact_vs_dictpath(strat):
    # test_spv_evolution.jl:54-75, tightened from 1e-12 to bit equality
    N = 8
    gens, angs = heisenberg_trotter(N)
    O0 = {(Pauli(N, Z=[4]).z, 0): 1.0}
    ref = dp.dict_evolve(N, O0, gens, angs, strat)
    v = _oracle_from(N, O0)
    v.evolve(gens, angs, strict=compile_filter(strat))
    assert v.check()
    got = {(z, x): c for z, x, c in v.terms()}
    assert set(got) == set(ref)
    assert all(got[k] == ref[k] for k in ref)


def test_adaptive_truncation_vs_dictpath():
    N = 8
    gens, angs = heisenberg_trotter(N)
    O0 = {(Pauli(N, Z=[4]).z, 0): 1.0}
    st = AdaptiveTruncation(50, 1e-12)
    ref = dp.dict_evolve(N, O0, gens, angs, st)
    v = _oracle_from(N, O0)
    v.evolve(gens, angs, adaptive=(50, 1e-12))
    got = {(z, x): c for z, x, c in v.terms()}
    assert set(got) == set(ref)
    assert max(abs(got[k] - ref[k]) for k in ref) == 0.0


def test_single_rotation_bitexact_vs_dictpath_random():
    # test_spv_evolution.jl:33-44 (θ list from there), random sums of 40 terms
    rng = random.Random(31)
    for N in (8, 64, 65, 128):
        for theta in (0.0, 0.3, math.pi / 2, math.pi, -1.234):
            d = {(rng.getrandbits(N), rng.getrandbits(N)): rng.gauss(0, 1) for _ in range(40)}
            G = (rng.getrandbits(N), rng.getrandbits(N))
            ref = dict(d)
            dp.dict_evolve_single(N, ref, G, theta)
            v = _oracle_from(N, d)
            v.evolve_single(*G, theta)
            assert v.check()
            got = {(z, x): c for z, x, c in v.terms()}
            assert set(got) == set(ref)
            assert all(got[k] == ref[k] for k in ref)


# --- invariants -----------------------------------------------------------------

def test_window_independence_without_truncation():
    # test_spv_evolution.jl:112-119
    N = 6
    gens, angs = heisenberg_trotter(N)
    O0 = {(Pauli(N, Z=[3]).z, 0): 1.0}
    v1 = _oracle_from(N, O0)
    v1.evolve(gens, angs, window=1)
    r1 = {(z, x): c for z, x, c in v1.terms()}
    for w in (3, 7, len(gens)):
        vw = _oracle_from(N, O0)
        vw.evolve(gens, angs, window=w)
        rw = {(z, x): c for z, x, c in vw.terms()}
        keys = set(r1) | set(rw)
        assert math.sqrt(sum((r1.get(k, 0.0) - rw.get(k, 0.0)) ** 2 for k in keys)) < 1e-10


def test_windowed_weight_truncation_equals_dict_every_w():
    # test_spv_evolution.jl:121-138
    N = 6
    gens, angs = heisenberg_trotter(N)
    O0 = {(Pauli(N, Z=[3]).z, 0): 1.0}
    w = 5
    ref = dict(O0)
    for i, (g, a) in enumerate(zip(gens, angs)):
        dp.dict_evolve_single(N, ref, g, a)
        if (i + 1) % w == 0 or i + 1 == len(gens):
            dp.dict_truncate(ref, WeightTruncation(3))
    v = po.OracleSPV(N, [(k[0], k[1], c) for k, c in O0.items()], capacity_factor=5000, append_factor=10)
    cnt = v.evolve(gens, angs, window=w, strict=compile_filter(WeightTruncation(3)))
    assert cnt.early_merges == 0
    got = {(z, x): c for z, x, c in v.terms()}
    assert set(got) == set(ref)
    assert max(abs(got[k] - ref[k]) for k in ref) < 1e-12


def test_early_merge_under_tight_capacity():
    # test_spv_evolution.jl:154-167
    N = 8
    gens, angs = heisenberg_trotter(N)
    O0 = [(Pauli(N, Z=[4]).z, 0, 1.0)]
    v = po.OracleSPV(N, O0, append_factor=0.01, min_capacity=4)
    cnt = v.evolve(gens, angs, window=10)
    assert cnt.early_merges > 0
    ref = po.OracleSPV(N, O0)
    ref.evolve(gens, angs, window=1)
    a = {(z, x): c for z, x, c in v.terms()}
    b = {(z, x): c for z, x, c in ref.terms()}
    keys = set(a) | set(b)
    assert math.sqrt(sum((a.get(k, 0.0) - b.get(k, 0.0)) ** 2 for k in keys)) < 1e-10


def test_statevector_expectation_end_to_end():
    # ⟨0|U^† O U|0⟩ by brute-force state vector, no truncation
    N = 10
    rng = random.Random(5)
    gens, angs = [], []
    for _ in range(3):
        for j in range(1, N):
            p = Pauli(N, Z=[j, j + 1])
            gens.append((p.z, p.x))
            angs.append(rng.uniform(0, math.pi))
        for j in range(1, N + 1):
            p = Pauli(N, X=[j])
            gens.append((p.z, p.x))
            angs.append(rng.uniform(0, math.pi))
    O0 = {(Pauli(N, Z=[5]).z, 0): 1.0}
    v = _oracle_from(N, O0)
    v.evolve(gens, angs)
    ref = dp.statevector_expectation(N, O0, gens, angs)
    assert abs(v.expectation(0) - ref) < 1e-12
    # nonzero ket: sign (-1)^popcount(z & k) (expectation_value.jl:8-11)
    ket = 0b1011001
    got = v.expectation(ket)
    exp = sum(c * (-1) ** bin(z & ket).count("1") for z, x, c in v.terms() if x == 0)
    assert abs(got - exp) < 1e-12


# --- windowed mode: boundary sort order and EnergyCorrection ---------------------

def _rand_circuit(N, L, seed, lo=1):
    """L random 1- and 2-site generators on sites lo..min(lo+11, N)."""
    rng = random.Random(seed)
    hi = min(lo + 11, N)
    gens, angs = [], []
    for _ in range(L):
        q = rng.randrange(lo, hi + 1)
        kind = rng.randrange(3)
        if kind == 0 and q < hi:
            p = Pauli(N, Z=[q, q + 1])
        elif kind == 1 or q == hi:
            p = Pauli(N, X=[q])
        else:
            p = Pauli(N, Y=[q], X=[q + 1])
        gens.append((p.z, p.x))
        angs.append(rng.uniform(-math.pi, math.pi))
    return gens, angs


def _pydrop(f):
    return lambda k, a: po.should_drop(_N_CUR[0], f, k[0], k[1], a)


_N_CUR = [0]


@pytest.mark.parametrize("N,window", [(10, 3), (70, 5), (12, 64)])
def test_window_stable_mode_bitexact_vs_python_restatement(N, window):
    """Stable-sort mode of the oracle against the independent Python restatement
    of the windowed driver: same string set, same coefficients bit for bit
    (duplicates summed in append order), same EnergyCorrection accumulation."""
    _N_CUR[0] = N
    lo = max(1, N - 11)
    gens, angs = _rand_circuit(N, 100, seed=N + window, lo=lo)
    fs = compile_filter(CompositeTruncation(CoeffTruncation(1e-7), WeightTruncation(6)))
    fl = compile_filter(WeightTruncation(7))
    terms = [((Pauli(N, Z=[lo + 5]).z, 0), 1.0), ((Pauli(N, X=[lo]).z, Pauli(N, X=[lo]).x), 0.5)]
    ket = random.Random(1).getrandbits(N)
    ref, ref_e = dp.window_evolve(N, list(terms), gens, angs, window, _pydrop(fs), _pydrop(fl), energy_ket=ket)
    v = po.OracleSPV(N, [(k[0], k[1], c) for k, c in terms], capacity_factor=1e5, append_factor=64)
    po.set_sort_mode(True)
    try:
        e = [ket]
        cnt = v.evolve(gens, angs, window=window, strict=fs, local=fl, energy=e)
    finally:
        po.set_sort_mode(False)
    assert cnt.early_merges == 0
    got = [((z, x), c) for z, x, c in v.terms()]
    assert [k for k, _ in got] == [k for k, _ in ref]
    assert all(a == b for (_, a), (_, b) in zip(got, ref))
    assert e[1] == ref_e
    assert len(ref) > 30


def test_window_quicksort_and_stable_agree_to_ulp():
    """The reference's quicksort and the stable sort differ only in the order
    equal keys are summed: coefficients agree to 1e-12 relative and the string
    sets agree away from the cut."""
    N = 12
    gens, angs = _rand_circuit(N, 120, seed=3)
    eps = 1e-6
    fs = compile_filter(CoeffTruncation(eps))
    res = []
    for stable in (False, True):
        v = po.OracleSPV(N, [(Pauli(N, Z=[6]).z, 0, 1.0)], append_factor=8)
        po.set_sort_mode(stable)
        try:
            v.evolve(gens, angs, window=16, strict=fs)
        finally:
            po.set_sort_mode(False)
        res.append({(z, x): c for z, x, c in v.terms()})
    a, b = res
    for k in set(a) | set(b):
        if k in a and k in b:
            assert abs(a[k] - b[k]) <= 1e-12 * abs(a[k])
        else:
            c = a.get(k, b.get(k))
            assert abs(abs(c) - eps) <= 1e-12 * eps
    assert len(a) > 100


def test_energy_correction_window1_matches_dict_path():
    """EnergyCorrection at window = 1 against evolve!; truncate!(O, s, corr) per
    rotation on the Dict path (truncation.jl:330-337): accumulated energy
    agrees to rounding; without truncation it is zero up to rounding."""
    N = 8
    gens, angs = heisenberg_trotter(N, dt=0.3, nsteps=3)
    O0 = {(Pauli(N, Z=[4]).z, 0): 1.0}
    strat = CoeffTruncation(1e-2)
    ket = 0b10110010
    ref = dict(O0)
    acc = 0.0

    def ev(O):
        return sum(c * (1 - 2 * (bin(z & ket).count("1") & 1)) for (z, x), c in O.items() if x == 0)

    for g, a in zip(gens, angs):
        dp.dict_evolve_single(N, ref, g, a)
        before = ev(ref)
        dp.dict_truncate(ref, strat)
        acc += ev(ref) - before
    v = _oracle_from(N, O0)
    e = [ket]
    v.evolve(gens, angs, strict=compile_filter(strat), energy=e)
    assert abs(acc) > 1e-4
    assert abs(e[1] - acc) < 1e-13
    assert abs(v.expectation(ket) - ev(ref)) < 1e-13
    v0 = _oracle_from(N, O0)
    e0 = [ket]
    v0.evolve(gens, angs, energy=e0)
    assert abs(e0[1]) < 1e-14
