"""The oracle's variance(O, ψ) (statistics.jl:6-22) and its
EnergyVarianceCorrection boundary (spv_evolve.jl:246-270,
truncation.jl:273-310), pinned by:

* the reference's known answers: Var(ZI, |00>) = 0, Var(XI, |00>) = 1
  (test_phase1.jl:131-150), and the truncate! case of
  test_truncation.jl:305-323 (IIII 1.0, ZIII 0.5, XIII 0.3, IIXI 0.01 clipped
  at 0.1: variance 0.0901 -> 0.09, by hand);
* brute force: O|k> built from the single-qubit actions X|b> = |1-b>,
  Z|b> = (-1)^b |b>, Y|b> = i (-1)^b |1-b> on 2^N amplitudes, and
  Var = ||O|k>||^2 - <k|O|k>^2, for random operators on N <= 5;
* the EnergyVarianceCorrection boundary: the final term set equals the run
  without a correction bit for bit; its accumulated energy equals
  EnergyCorrection's; with one boundary its accumulated variance equals
  Var(truncated) - Var(untruncated) computed densely; nothing accumulates
  without truncation.
"""
import math
import random

import numpy as np
import pytest

import pyoracle as po
from pauliprop import CoeffTruncation, CompositeTruncation, WeightTruncation, compile_filter


def dense_apply(N, terms, k):
    """O|k> for O = sum c P over (z, x, c), as a dict basis -> amplitude."""
    out = {}
    for z, x, c in terms:
        amp = complex(c)
        b = k
        for q in range(N):
            zq, xq, bq = (z >> q) & 1, (x >> q) & 1, (k >> q) & 1
            if zq and xq:       # Y
                amp *= 1j * (-1) ** bq
            elif zq:            # Z
                amp *= (-1) ** bq
        b = k ^ x
        out[b] = out.get(b, 0) + amp
    return out


def dense_var(N, terms, k):
    v = dense_apply(N, terms, k)
    e2 = sum(abs(a) ** 2 for a in v.values())
    e1 = v.get(k, 0)
    return e2 - (e1 * e1).real


def test_reference_known_answers():
    o = po.OracleSPV(2, [(0b01, 0, 1.0)])  # Z on qubit 1 (ZI)
    assert abs(o.variance(0)) < 1e-12
    o = po.OracleSPV(2, [(0, 0b01, 1.0)])  # XI
    assert o.variance(0) == 1.0
    # test_truncation.jl:305-323
    N = 4
    terms = [(0, 0, 1.0), (0b0001, 0, 0.5), (0, 0b0001, 0.3), (0, 0b0100, 0.01)]
    o = po.OracleSPV(N, terms)
    v_before = o.variance(0)
    assert abs(v_before - 0.0901) < 1e-15
    o.compact(compile_filter(CoeffTruncation(0.1)))
    assert abs(o.variance(0) - 0.09) < 1e-15


@pytest.mark.parametrize("N", [1, 2, 3, 5, 70])
def test_variance_matches_brute_force(N):
    rng = random.Random(N)
    for trial in range(20):
        nt = rng.randint(1, min(40, 4 ** min(N, 5)))
        seen, terms = set(), []
        while len(terms) < nt:
            z, x = rng.getrandbits(N), rng.getrandbits(N)
            if N > 5:  # keep the dense action cheap: support on a few qubits
                z &= 0b11111
                x &= 0b11111
            if (z, x) in seen:
                continue
            seen.add((z, x))
            terms.append((z, x, rng.gauss(0, 1)))
        k = rng.getrandbits(N)
        o = po.OracleSPV(N, terms)
        got = o.variance(k)
        ref = dense_var(N, terms, k)
        scale = sum(abs(c) for _, _, c in terms) ** 2
        assert abs(got - ref) <= 1e-13 * scale, (N, trial, got, ref)
        assert got >= -1e-13 * scale


def test_variance_invariants():
    rng = random.Random(4)
    N = 6
    diag = [(rng.getrandbits(N), 0, rng.gauss(0, 1)) for _ in range(30)]
    o = po.OracleSPV(N, list({(z, x): (z, x, c) for z, x, c in diag}.values()))
    assert abs(o.variance(rng.getrandbits(N))) < 1e-12  # basis states are eigenstates
    terms = list({(z, x): (z, x, rng.gauss(0, 1)) for z, x in
                  [(rng.getrandbits(N), rng.getrandbits(N)) for _ in range(50)]}.values())
    a = po.OracleSPV(N, terms).variance(5)
    b = po.OracleSPV(N, [(z, x, 2.0 * c) for z, x, c in terms]).variance(5)
    assert abs(b - 4.0 * a) <= 1e-12 * max(1.0, abs(a))


def chain_gens(N, reps=2):
    gens = []
    for _ in range(reps):
        gens += [(0b11 << (i - 1), 0) for i in range(1, N)]   # ZZ bonds
        gens += [(0, 1 << (i - 1)) for i in range(1, N + 1)]  # X sites
    return gens


@pytest.mark.parametrize("window", [1, 4, 1000])
def test_variance_correction_boundary(window):
    N = 8
    ket = 0b10110010
    gens = chain_gens(N)
    rng = random.Random(window)
    angs = [rng.uniform(-1.2, 1.2) for _ in gens]
    O0 = [(1 << 3, 0, 1.0), (0, 1 << 5, 0.4)]
    st = CompositeTruncation(CoeffTruncation(2e-3), WeightTruncation(5))
    f = compile_filter(st)
    big = {"min_capacity": 1 << 16}  # no early merge: a window of 1000 is one boundary
    plain = po.OracleSPV(N, O0, **big)
    c_plain = plain.evolve(gens, angs, window=window, strict=f)
    ec = po.OracleSPV(N, O0, **big)
    e_ec = [ket]
    ec.evolve(gens, angs, window=window, strict=f, energy=e_ec)
    evc = po.OracleSPV(N, O0, **big)
    e_v, var = [ket], []
    c_v = evc.evolve(gens, angs, window=window, strict=f, energy=e_v, variance=var)
    a, b = plain.arrays(), evc.arrays()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))
    assert (c_v.rotations, c_v.terms_swept, c_v.terms_created) == (c_plain.rotations, c_plain.terms_swept,
                                                                   c_plain.terms_created)
    assert c_v.merge_out >= c_plain.merge_out  # the boundary merges with NOFILTER, then compacts
    assert abs(e_v[1] - e_ec[1]) <= 1e-14
    if window < len(gens):  # (the single final boundary happens to drop only off-diagonal terms)
        assert abs(e_v[1]) > 1e-6
    assert abs(var[0]) > 1e-6
    if window >= len(gens):  # one boundary: accumulated = Var(truncated) - Var(untruncated), densely
        assert c_v.early_merges == 0
        full = po.OracleSPV(N, O0, **big)
        full.evolve(gens, angs, window=window)
        fz, fx, fc = full.arrays()
        tz, tx, tc = evc.arrays()
        ref = (dense_var(N, list(zip(tz[:, 0].tolist(), tx[:, 0].tolist(), tc.tolist())), ket)
               - dense_var(N, list(zip(fz[:, 0].tolist(), fx[:, 0].tolist(), fc.tolist())), ket))
        assert abs(var[0] - ref) <= 1e-12, (var[0], ref)
    # no truncation: nothing accumulates
    nt = po.OracleSPV(N, O0)
    e0, v0 = [ket], []
    nt.evolve(gens, angs, window=window, energy=e0, variance=v0)
    assert e0[1] == 0.0 and v0[0] == 0.0
