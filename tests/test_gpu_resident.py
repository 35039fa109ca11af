"""GPU parity of the resident layer launches (one kernel launch per evolve
call, terms held in LDS across the rotations of the call).

Every case forces the resident mode and compares with the CPU oracle bit for
bit (keys, order, coefficient bit patterns) and counter for counter.  Small
chunk capacities (cap_limit) force multi-workgroup clusters, chunk
rebalances inside a launch, and vectors that outgrow their cluster and
finish the call on the per-step kernels.
"""
import math
import random

import numpy as np
import pytest

import pyoracle as po
from pauliprop import (CoeffTruncation, CompositeTruncation, MajoranaWeightTruncation, NoTruncation, Pauli,
                       PauliBasis, PPCounters, SparsePauliBatch, SparsePauliVector, WeightDampedTruncation,
                       WeightTruncation, XWeightTruncation, compile_filter)
from pauliprop import workloads as wl

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def same_arrays(a, b):
    gz, gx, gc = a
    oz, ox, oc = b
    return (gz.shape == oz.shape and np.array_equal(gz, oz) and np.array_equal(gx, ox)
            and np.array_equal(gc.view(np.uint64), oc.view(np.uint64)))


def heisenberg_gens(N, dt=0.05, nsteps=2):
    terms = []
    for i in range(1, N):
        terms += [(Pauli(N, X=[i, i + 1]), 1.0), (Pauli(N, Y=[i, i + 1]), 0.7), (Pauli(N, Z=[i, i + 1]), 1.1)]
    terms += [(Pauli(N, Z=[i]), 0.3) for i in range(1, N + 1)]
    gens, angs = [], []
    for _ in range(nsteps):
        half = [(p, dt * c) for p, c in terms]
        for p, a in half + half[::-1]:
            gens.append(p)
            angs.append(a)
    return gens, angs


STRATS = [NoTruncation(), CoeffTruncation(1e-4), WeightTruncation(3), XWeightTruncation(2),
          MajoranaWeightTruncation(6), WeightDampedTruncation(0.5, 1e-4),
          CompositeTruncation(WeightTruncation(4), CoeffTruncation(1e-6))]


@pytest.mark.parametrize("cap", [0, 48, 200])
@pytest.mark.parametrize("strat", STRATS, ids=lambda s: type(s).__name__)
def test_resident_single_vector_sequences(strat, cap):
    """Heisenberg Trotter sequences (XX, YY, ZZ, Z: 1, 2 and 4 mask levels) on
    one vector, N = 8 (one key word) and 70 (two)."""
    for N in (8, 70):
        gens, angs = heisenberg_gens(8, nsteps=3)
        gens = [PauliBasis(N, p.z, p.x) for p in gens]
        O0 = [(Pauli(N, Z=[4]).z, 0, 1.0)]
        g = SparsePauliVector(N, [((z, x), c) for z, x, c in O0]).set_resident(2, cap)
        o = po.OracleSPV(N, O0)
        cnt = PPCounters()
        g.evolve_(gens, angs, truncation=strat, counters=cnt)
        oc = o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(strat))
        log(f"{type(strat).__name__} N={N} cap={cap}: n={len(g)}")
        assert same_arrays(g.arrays(), o.arrays()), f"{type(strat).__name__} N={N} cap={cap}"
        assert (cnt.rotations, cnt.terms_swept, cnt.terms_created, cnt.merge_in, cnt.merge_out,
                cnt.early_merges) == (oc.rotations, oc.terms_swept, oc.terms_created, oc.merge_in, oc.merge_out,
                                      oc.early_merges)
        assert g.capacities() == o.capacities()


@pytest.mark.parametrize("cap", [0, 48, 200])
@pytest.mark.parametrize("strat", [NoTruncation(), CoeffTruncation(1e-4), WeightTruncation(3),
                                   CompositeTruncation(WeightTruncation(4), CoeffTruncation(1e-6))],
                         ids=lambda s: type(s).__name__)
def test_resident_complex_single_vector_sequences(strat, cap):
    """ComplexF64 coefficients in the resident launch (16-byte coefficients
    in LDS; 3072-term chunks at two key words): the Heisenberg sequences with
    a complex observable, against the ComplexF64 oracle bit for bit (both
    parts) and counter for counter."""
    for N in (8, 70):
        gens, angs = heisenberg_gens(8, nsteps=3)
        gens = [PauliBasis(N, p.z, p.x) for p in gens]
        p2 = Pauli(N, X=[2], Y=[3])
        O0 = [(Pauli(N, Z=[4]).z, 0, complex(0.8, 0.6)), (p2.z, p2.x, complex(0.3, -0.2))]
        g = SparsePauliVector(N, [((z, x), c) for z, x, c in O0], T=complex).set_resident(2, cap)
        o = po.OracleSPV(N, O0, T=complex)
        cnt = PPCounters()
        g.evolve_(gens, angs, truncation=strat, counters=cnt)
        oc = o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(strat))
        log(f"complex {type(strat).__name__} N={N} cap={cap}: n={len(g)}")
        assert g.arrays()[2].dtype == np.complex128
        assert same_arrays(g.arrays(), o.arrays()), f"{type(strat).__name__} N={N} cap={cap}"
        assert (cnt.rotations, cnt.terms_swept, cnt.terms_created, cnt.merge_in, cnt.merge_out,
                cnt.early_merges) == (oc.rotations, oc.terms_swept, oc.terms_created, oc.merge_in, oc.merge_out,
                                      oc.early_merges)
        assert g.capacities() == o.capacities()


@pytest.mark.parametrize("cap", [0, 256])
def test_resident_complex_lattice121(cap):
    """The 121-qubit lattice (two key words) as one ComplexF64 vector with a
    complex observable, three layers with dephasing: multi-workgroup
    clusters at cap 256, against the ComplexF64 oracle."""
    w = wl.lattice121(0.3, n_layers=3)
    terms = [(p.z, p.x, complex(c, 0.5 * c)) for p, c in w.observable]
    g = SparsePauliVector(w.N, [((z, x), c) for z, x, c in terms], T=complex).set_resident(2, cap)
    o = po.OracleSPV(w.N, terms, T=complex)
    f = compile_filter(w.truncation)
    for li in range(3):
        L = w.layers[li]
        g.evolve_(L.generators, L.angles, truncation=w.truncation)
        assert w.channel[0] == "dephasing"
        g.dephasing_channel_(w.channel[1])
        o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f)
        o.channel(*w.channel)
        assert same_arrays(g.arrays(), o.arrays()), f"layer {li}"
        log(f"complex lattice cap={cap} layer {li}: n={len(g)}")


@pytest.mark.parametrize("cap", [0, 64, 300])
def test_resident_dense_collisions(cap):
    """N = 6 (4096 keys): most children land on live keys; random generators."""
    rng = random.Random(99 + cap)
    N = 6
    seen, terms = set(), []
    while len(terms) < 1500:
        z, x = rng.getrandbits(N), rng.getrandbits(N)
        if (z, x) not in seen:
            seen.add((z, x))
            terms.append((z, x, rng.gauss(0, 1)))
    g = SparsePauliVector(N, [((z, x), c) for z, x, c in terms]).set_resident(2, cap)
    o = po.OracleSPV(N, terms)
    gens, angs = [], []
    for _ in range(40):
        gz, gx = rng.getrandbits(N), rng.getrandbits(N)
        if bin(gz | gx).count("1") > 2:  # at most 4 mask levels
            gz &= 0b11
            gx &= 0b11
        gens.append(PauliBasis(N, gz, gx))
        angs.append(rng.uniform(-3, 3))
    g.evolve_(gens, angs, truncation=CoeffTruncation(1e-3))
    o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(CoeffTruncation(1e-3)))
    assert same_arrays(g.arrays(), o.arrays())


def test_resident_kicked_ising_chain16():
    w = wl.kicked_ising_16()
    gens, angs = w.flat()
    obs = [(p.z, p.x, c) for p, c in w.observable]
    g = SparsePauliVector(w.N, [((z, x), c) for z, x, c in obs]).set_resident(2, 0)
    o = po.OracleSPV(w.N, obs)
    for li, L in enumerate(w.layers):
        g.evolve_(L.generators, L.angles, truncation=w.truncation)
        o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=compile_filter(w.truncation))
        assert same_arrays(g.arrays(), o.arrays()), f"layer {li}"
    assert g.capacities() == o.capacities()
    assert abs(g.expectation_value(0) - o.expectation(0)) <= 1e-12


def lattice_batch(cap, n_layers, thetas, mode=2):
    ws = [wl.lattice121(th, n_layers=n_layers) for th in thetas]
    bt = SparsePauliBatch([w.N for w in ws], [[((p.z, p.x), c) for p, c in w.observable] for w in ws])
    bt.set_resident(mode, cap)
    orcs = [po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable]) for w in ws]
    f = compile_filter(ws[0].truncation)
    for li in range(n_layers):
        cnts = bt.evolve_([w.layers[li].generators for w in ws], [w.layers[li].angles for w in ws],
                          truncation=ws[0].truncation, counters=True)
        bt.channel_(*ws[0].channel)
        for b, (w, o) in enumerate(zip(ws, orcs)):
            L = w.layers[li]
            oc = o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f)
            o.channel(*w.channel)
            c = cnts[b]
            assert (c.rotations, c.terms_swept, c.terms_created, c.merge_in, c.merge_out, c.early_merges) == (
                oc.rotations, oc.terms_swept, oc.terms_created, oc.merge_in, oc.merge_out, oc.early_merges), \
                f"layer {li} theta {b} counters"
        for b, o in enumerate(orcs):
            assert same_arrays(bt.arrays(b), o.arrays()), f"layer {li} theta {b}"
        log(f"cap={cap} layer {li}: sizes {bt.sizes()} resident {bt.resident_stats()} work {bt.resident_work()}")
    evs = bt.expectation_values()
    for e, o in zip(evs, orcs):
        assert abs(e - o.expectation(0)) <= 1e-12
    return bt


def test_resident_lattice121_first_layers():
    """The 121-qubit sweep's first 4 layers for 8 of its 16 θ_h: clusters of
    several workgroups, vectors that outgrow them mid-layer (resumed from
    that rotation with more workgroups, or on the per-step kernels),
    dephasing between layers."""
    ths = wl.theta_h_values()[::2]
    bt = lattice_batch(0, 4, ths)
    launches, fallbacks, aborts = bt.resident_stats()
    assert launches >= 4 and aborts == 0
    assert bt.resident_work()[3] > 0  # chunks rebalanced inside launches


@pytest.mark.parametrize("cap", [40, 256])
def test_resident_lattice121_small_chunks(cap):
    """Tiny chunks: up to 32 workgroups per vector, frequent rollbacks."""
    ths = wl.theta_h_values()[1:5]
    bt = lattice_batch(cap, 3, ths)
    launches, fallbacks, aborts = bt.resident_stats()
    assert launches >= 3 and aborts == 0 and fallbacks > 0
    assert bt.resident_work()[3] > 0


@pytest.mark.parametrize("count", [12, 64])
def test_resident_batched_instances(count):
    """The first `count` seed-7 batched instances (uneven N, chain and grid
    topologies, RZZ/RX/RZ) for 2 layers, per key-width group.  With 64
    instances the XCDs cannot give every vector the workgroups a
    quarter-full start asks for, so the planner falls back to fuller
    chunks."""
    ws = wl.batched_instances(count, seed=7, max_depth=2)
    for sel in ([w for w in ws if w.N <= 64], [w for w in ws if w.N > 64]):
        if not sel:
            continue
        bt = SparsePauliBatch([w.N for w in sel], [[((p.z, p.x), c) for p, c in w.observable] for w in sel])
        bt.set_resident(2, 512)
        orcs = [po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable]) for w in sel]
        for li in range(2):
            bt.evolve_([w.layers[li].generators for w in sel], [w.layers[li].angles for w in sel],
                       truncation=sel[0].truncation)
            for w, o in zip(sel, orcs):
                L = w.layers[li]
                o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=compile_filter(w.truncation))
            log(f"batched W={bt.W} layer {li}: sizes {bt.sizes()} resident {bt.resident_stats()}")
        for b, o in enumerate(orcs):
            assert same_arrays(bt.arrays(b), o.arrays()), f"instance {sel[b].name}"
        assert bt.resident_stats()[2] == 0
