"""GPU parity: the HIP path against the CPU oracle, bit for bit.

Every comparison is exact: the surviving key set, key order and every FP64
coefficient (compared as bit patterns).  ⟨O⟩ is compared with the exactly
rounded sum of the (bit-identical) coefficients, to 1 ulp, and with the
oracle's sequential sum within its condition number (1e-12 relative to sum |c|).
"""
import math
import random

import numpy as np
import pytest

import pyoracle as po
from pauliprop import (CoeffTruncation, CompositeTruncation, MajoranaWeightTruncation, NoTruncation, Pauli,
                       PauliBasis, PPCounters, SparsePauliVector, WeightDampedTruncation, WeightTruncation,
                       XWeightDampedTruncation, XWeightTruncation, compile_filter)
from pauliprop import workloads as wl

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def assert_same(gpu: SparsePauliVector, orc: po.OracleSPV, what=""):
    gz, gx, gc = gpu.arrays()
    oz, ox, oc = orc.arrays()
    assert gz.shape == oz.shape, f"{what}: n gpu={len(gc)} oracle={len(oc)}"
    assert np.array_equal(gz, oz), f"{what}: z keys differ"
    assert np.array_equal(gx, ox), f"{what}: x keys differ"
    assert np.array_equal(gc.view(np.uint64), oc.view(np.uint64)), f"{what}: coefficients differ"


def random_terms(N, n, rng, support=None):
    """n distinct random (z, x, c) with bits only inside `support` (a bit mask)."""
    support = support if support is not None else (1 << N) - 1
    seen = set()
    out = []
    while len(out) < n:
        z = rng.getrandbits(N) & support
        x = rng.getrandbits(N) & support
        if (z, x) in seen:
            continue
        seen.add((z, x))
        out.append((z, x, rng.gauss(0, 1)))
    return out


def pair(N, terms, **kw):
    g = SparsePauliVector(N, [((z, x), c) for z, x, c in terms], **kw)
    o = po.OracleSPV(N, terms, **{k: v for k, v in kw.items()})
    return g, o


def spread_support(N, bits, rng):
    return sum(1 << q for q in rng.sample(range(N), bits))


@pytest.mark.parametrize("N", [6, 8, 16, 64, 65, 121, 128])
def test_single_rotation_random_generators(N):
    rng = random.Random(1000 + N)
    for trial in range(6):
        bits = min(N, 7)
        sup = spread_support(N, bits, rng)
        n = rng.choice([1, 2, 63, 64, 65, 2047, 2048, 2049, 5000, 9000])
        n = min(n, 2 ** (2 * bits) - 1)
        terms = random_terms(N, n, rng, sup)
        g, o = pair(N, terms)
        for k in range(4):
            kind = rng.randrange(4)
            if kind == 0:
                gz, gx = rng.getrandbits(N) & sup, rng.getrandbits(N) & sup
            elif kind == 1:
                q = rng.choice([b for b in range(N) if sup >> b & 1])
                gz, gx = 0, 1 << q
            elif kind == 2:
                qs = rng.sample([b for b in range(N) if sup >> b & 1], 2)
                gz, gx = (1 << qs[0]) | (1 << qs[1]), 0
            else:
                gz, gx = rng.getrandbits(N), rng.getrandbits(N)
            theta = rng.choice([0.0, 0.3, math.pi / 2, math.pi, -1.234, rng.uniform(-3, 3)])
            log(f"N={N} trial={trial} n={len(terms)} step={k} G=({gz:x},{gx:x}) θ={theta}")
            g.evolve_(PauliBasis(N, gz, gx), theta)
            o.evolve_single(gz, gx, theta)
            assert_same(g, o, f"N={N} trial={trial} step={k}")
            g.check()


def test_dense_small_keyspace_collisions():
    # N = 6: 4096 possible keys, 3000 live -> most children collide with live terms
    rng = random.Random(7)
    N = 6
    terms = random_terms(N, 3000, rng)
    g, o = pair(N, terms)
    for step in range(40):
        gz, gx = rng.getrandbits(N), rng.getrandbits(N)
        th = rng.uniform(-3, 3)
        g.evolve_(PauliBasis(N, gz, gx), th)
        o.evolve_single(gz, gx, th)
        assert_same(g, o, f"step {step}")


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
          MajoranaWeightTruncation(6), WeightDampedTruncation(0.5, 1e-4), XWeightDampedTruncation(0.5, 1e-4),
          CompositeTruncation(WeightTruncation(4), CoeffTruncation(1e-6))]


@pytest.mark.parametrize("strat", STRATS, ids=lambda s: type(s).__name__)
def test_sequence_strategies(strat):
    for N in (8, 70):
        gens, angs = heisenberg_gens(N if N == 8 else 8)
        gens = [PauliBasis(N, p.z, p.x) for p in gens]
        O0 = [(Pauli(N, Z=[4]).z, 0, 1.0)]
        g, o = pair(N, O0)
        g.evolve_(gens, angs, truncation=strat)
        o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(strat))
        assert_same(g, o, f"{type(strat).__name__} N={N}")


def test_local_truncation_and_counters():
    N = 10
    gens, angs = heisenberg_gens(N)
    O0 = [(Pauli(N, Z=[5]).z, 0, 1.0)]
    g, o = pair(N, O0)
    cnt = PPCounters()
    st = CoeffTruncation(1e-5)
    loc = WeightTruncation(4)
    g.evolve_(gens, angs, truncation=st, local_truncation=loc, counters=cnt)
    oc = o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(st), local=compile_filter(loc))
    assert_same(g, o, "local truncation")
    assert cnt.rotations == oc.rotations
    assert cnt.terms_swept == oc.terms_swept
    assert cnt.terms_created == oc.terms_created
    assert cnt.merge_out == oc.merge_out
    assert cnt.early_merges == oc.early_merges


def test_kicked_ising_chain16_end_to_end():
    w = wl.kicked_ising_16()
    gens, angs = w.flat()
    obs = [(p.z, p.x, c) for p, c in w.observable]
    g, o = pair(w.N, obs)
    cnt = PPCounters()
    g.evolve_(gens, angs, truncation=w.truncation, counters=cnt)
    oc = o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(w.truncation))
    assert_same(g, o, "kicked ising 16")
    assert cnt.early_merges == oc.early_merges and oc.early_merges > 0
    assert g.capacities() == o.capacities()
    _, _, c = g.arrays()
    zz, xx, _ = g.arrays()
    diag = c[np.all(xx == 0, axis=1)]
    exact = math.fsum(diag.tolist())
    ev = g.expectation_value(0)
    assert abs(ev - exact) <= 2 * np.spacing(abs(exact))
    assert abs(ev - o.expectation(0)) <= 1e-12 * max(1.0, float(np.sum(np.abs(diag))))


def run_layers(w, layers, g, o):
    for li in range(layers):
        L = w.layers[li]
        g.evolve_(L.generators, L.angles, truncation=w.truncation)
        o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=compile_filter(w.truncation))
        if w.channel:
            kind, p = w.channel
            getattr(g, kind + "_channel_")(p)
            o.channel(kind, p)
        log(f"{w.name} layer {li}: n={len(g)}")
        assert_same(g, o, f"{w.name} layer {li}")


def test_chain64_first_layers():
    w = wl.chain64()
    obs = [(p.z, p.x, c) for p, c in w.observable]
    g, o = pair(w.N, obs)
    run_layers(w, 4, g, o)
    ev = g.expectation_value(0)
    assert abs(ev - o.expectation(0)) <= 1e-12


def test_lattice121_first_layers():
    w = wl.lattice121(wl.theta_h_values()[9], n_layers=3)
    obs = [(p.z, p.x, c) for p, c in w.observable]
    g, o = pair(w.N, obs)
    run_layers(w, 3, g, o)
    assert abs(g.expectation_value(0) - o.expectation(0)) <= 1e-12


@pytest.mark.parametrize("kind", ["depolarizing", "dephasing", "bit_flip", "bit_phase_flip", "pauli"])
def test_channels_bitexact(kind):
    rng = random.Random(11)
    for N in (5, 64, 100):
        bits = min(N, 8)
        terms = random_terms(N, min(3000, 4 ** bits // 2), rng, spread_support(N, bits, rng))
        g, o = pair(N, terms)
        qs = sorted(rng.sample(range(1, N + 1), 3))
        mask = sum(1 << (q - 1) for q in qs)
        if kind == "pauli":
            g.pauli_channel_(0.05, 0.1, 0.15, qubits=qs)
            o.channel("pauli", pX=0.05, pY=0.1, pZ=0.15, mask=mask)
            g.pauli_channel_(0.01, 0.02, 0.03)
            o.channel("pauli", pX=0.01, pY=0.02, pZ=0.03)
        else:
            getattr(g, kind + "_channel_")(0.13, qubits=qs)
            o.channel(kind, 0.13, mask=mask)
            getattr(g, kind + "_channel_")(0.07)
            o.channel(kind, 0.07)
        assert_same(g, o, f"{kind} N={N}")


def test_clips_bitexact():
    rng = random.Random(12)
    for N in (8, 90):
        terms = random_terms(N, 5000, rng, spread_support(N, 8, rng))
        for strat in STRATS[1:]:
            g, o = pair(N, terms)
            g.truncate_(strat)
            o.compact(compile_filter(strat))
            assert_same(g, o, f"{type(strat).__name__} N={N}")


def test_edge_cases():
    N = 12
    # empty vector: evolution and expectation are no-ops
    g = SparsePauliVector(N, [])
    g.evolve_([Pauli(N, X=[1])], [0.4], truncation=CoeffTruncation(1e-6))
    assert len(g) == 0 and g.expectation_value(0) == 0.0
    # identity generator and commuting generator: values unchanged
    terms = [(Pauli(N, Z=[2]).z, 0, 0.5), (Pauli(N, Z=[3]).z, 0, -0.25)]
    g, o = pair(N, terms)
    g.evolve_(PauliBasis(N, 0, 0), 0.7)
    g.evolve_(Pauli(N, Z=[1, 4]), 0.7)
    o.evolve_single(0, 0, 0.7)
    p = Pauli(N, Z=[1, 4])
    o.evolve_single(p.z, p.x, 0.7)
    assert_same(g, o, "commuting")
    # everything truncated away
    g.evolve_([Pauli(N, X=[2])], [0.3], truncation=CoeffTruncation(10.0))
    assert len(g) == 0


def test_tight_capacity_early_merges():
    N = 8
    gens, angs = heisenberg_gens(N)
    O0 = [(Pauli(N, Z=[4]).z, 0, 1.0)]
    g, o = pair(N, O0, append_factor=0.01, min_capacity=4)
    cnt = PPCounters()
    st = CoeffTruncation(1e-4)
    g.depolarizing_channel_(0.2)
    o.channel("depolarizing", 0.2)
    g.evolve_(gens, angs, truncation=st, counters=cnt)
    oc = o.evolve([(p.z, p.x) for p in gens], angs, strict=compile_filter(st))
    assert_same(g, o, "tight capacity")
    assert cnt.early_merges == oc.early_merges > 0


def test_child_buffer_regrowth_after_ratio_jump():
    # Rotation 1 anticommutes with ~1% of the terms, so the child buffer is
    # sized small for rotation 2, which anticommutes with every term: the
    # overflow must be detected and the rotation re-run with exact sizes.
    rng = random.Random(21)
    N = 40
    q, r = 3, 17
    terms, seen = [], set()
    while len(terms) < 300000:
        z = rng.getrandbits(N) & ~(1 << q)
        x = rng.getrandbits(N) | (1 << r)
        if rng.random() < 0.01:
            z |= 1 << q
        if (z, x) in seen:
            continue
        seen.add((z, x))
        terms.append((z, x, rng.gauss(0, 1)))
    g, o = pair(N, terms)
    steps = [(0, 1 << q, 0.4), (1 << r, 0, 0.7), (1 << (q + 1), 1 << (q + 1), 0.2)]
    for gz, gx, th in steps:
        log(f"regrowth step G=({gz:x},{gx:x}) n={len(g)}")
        g.evolve_(PauliBasis(N, gz, gx), th)
        o.evolve_single(gz, gx, th)
        assert_same(g, o, f"regrowth G=({gz:x},{gx:x})")


def test_batch_uneven_vectors_bitexact():
    # Vectors of different qubit counts (one word type), circuits of different
    # lengths, per-vector truncation and channels: each must equal its own
    # single-vector oracle run.
    from pauliprop import SparsePauliBatch
    rng = random.Random(33)
    Ns = [8, 20, 64, 40, 13, 64]
    B = len(Ns)
    gens, angs, strats = [], [], []
    obs = []
    for b, N in enumerate(Ns):
        q = rng.randint(1, N)
        obs.append([(Pauli(N, Z=[q]).z, 0, 1.0)])
        L = rng.randint(30, 200)
        g, a = [], []
        for _ in range(L):
            kind = rng.randrange(3)
            j = rng.randint(1, N - 1)
            p = Pauli(N, Z=[j, j + 1]) if kind == 0 else (Pauli(N, X=[j]) if kind == 1 else Pauli(N, Z=[j]))
            g.append(p)
            a.append(rng.uniform(0, math.pi))
        gens.append(g)
        angs.append(a)
        strats.append(CompositeTruncation(CoeffTruncation(10 ** rng.uniform(-6, -3)), WeightTruncation(rng.randint(4, 12))))
    bt = SparsePauliBatch(Ns, [[((z, x), c) for z, x, c in o] for o in obs])
    orcs = [po.OracleSPV(N, o) for N, o in zip(Ns, obs)]
    for layer in range(3):
        log(f"batch layer {layer} sizes={bt.sizes()}")
        bt.evolve_(gens, angs, truncation=strats)
        bt.channel_(["depolarizing" if b % 2 == 0 else "dephasing" for b in range(B)], [1e-3] * B)
        for b in range(B):
            orcs[b].evolve([(p.z, p.x) for p in gens[b]], angs[b], strict=compile_filter(strats[b]))
            orcs[b].channel("depolarizing" if b % 2 == 0 else "dephasing", 1e-3)
    evs = bt.expectation_values()
    for b in range(B):
        gz, gx, gc = bt.arrays(b)
        oz, ox, oc = orcs[b].arrays()
        assert np.array_equal(gz, oz) and np.array_equal(gx, ox), f"vector {b} keys"
        assert np.array_equal(gc.view(np.uint64), oc.view(np.uint64)), f"vector {b} coefficients"
        assert abs(evs[b] - orcs[b].expectation(0)) <= 1e-12


def random_two_qubit_pauli(N, sup, rng):
    """A Pauli with random letters on two qubits of the support: 1 to 4 mask levels."""
    qs = rng.sample([b for b in range(N) if sup >> b & 1], 2)
    gz = gx = 0
    for q in qs:
        letter = rng.randint(1, 3)  # 1 X, 2 Z, 3 Y
        gx |= (letter & 1) << q
        gz |= (letter >> 1) << q
    return gz, gx


@pytest.mark.parametrize("N,tile,mtile", [(40, 512, 512), (40, 2048, 2048), (100, 512, 1024), (100, 2048, 1024)])
def test_segment_summaries_mixed_levels(N, tile, mtile):
    """Masks of 1 to 4 levels (children ranked from kb_flags' per-tile segment
    summaries) and many-level masks (kb_bounds' key searches) in the same
    steps of one batch, sweep tiles of 512 and 2048 terms, segments spanning
    many tiles: every vector equals its oracle run bit for bit."""
    from pauliprop import SparsePauliBatch
    rng = random.Random(N * 7 + tile)
    sizes = [20000, 9000, 600, 3000]
    B = len(sizes)
    sups = [spread_support(N, 9, rng) for _ in range(B)]
    init = [random_terms(N, n, rng, sup) for n, sup in zip(sizes, sups)]
    gens, angs = [], []
    for b in range(B):
        g, a = [], []
        for k in range(24):
            if b == 3 and k % 3 == 0:
                gz, gx = rng.getrandbits(N) & sups[b], rng.getrandbits(N) & sups[b]
            else:
                gz, gx = random_two_qubit_pauli(N, sups[b], rng)
            g.append(PauliBasis(N, gz, gx))
            a.append(rng.uniform(-3, 3))
        gens.append(g)
        angs.append(a)
    strat = CoeffTruncation(1e-9)
    bt = SparsePauliBatch([N] * B, [[((z, x), c) for z, x, c in t] for t in init])
    bt.set_tiles(tile, mtile)
    log(f"N={N} tiles {tile}/{mtile}: sizes {bt.sizes()}")
    bt.evolve_(gens, angs, truncation=[strat] * B)
    log(f"sizes after {bt.sizes()}")
    for b in range(B):
        o = po.OracleSPV(N, init[b])
        o.evolve([(p.z, p.x) for p in gens[b]], angs[b], strict=compile_filter(strat))
        gz, gx, gc = bt.arrays(b)
        oz, ox, oc = o.arrays()
        assert np.array_equal(gz, oz) and np.array_equal(gx, ox), f"vector {b} keys"
        assert np.array_equal(gc.view(np.uint64), oc.view(np.uint64)), f"vector {b} coefficients"


def _batch_vs_oracle(N, init, gens, angs, strat, reserve=None, tiles=None):
    from pauliprop import SparsePauliBatch
    B = len(init)
    bt = SparsePauliBatch([N] * B, [[((z, x), c) for z, x, c in t] for t in init])
    if reserve:
        bt.reserve(reserve)
    if tiles:
        bt.set_tiles(*tiles)
    bt.evolve_(gens, angs, truncation=[strat] * B)
    for b in range(B):
        o = po.OracleSPV(N, init[b])
        o.evolve([(p.z, p.x) for p in gens[b]], angs[b], strict=compile_filter(strat))
        gz, gx, gc = bt.arrays(b)
        oz, ox, oc = o.arrays()
        assert np.array_equal(gz, oz) and np.array_equal(gx, ox), f"vector {b} keys"
        assert np.array_equal(gc.view(np.uint64), oc.view(np.uint64)), f"vector {b} coefficients"


def _two_qubit_circuits(N, sups, steps, rng):
    gens, angs = [], []
    for sup in sups:
        g, a = [], []
        for _ in range(steps):
            gz, gx = random_two_qubit_pauli(N, sup, rng)
            g.append(PauliBasis(N, gz, gx))
            a.append(rng.uniform(-3, 3))
        gens.append(g)
        angs.append(a)
    return gens, angs


def test_large_vector_switches_every_vector_to_the_bounds_table():
    """A call whose largest vector exceeds 256 sweep tiles ranks the children
    of every vector (small ones included) from kb_bounds' key searches, with
    kb_scan's tile offsets: the same terms as the oracle."""
    rng = random.Random(404)
    N = 40
    sups = [spread_support(N, 11, rng) for _ in range(3)]
    init = [random_terms(N, n, rng, sup) for n, sup in zip([160000, 3000, 700], sups)]
    gens, angs = _two_qubit_circuits(N, sups, 8, rng)
    log(f"large call: sizes {[len(t) for t in init]}")
    _batch_vs_oracle(N, init, gens, angs, CoeffTruncation(1e-9))


def test_large_capacity_uses_the_scan_kernel():
    """Buffers sized beyond 512 sweep tiles make every step run kb_scan for
    the child offsets instead of kb_emit's own sums: same terms."""
    rng = random.Random(405)
    N = 70
    sups = [spread_support(N, 9, rng) for _ in range(2)]
    init = [random_terms(N, n, rng, sup) for n, sup in zip([20000, 5000], sups)]
    gens, angs = _two_qubit_circuits(N, sups, 12, rng)
    _batch_vs_oracle(N, init, gens, angs, CoeffTruncation(1e-9), reserve=600000)


def test_many_tiles_contiguous_block_layout():
    """More tiles than the XCD-row layout takes (over 16 per CU): blocks are
    laid out contiguously; two vectors of ~1.2e6 terms against the oracle."""
    rng = random.Random(406)
    N = 40
    sups = [spread_support(N, 12, rng) for _ in range(2)]
    init = [random_terms(N, n, rng, sup) for n, sup in zip([1200000, 1100000], sups)]
    gens, angs = _two_qubit_circuits(N, sups, 4, rng)
    _batch_vs_oracle(N, init, gens, angs, CoeffTruncation(1e-7), tiles=(512, 512))


def test_folded_sweep_workgroups():
    """A call with more than 32 sweep tiles per CU gives each flags/emit
    workgroup several tiles in turn (fewer workgroups than tiles by a fixed
    factor): twenty vectors of 1.8e5 terms against the oracle."""
    rng = random.Random(408)
    N = 40
    sups = [spread_support(N, 12, rng) for _ in range(20)]
    init = [random_terms(N, 180000, rng, sup) for sup in sups]
    gens, angs = _two_qubit_circuits(N, sups, 3, rng)
    log("folded sweep call: 20 x 180000 terms")
    _batch_vs_oracle(N, init, gens, angs, CoeffTruncation(1e-7), tiles=(512, 512))


def test_batch_graph_replays_match_eager():
    """hipGraph-captured sequences: the same layer circuit replayed over
    several calls, with eager channel work and host synchronisation between
    the replays; every replay's output equals an eager twin batch and the
    oracle bit for bit."""
    from pauliprop import SparsePauliBatch
    Ns = [64, 40, 64, 24]
    B = len(Ns)
    rng = random.Random(71)
    gens, angs = [], []
    for N in Ns:
        g = [Pauli(N, Z=[j, j + 1]) for j in range(1, N)] + [Pauli(N, X=[j]) for j in range(1, N + 1)]
        gens.append(g + [Pauli(N, Z=[1])] * (127 - len(g)))  # equal step counts per call
        angs.append([rng.uniform(0, math.pi) for _ in range(127)])
    obs = [[(Pauli(N, Z=[N // 2]).z, 0, 1.0)] for N in Ns]
    st = CoeffTruncation(1e-5)
    g_bt = SparsePauliBatch(Ns, [[((z, x), c) for z, x, c in o] for o in obs]).set_graphs(True)
    e_bt = SparsePauliBatch(Ns, [[((z, x), c) for z, x, c in o] for o in obs])
    g_bt.reserve(1 << 16)
    e_bt.reserve(1 << 16)
    orcs = [po.OracleSPV(N, o) for N, o in zip(Ns, obs)]
    for layer in range(6):
        g_bt.evolve_(gens, angs, truncation=st)
        e_bt.evolve_(gens, angs, truncation=st)
        g_bt.channel_("depolarizing", 1e-3)
        e_bt.channel_("depolarizing", 1e-3)
        for b in range(B):
            orcs[b].evolve([(p.z, p.x) for p in gens[b]], angs[b], strict=compile_filter(st))
            orcs[b].channel("depolarizing", 1e-3)
        assert g_bt.expectation_values() == e_bt.expectation_values()
        for b in range(B):
            gz, gx, gc = g_bt.arrays(b)
            ez, ex, ec = e_bt.arrays(b)
            assert np.array_equal(gz, ez) and np.array_equal(gx, ex), f"layer {layer} vector {b} keys"
            assert np.array_equal(gc.view(np.uint64), ec.view(np.uint64)), f"layer {layer} vector {b} coefficients"
            oz, ox, oc = orcs[b].arrays()
            assert np.array_equal(gz, oz) and np.array_equal(gc.view(np.uint64), oc.view(np.uint64))
        log(f"graph layer {layer}: sizes {g_bt.sizes()} stats {g_bt.graph_stats()}")
    caps, launches = g_bt.graph_stats()
    assert launches >= 6  # one per layer call, plus replays of a call's tail after a vector outgrew its cluster
    # zero angles: every child has c = 0 and is cut, sizes and grids stay
    # fixed, so the captured sequence is replayed
    zero = [[0.0] * 127 for _ in range(B)]
    for rep in range(3):
        g_bt.evolve_(gens, zero, truncation=st)
        e_bt.evolve_(gens, zero, truncation=st)
        assert g_bt.expectation_values() == e_bt.expectation_values()
        for b in range(B):
            assert np.array_equal(g_bt.arrays(b)[2].view(np.uint64), e_bt.arrays(b)[2].view(np.uint64))
    caps2, launches2 = g_bt.graph_stats()
    log(f"graph stats after replays: {caps2} captures, {launches2} launches")
    assert launches2 == launches + 3 and caps2 <= caps + 1


def test_batched_instances_bitexact_with_graphs_and_chunks():
    """The first 12 instances of the seed-7 batched set (2 layers each), run
    as in the batched driver (key-width groups, 64-step chunks, captured
    graphs), equal their single-vector oracle runs bit for bit."""
    from pauliprop import SparsePauliBatch
    from pauliprop import batched_bench as bb
    ws = wl.batched_instances(12, seed=7, max_depth=2)
    for sel in ([w for w in ws if w.N <= 64], [w for w in ws if w.N > 64]):
        if not sel:
            continue
        bt = SparsePauliBatch([w.N for w in sel], [[((p.z, p.x), c) for p, c in w.observable] for w in sel])
        bt.set_graphs(True)
        gz, gx, layers = bb._group_arrays(sel, bt.W)
        orcs = [po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable]) for w in sel]
        for li, (cv, sv, act) in enumerate(layers):
            for s0 in range(0, act.shape[0], 64):
                bt.evolve_packed((gz[s0:s0 + 64], gx[s0:s0 + 64], cv[s0:s0 + 64], sv[s0:s0 + 64],
                                  np.ascontiguousarray(act[s0:s0 + 64])), truncation=sel[0].truncation)
            for w, o in zip(sel, orcs):
                L = w.layers[li]
                o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=compile_filter(w.truncation))
            log(f"batched instances W={bt.W} layer {li}: sizes {bt.sizes()}")
        for b, o in enumerate(orcs):
            gz_, gx_, gc_ = bt.arrays(b)
            oz, ox, oc = o.arrays()
            assert np.array_equal(gz_, oz) and np.array_equal(gx_, ox), f"instance {sel[b].name} keys"
            assert np.array_equal(gc_.view(np.uint64), oc.view(np.uint64)), f"instance {sel[b].name} coefficients"
