"""Full-depth GPU parity on the configured workloads.

* The 121-qubit sweep exactly as bench.py runs it (16 θ_h vectors in one
  batch, 20 layers, dephasing between layers), with the resident layer
  launches on (the default) and off: every final term set equals the
  oracle's bit for bit (keys, order, coefficient bit patterns), and the
  per-layer counts, final ⟨O⟩ and array digests equal the committed fixture
  tests/golden/lattice121_full.json.
* The 64-qubit chain (depolarizing 1e-3, |c| <= 1e-6, weight <= 16) for 8
  layers (9.0e5 live terms at the end), layer by layer.
* The first 32 seed-7 batched instances as key-width batches: chains for 4
  layers, grids for 3 (grids reach 1e6-8e7 terms by layer 4, beyond what the
  oracle finishes in a test).
"""
import json
import os
import sys

import numpy as np
import pytest

import pyoracle as po
from pauliprop import SparsePauliBatch, SparsePauliVector, compile_filter
from pauliprop import workloads as wl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def same(a, b):
    return (a[0].shape == b[0].shape and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            and np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64)))


@pytest.fixture(scope="module")
def lattice_oracle():
    ws = [wl.lattice121(th) for th in wl.theta_h_values()]
    out = []
    for w in ws:
        o = mg.run_oracle(w, 20)
        out.append((o.arrays(), o.expectation(0)))
    return ws, out


@pytest.mark.parametrize("resident", [1, 0])
def test_lattice121_full_sweep_bitexact(lattice_oracle, resident):
    ws, orc = lattice_oracle
    fx = json.load(open(os.path.join(HERE, "golden", "lattice121_full.json")))
    bt = SparsePauliBatch([w.N for w in ws], [[((p.z, p.x), c) for p, c in w.observable] for w in ws])
    bt.reserve(1 << 18)
    bt.set_resident(resident, 0)
    packed = [bt.pack([w.layers[li].generators for w in ws], [w.layers[li].angles for w in ws]) for li in range(20)]
    for li in range(20):
        bt.evolve_packed(packed[li], truncation=ws[0].truncation)
        bt.channel_(*ws[0].channel)
        sz = bt.sizes()
        assert sz == [f["n_by_layer"][li] for f in fx], f"layer {li} sizes"
    log(f"resident={resident}: final sizes {bt.sizes()} launches/fallbacks/aborts {bt.resident_stats()}")
    evs = bt.expectation_values()
    for b, ((oz, ox, oc), oe) in enumerate(orc):
        got = bt.arrays(b)
        assert same(got, (oz, ox, oc)), f"theta_h index {b}: final term set differs"
        assert mg.digest(*got) == fx[b]["digest"]
        assert abs(evs[b] - oe) <= 1e-12 * max(1.0, abs(oe))
        assert abs(evs[b] - float.fromhex(fx[b]["expectation"])) <= 1e-12
    if resident:
        assert bt.resident_stats()[0] > 0 and bt.resident_stats()[2] == 0


def test_chain64_eight_layers_bitexact():
    w = wl.chain64(n_layers=8)
    obs = [(p.z, p.x, c) for p, c in w.observable]
    g = SparsePauliVector(w.N, [((z, x), c) for z, x, c in obs])
    o = po.OracleSPV(w.N, obs)
    f = compile_filter(w.truncation)
    for li, L in enumerate(w.layers):
        g.evolve_(L.generators, L.angles, truncation=w.truncation)
        g.depolarizing_channel_(w.channel[1])
        o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f)
        o.channel(*w.channel)
        log(f"chain64 layer {li}: n={len(g)}")
        assert same(g.arrays(), o.arrays()), f"layer {li}"
    assert abs(g.expectation_value(0) - o.expectation(0)) <= 1e-12


def test_batched_first32_instances_bitexact():
    ws = wl.batched_instances(32, seed=7, max_depth=4)
    depth = [4 if w.meta["topology"] == "chain" else 3 for w in ws]
    orcs = []
    for w, d in zip(ws, depth):
        o = po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable])
        f = compile_filter(w.truncation)
        for L in w.layers[:d]:
            o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f)
        orcs.append(o)
    for wide in (False, True):
        idx = [i for i, w in enumerate(ws) if (w.N > 64) == wide]
        bt = SparsePauliBatch([ws[i].N for i in idx], [[((p.z, p.x), c) for p, c in ws[i].observable] for i in idx])
        for li in range(4):
            gens = [ws[i].layers[li].generators if li < depth[i] else [] for i in idx]
            angs = [ws[i].layers[li].angles if li < depth[i] else [] for i in idx]
            bt.evolve_(gens, angs, truncation=ws[0].truncation)
            log(f"batched wide={wide} layer {li}: max n {max(bt.sizes())}")
        for b, i in enumerate(idx):
            assert same(bt.arrays(b), orcs[i].arrays()), f"instance {i} (N={ws[i].N}, {ws[i].meta['topology']})"
