"""GPU: AdaptiveTruncation's known answers on the device, and the merge
look-back tags across their 22-bit wrap.

* The device's radix-select threshold on the hand-written known answers of
  tests/test_adaptive_kat.py (thresh = sorted |c|[end - max_terms], drop
  |c| <= thresh; spv_evolve.jl:103-114), and exactly the k largest of 1e5
  distinct magnitudes against numpy.
* Look-back tags: every per-step and windowed merge tags its status words
  with a 22-bit launch tag; before the tag would wrap, the library clears the
  status arrays and restarts at 1.  pp_batch_set_epoch / pp_set_epoch move the
  tag base to just below 2^22 - 1, so a few calls run with tags at the top of
  the range and then across the wrap; the results stay bit-exact with the
  oracle.
"""
import numpy as np
import pytest

import pyoracle as po
from pauliprop import (AdaptiveTruncation, CoeffTruncation, Pauli, PauliBasis, SparsePauliBatch,
                       SparsePauliVector, compile_filter)
from pauliprop import workloads as wl
from test_adaptive_kat import KATS

pytestmark = pytest.mark.gpu
TAG_TOP = (1 << 22) - 1


def log(msg):
    print(msg, flush=True)


def same(a, b):
    return (a[0].shape == b[0].shape and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            and np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64)))


@pytest.mark.parametrize("coeffs,k,mn,expect", KATS)
def test_device_adaptive_known_answers(coeffs, k, mn, expect):
    v = SparsePauliVector(8, [((i + 1, 0), c) for i, c in enumerate(coeffs)])
    v.truncate_(AdaptiveTruncation(k, mn))
    z, x, c = v.arrays()
    keys = z[:, 0].tolist()
    assert list(c) == expect
    assert keys == sorted(keys) and [coeffs[q - 1] for q in keys] == expect


def test_device_adaptive_top_k_of_distinct_magnitudes():
    rng = np.random.default_rng(11)
    n = 100000
    mags = rng.permutation(np.arange(1, n + 1)).astype(np.float64) * 1e-6
    coeffs = mags * rng.choice([-1.0, 1.0], size=n)
    N = 40
    z = np.arange(1, n + 1, dtype=np.uint64).reshape(-1, 1)
    x = np.zeros((n, 1), dtype=np.uint64)
    for k in (1, 3, 4096, 65537, n - 1):
        v = SparsePauliVector(N, (z, x, coeffs.copy()))
        v.truncate_(AdaptiveTruncation(k, 0.0))
        gz, _, gc = v.arrays()
        assert len(gc) == k, k
        top = np.sort(np.argsort(mags)[-k:] + 1).astype(np.uint64)
        assert np.array_equal(gz[:, 0], top), k
        assert np.array_equal(gc.view(np.uint64), coeffs[top.astype(np.int64) - 1].view(np.uint64))


def test_batch_merge_tags_across_the_wrap():
    """Per-step kernels (resident launches off): layer 0 at the start of the
    tag range, layer 1 with tags just below 2^22 - 1, layer 2 after the clear."""
    ths = wl.theta_h_values()[1:5]
    ws = [wl.lattice121(th, n_layers=4) for th in ths]
    bt = SparsePauliBatch([w.N for w in ws], [[((p.z, p.x), c) for p, c in w.observable] for w in ws])
    bt.set_resident(0, 0)
    orcs = [po.OracleSPV(w.N, [(p.z, p.x, c) for p, c in w.observable]) for w in ws]
    f = compile_filter(ws[0].truncation)
    for li in range(4):
        if li == 1:
            bt.set_epoch(TAG_TOP - 400)  # one 341-step call fits below the top, the next wraps
        bt.evolve_([w.layers[li].generators for w in ws], [w.layers[li].angles for w in ws],
                   truncation=ws[0].truncation)
        bt.channel_(*ws[0].channel)
        for w, o in zip(ws, orcs):
            L = w.layers[li]
            o.evolve([(p.z, p.x) for p in L.generators], L.angles, strict=f)
            o.channel(*w.channel)
        for b, o in enumerate(orcs):
            assert same(bt.arrays(b), o.arrays()), f"layer {li} vector {b}"
        log(f"tags: layer {li} sizes {bt.sizes()}")


def test_windowed_merge_tags_across_the_wrap():
    """Windowed merges (window 3) on one vector: 40 boundaries per call, the
    tag base set 25 below the top, so the wrap falls inside a call."""
    N = 12
    terms = []
    for i in range(1, N):
        terms += [Pauli(N, X=[i, i + 1]), Pauli(N, Z=[i, i + 1])]
    terms += [Pauli(N, Z=[i]) for i in range(1, N + 1)]
    gens = [PauliBasis(N, p.z, p.x) for p in terms * 4][:120]
    rng = np.random.default_rng(5)
    angs = list(rng.uniform(-1.0, 1.0, size=len(gens)))
    O0 = [(Pauli(N, Z=[6]).z, 0, 1.0)]
    g = SparsePauliVector(N, [((z, x), c) for z, x, c in O0])
    o = po.OracleSPV(N, O0)
    st = CoeffTruncation(1e-6)
    po.set_sort_mode(True)  # the device's stable order of pending appends
    try:
        for call in range(3):
            if call == 1:
                g.set_epoch(TAG_TOP - 25)
            g.evolve_(gens, angs, window=3, truncation=st)
            o.evolve([(p.z, p.x) for p in gens], angs, window=3, strict=compile_filter(st))
            assert same(g.arrays(), o.arrays()), f"call {call}"
            log(f"windowed tags: call {call} n={len(g)}")
    finally:
        po.set_sort_mode(False)
