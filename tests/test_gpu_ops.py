"""Device lookups and reductions: get/getindex/haskey by binary search
(type_SparsePauliVector.jl:286-311), norm (spv_ops.jl:57-84), isapprox
(:88-107), inner_product (:378-396), tr (:52-55), against plain Python on
the same data.  Sums: the device accumulates in double-double, so results
equal the exactly rounded sum (math.fsum) to a few ulps and the reference's
left-to-right sum to 1e-12 relative."""
import math
import random

import numpy as np
import pytest

from pauliprop import Pauli, PauliBasis, SparsePauliVector, inner_product, isapprox, norm

pytestmark = pytest.mark.gpu


def rand_vec(N, n, rng, keys=None):
    keys = keys if keys is not None else set()
    out = {}
    while len(out) < n:
        k = (rng.getrandbits(N), rng.getrandbits(N))
        out[k] = rng.gauss(0, 1) * 10 ** rng.uniform(-6, 0)
    return out


def close(a, b, ulps=4):
    return abs(a - b) <= ulps * np.spacing(max(abs(a), abs(b), 1e-300))


@pytest.mark.parametrize("N", [12, 64, 65, 128])
def test_lookup_get_getindex_haskey(N):
    rng = random.Random(N)
    d = rand_vec(N, 20000, rng)
    v = SparsePauliVector(N, list(d.items()))
    keys = list(d.keys())
    vals, found = v.lookup(keys)
    assert found.all()
    assert all(vals[i] == d[k] for i, k in enumerate(keys))
    absent = []
    while len(absent) < 5000:
        k = (rng.getrandbits(N), rng.getrandbits(N))
        if k not in d:
            absent.append(k)
    vals, found = v.lookup(absent)
    assert not found.any() and not vals.any()
    k0 = keys[17]
    p = PauliBasis(N, *k0)
    assert v[p] == d[k0] and v.get(p) == d[k0] and v.haskey(p) and p in v
    q = PauliBasis(N, *absent[0])
    with pytest.raises(KeyError):
        v[q]
    assert v.get(q, 7.5) == 7.5 and q not in v
    assert v.tr() == 0.0 or (0, 0) in d
    w = SparsePauliVector(N, [((0, 0), 0.25), (k0, 1.0)])
    assert w.tr() == 0.25 * 2.0 ** N


@pytest.mark.parametrize("N", [20, 100])
def test_norms(N):
    rng = random.Random(3 + N)
    d = rand_vec(N, 50000, rng)
    v = SparsePauliVector(N, list(d.items()))
    cs = [d[k] for k in sorted(d)]
    assert close(v.norm(2), math.sqrt(math.fsum(c * c for c in cs)))
    assert close(v.norm(1), math.fsum(abs(c) for c in cs))
    assert v.norm(math.inf) == max(abs(c) for c in cs)
    ref3 = math.fsum(abs(c) ** 3 for c in cs) ** (1 / 3)
    assert abs(v.norm(3) - ref3) <= 1e-13 * ref3
    seq = 0.0
    for c in cs:
        seq += c * c
    assert abs(norm(v) - math.sqrt(seq)) <= 1e-12 * math.sqrt(seq)


@pytest.mark.parametrize("N", [30, 90])
def test_inner_product_and_isapprox(N):
    rng = random.Random(11 + N)
    a = rand_vec(N, 30000, rng)
    b = {}
    for k, c in a.items():
        if rng.random() < 0.5:
            b[k] = c + rng.gauss(0, 1e-3)
    b.update(rand_vec(N, 10000, rng))
    va = SparsePauliVector(N, list(a.items()))
    vb = SparsePauliVector(N, list(b.items()))
    ip = math.fsum(a[k] * b[k] for k in a if k in b)
    assert close(inner_product(va, vb), ip, ulps=16)
    assert close(inner_product(vb, va), ip, ulps=16)
    keys = set(a) | set(b)
    dist = math.sqrt(math.fsum((a.get(k, 0.0) - b.get(k, 0.0)) ** 2 for k in keys))
    assert isapprox(va, vb, atol=dist * (1 + 1e-12)) and not isapprox(va, vb, atol=dist * (1 - 1e-9))
    va2 = SparsePauliVector(N, list(a.items()))
    assert isapprox(va, va2, atol=0.0)
    assert isapprox(va, vb, rtol=1.0)


@pytest.mark.parametrize("N", [16, 100])
def test_setindex_and_delete(N):
    """setindex! / delete! (type_SparsePauliVector.jl:315-358) against dict
    semantics, with the reference's live-capacity growth per inserted key."""
    rng = random.Random(100 + N)
    d = rand_vec(N, 3000, rng)
    v = SparsePauliVector(N, list(d.items()))
    cap0 = v.capacities()[0]
    keys = list(d)
    upd = []
    for k in rng.sample(keys, 500):
        upd.append((k, rng.gauss(0, 1)))
    new = []
    while len(new) < 7000:
        k = (rng.getrandbits(N), rng.getrandbits(N))
        if k not in d:
            new.append((k, rng.gauss(0, 1)))
    batch = upd + new + [(new[0][0], 42.0)]  # repeated key: last value wins
    rng.shuffle(batch)
    batch.append((new[1][0], -3.0))
    ref = dict(d)
    n, cap = len(ref), cap0
    for k, c in batch:
        if k not in ref:
            if n == cap:
                cap = max(2 * cap, n + 1)
            n += 1
        ref[k] = c
    v.setindex_(batch)
    assert v.check()
    got = {(z, x): c for z, x, c in v.terms()}
    assert got == ref
    assert v.capacities()[0] == cap
    p = PauliBasis(N, *keys[0])
    v[p] = 0.5
    ref[keys[0]] = 0.5
    assert v[p] == 0.5
    dele = rng.sample(list(ref), 1000) + [(rng.getrandbits(N) | 1, 0)]
    v.delete_(dele)
    for k in dele:
        ref.pop(k, None)
    got = {(z, x): c for z, x, c in v.terms()}
    assert got == ref and v.check()
    del v[PauliBasis(N, *next(iter(ref)))]
    assert len(v) == len(ref) - 1


@pytest.mark.parametrize("N", [40, 120])
def test_sum_mul_eq(N):
    """sum! (spv_ops.jl:182-193, a merge summing shared keys), mul! (:12-17),
    == (type_SparsePauliVector.jl:380-388) against dict arithmetic; shared
    keys are summed once (c1 + c2), so the values are bit-exact."""
    rng = random.Random(50 + N)
    a = rand_vec(N, 20000, rng)
    b = {k: rng.gauss(0, 1) for k in rng.sample(list(a), 5000)}
    b.update(rand_vec(N, 8000, rng))
    for k in rng.sample(list(a), 300):
        b[k] = -a[k]  # NOFILTER keeps the exact zeros
    va = SparsePauliVector(N, list(a.items()))
    vb = SparsePauliVector(N, list(b.items()))
    cap = va.capacities()[0]
    va.sum_(vb)
    ref = dict(a)
    for k, c in b.items():
        ref[k] = ref[k] + c if k in ref else c
    got = {(z, x): c for z, x, c in va.terms()}
    assert got == ref and va.check()
    assert sum(1 for c in got.values() if c == 0.0) >= 300
    n_in = len(a) + len(b)
    assert va.capacities()[0] == (max(2 * cap, n_in) if len(ref) > cap else cap)
    vb.mul_(-0.5)
    assert {(z, x): c for z, x, c in vb.terms()} == {k: c * -0.5 for k, c in b.items()}
    vs = va + vb
    assert len(vs) == len(set(ref) | set(b))
    v2 = SparsePauliVector(N, list(a.items()))
    assert v2 == SparsePauliVector(N, list(a.items()))
    assert not (v2 == va)
    assert (-v2).terms()[0][2] == -v2.terms()[0][2]


def test_filter():
    """filter! (type_SparsePauliVector.jl:389-404) with a host predicate."""
    rng = random.Random(77)
    d = rand_vec(70, 5000, rng)
    v = SparsePauliVector(70, list(d.items()))
    keep = lambda pr: abs(pr[1]) > 1e-3 and pr[0].z % 3 != 0
    w = v.filter(keep)
    assert len(v) == len(d)
    ref = {k: c for k, c in d.items() if abs(c) > 1e-3 and k[0] % 3 != 0}
    assert {(z, x): c for z, x, c in w.terms()} == ref and w.check()
    assert v.filter_(lambda pr: True) == SparsePauliVector(70, list(d.items()))
