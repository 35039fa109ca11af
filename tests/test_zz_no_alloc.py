"""pp_lookup_noalloc (lookups with the staged queries and results in a fixed
per-GPU scratch array, no device allocation) against pp_lookup, through the
C ABI: identical found flags and coefficient bit patterns, on two calls in a
row with the second smaller, plus the size limit."""
import ctypes
import random

import numpy as np
import pytest

from pauliprop import SparsePauliVector, _lib

pytestmark = pytest.mark.gpu

MAX_QUERIES = 1 << 16


def _queries(N, live, count, rng):
    W = 1 if N <= 64 else 2
    z = np.zeros((count, W), dtype=np.uint64)
    x = np.zeros((count, W), dtype=np.uint64)
    for i in range(count):
        if i % 2 == 0 and live:
            kz, kx = live[rng.randrange(len(live))]
        else:
            kz, kx = rng.getrandbits(N), rng.getrandbits(N)
        for w in range(W):
            z[i, w] = (kz >> (64 * w)) & ((1 << 64) - 1)
            x[i, w] = (kx >> (64 * w)) & ((1 << 64) - 1)
    return z, x


def _call(name, g, z, x):
    count = z.shape[0]
    c = np.full(count, np.nan)
    found = np.full(count, -1, dtype=np.int32)
    rc = getattr(_lib.lib(), name)(g._h, _lib.u64p(z), _lib.u64p(x), count, _lib.dp(c),
                                   found.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return rc, c, found


@pytest.mark.parametrize("N", [40, 121])
def test_lookup_noalloc_matches_lookup(N):
    rng = random.Random(2024 + N)
    live = list({(rng.getrandbits(N), rng.getrandbits(N)) for _ in range(6000)})
    g = SparsePauliVector(N, [(k, rng.gauss(0, 1)) for k in live])
    for count in (40000, 700):  # the second call is smaller: stale scratch entries must not leak
        z, x = _queries(N, live, count, rng)
        rc0, c0, f0 = _call("pp_lookup", g, z, x)
        rc1, c1, f1 = _call("pp_lookup_noalloc", g, z, x)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(f0, f1) and int(f1.sum()) >= count // 2 - 1
        assert np.array_equal(c0.view(np.uint64), c1.view(np.uint64))


def test_lookup_noalloc_limit_and_empty():
    g = SparsePauliVector(20, [((1, 0), 0.5)])
    z = np.zeros((MAX_QUERIES + 1, 1), dtype=np.uint64)
    rc, _, _ = _call("pp_lookup_noalloc", g, z, z.copy())
    assert rc == _lib.PP_EARG
    assert b"scratch capacity" in _lib.lib().pp_last_error()
    z0 = np.zeros((MAX_QUERIES, 1), dtype=np.uint64)
    z0[0, 0] = 1
    rc, c, f = _call("pp_lookup_noalloc", g, z0, np.zeros_like(z0))
    assert rc == 0 and f[0] == 1 and c[0] == 0.5 and int(f.sum()) == 1
