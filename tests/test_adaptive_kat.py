"""Known answers for AdaptiveTruncation's threshold on the oracle.

The reference (spv_evolve.jl:103-114, parity with the PauliSum method) sorts
|c|, takes thresh = coeffs[end - max_terms] (1-based) when n > max_terms and
drops every |c| <= thresh; otherwise it clips at min_thresh.  With distinct
magnitudes exactly max_terms terms survive: the largest ones.  The expected
survivors below are written out by hand, so an index off by one in either
direction (keeping max_terms - 1 or max_terms + 1 terms) fails.  The
reference's own test (test_truncation.jl:237-248) checks only
length <= max_terms."""
import numpy as np
import pytest

import pyoracle as po

# (coefficients, max_terms, min_thresh, surviving coefficients in key order)
KATS = [
    ([1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0], 3, 0.0, [8.0, 9.0, 10.0]),
    ([-1.0, 2.0, -3.0, 4.0, -5.0, 6.0, -7.0, 8.0, -9.0, 10.0], 4, 0.0, [-7.0, 8.0, -9.0, 10.0]),
    # ties at the threshold are all dropped: fewer than max_terms survive
    ([1.0, 2.0, 3.0, 3.0, 3.0, 4.0], 2, 0.0, [4.0]),
    ([1.0, 2.0, 3.0, 3.0, 3.0, 4.0], 4, 0.0, [3.0, 3.0, 3.0, 4.0]),
    ([0.0, -5.0, 3.0, -1.0, 2.0], 2, 0.0, [-5.0, 3.0]),
    # n == max_terms + 1: only the smallest goes
    ([0.5, 0.25, 0.125, 0.0625], 3, 0.0, [0.5, 0.25, 0.125]),
    # n <= max_terms: clip at min_thresh (|c| <= min_thresh dropped)
    ([1e-9, 1e-7, 1e-3, 0.5], 4, 1e-6, [1e-3, 0.5]),
    ([1e-9, 1e-7, 1e-3, 0.5], 10, 1e-3, [0.5]),
]


def run_oracle(coeffs, k, mn):
    N = 8
    terms = [(i + 1, 0, c) for i, c in enumerate(coeffs)]  # keys z = 1, 2, ... in key order
    o = po.OracleSPV(N, terms)
    o.adaptive(k, mn)
    return o.arrays()


@pytest.mark.parametrize("coeffs,k,mn,expect", KATS)
def test_oracle_adaptive_known_answers(coeffs, k, mn, expect):
    z, x, c = run_oracle(coeffs, k, mn)
    assert list(c) == expect
    # survivors keep their own keys, in key order
    keys = z[:, 0].tolist()
    assert keys == sorted(keys) and [coeffs[k - 1] for k in keys] == expect


def test_oracle_adaptive_top_k_of_distinct_magnitudes():
    """Distinct |c|: exactly the k largest survive (numpy argsort as the
    independent check), for k across the range."""
    rng = np.random.default_rng(3)
    n = 5000
    mags = rng.permutation(np.arange(1, n + 1)) * 1e-3
    signs = rng.choice([-1.0, 1.0], size=n)
    coeffs = list(mags * signs)
    for k in (1, 2, 17, 2500, n - 1):
        z, x, c = run_oracle(coeffs, k, 0.0)
        assert len(c) == k
        top = set(np.argsort(mags)[-k:] + 1)
        assert set(z[:, 0].tolist()) == top
