"""CPU oracle of the Pauli-sum products (spv_ops.jl:255-484).

TEST INFRASTRUCTURE ONLY.  Imported by tests/ as the checker of pp_algebra;
the product package never imports it.

Operands are (z, x, c) arrays: z, x of shape (n, W) uint64 (little-endian
words, site 1 = bit 0), c float64 (the reference's Float64 coefficient type)
or complex128 (ComplexF64).

* ``pairs``       every pair (i, j) in the reference's generation order
                  (i over A, j over B: spv_ops.jl:262-279, 418-445) with the
                  product key and coefficient, evaluated expression by
                  expression as Julia does (left to right; complex times
                  complex = (ar br - ai bi, ar bi + ai br); an integer or real
                  times a complex scales both parts).  numpy float64 products
                  and sums are single IEEE operations, so values match to the
                  bit.
* ``reduce_runs`` sort the pairs by key, then sum equal-key runs left to
                  right (_triples_to_live!, spv_ops.jl:229-251; zero sums
                  kept).  order="stable": equal keys in generation order (the
                  device's stable radix sort).  order="reference": the
                  reference's quicksort _sort_ws! (spv_kernels.jl:202-247),
                  restated with its pivot and insertion cutoff (pure Python:
                  small inputs only; 0-based bounds, whose differences and
                  midpoints equal the 1-based ones shifted by one).
"""
from __future__ import annotations

import numpy as np

MUL, COMMUTATOR, ANTICOMMUTATOR = 0, 1, 2


def _popcount(a: np.ndarray) -> np.ndarray:
    """Per-row popcount of (m, W) uint64 words."""
    a = a.astype(np.uint64, copy=True)
    out = np.zeros(a.shape[0], dtype=np.int64)
    for w in range(a.shape[1]):
        v = a[:, w]
        v = v - ((v >> np.uint64(1)) & np.uint64(0x5555555555555555))
        v = (v & np.uint64(0x3333333333333333)) + ((v >> np.uint64(2)) & np.uint64(0x3333333333333333))
        v = (v + (v >> np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
        out += ((v * np.uint64(0x0101010101010101)) >> np.uint64(56)).astype(np.int64)
    return out


def pairs(A, B, op: int, conj_a: bool = False, conj_b: bool = False):
    """(z, x, re, im, complex) of the kept pairs in generation order.  With
    real coefficients `im` is the product's imaginary part, which the
    reference's convert to Float64 requires to be zero (multiplication)."""
    az, ax, ac = A
    bz, bx, bc = B
    cplx = np.iscomplexobj(ac) or np.iscomplexobj(bc)
    if op == COMMUTATOR and not cplx:
        raise ValueError("commutator coefficients are imaginary; use a complex coefficient type")
    na, nb = len(ac), len(bc)
    ii = np.repeat(np.arange(na), nb)
    jj = np.tile(np.arange(nb), na)
    z1, x1, z2, x2 = az[ii], ax[ii], bz[jj], bx[jj]
    zp, xp = z1 ^ z2, x1 ^ x2
    n_a = _popcount(z1 & x1)
    n_b = _popcount(z2 & x2)
    m1 = _popcount(x1 & z2)
    m2 = _popcount(z1 & x2)
    n_p = _popcount(zp & xp)
    k = (n_p - n_a - n_b + 2 * m1) % 4
    ar = np.real(ac)[ii].astype(np.float64)
    br = np.real(bc)[jj].astype(np.float64)
    ai = (np.imag(ac)[ii] if cplx else np.zeros(len(ii))).astype(np.float64)
    bi = (np.imag(bc)[jj] if cplx else np.zeros(len(jj))).astype(np.float64)
    if op == MUL:
        if conj_a:
            ai = -ai
        if conj_b:
            bi = -bi
        keep = np.ones(len(ii), dtype=bool)
        pr = np.select([k == 0, k == 2], [1.0, -1.0], 0.0)
        pi = np.select([k == 1, k == 3], [1.0, -1.0], 0.0)
        if cplx:
            zr = pr * ar - pi * ai
            zi = pr * ai + pi * ar
            re = zr * br - zi * bi
            im = zr * bi + zi * br
        else:
            re = (pr * ar) * br
            im = (pi * ar) * br
    elif op == ANTICOMMUTATOR:
        keep = (k & 1) == 0
        f = (2 * (1 - k)).astype(np.float64)
        if cplx:
            zr = f * ar
            zi = f * ai
            re = zr * br - zi * bi
            im = zr * bi + zi * br
        else:
            re = (f * ar) * br
            im = np.zeros(len(ii))
    else:
        keep = ((m1 - m2) & 1) == 1
        f = (2 * (2 - k)).astype(np.float64)
        zr = 0.0 * ar - f * ai
        zi = 0.0 * ai + f * ar
        re = zr * br - zi * bi
        im = zr * bi + zi * br
    return zp[keep], xp[keep], re[keep], im[keep], cplx


def _key_tuple(z_row, x_row):
    # z-major, high word first (spv_kernels.jl:14-16)
    return tuple(int(v) for v in z_row[::-1]) + tuple(int(v) for v in x_row[::-1])


def reduce_runs(z, x, re, im, cplx: bool, order: str = "stable"):
    """Sorted unique (z, x, c) with equal-key runs summed left to right."""
    m = len(re)
    W = z.shape[1] if z.ndim == 2 else 1
    if m == 0:
        dt = np.complex128 if cplx else np.float64
        return np.zeros((0, W), np.uint64), np.zeros((0, W), np.uint64), np.zeros(0, dt)
    if order == "stable":
        cols = [x[:, w] for w in range(W)] + [z[:, w] for w in range(W)]  # least significant first
        perm = np.lexsort(cols)
    elif order == "reference":
        ws = [(_key_tuple(z[p], x[p]), p) for p in range(m)]
        _sort_ws_ref(ws, 0, m - 1)
        perm = np.array([p for _, p in ws], dtype=np.int64)
    else:
        raise ValueError(order)
    zs, xs, rs, ims = z[perm], x[perm], re[perm], im[perm]
    # runs of equal keys; each summed left to right: step t adds the run's
    # (t+1)-th element to the runs that long (one IEEE add per element, in
    # the same order as a sequential loop)
    diff = np.any(zs[1:] != zs[:-1], axis=1) | np.any(xs[1:] != xs[:-1], axis=1)
    starts = np.concatenate(([0], np.nonzero(diff)[0] + 1))
    lens = np.diff(np.concatenate((starts, [m])))
    sr = rs[starts].copy()
    si = ims[starts].copy()
    for t in range(1, int(lens.max())):
        sel = lens > t
        sr[sel] = sr[sel] + rs[starts[sel] + t]
        si[sel] = si[sel] + ims[starts[sel] + t]
    oz, ox = zs[starts], xs[starts]
    if cplx:
        oc = np.empty(len(sr), dtype=np.complex128)
        oc.real = sr
        oc.imag = si
    else:
        oc = sr
    return oz, ox, oc


def _sort_ws_ref(ws, lo, hi):
    """_sort_ws! with the reference's bounds: the loop runs while
    hi - lo >= 24 for 1-based inclusive lo, hi, i.e. the same difference of
    0-based inclusive bounds."""
    lt = lambda a, b: a[0] < b[0]
    while hi - lo >= 24:
        mid = (lo + hi) >> 1
        if lt(ws[mid], ws[lo]):
            ws[mid], ws[lo] = ws[lo], ws[mid]
        if lt(ws[hi], ws[mid]):
            ws[hi], ws[mid] = ws[mid], ws[hi]
            if lt(ws[mid], ws[lo]):
                ws[mid], ws[lo] = ws[lo], ws[mid]
        pivot = ws[mid]
        i, j = lo, hi
        while i <= j:
            while lt(ws[i], pivot):
                i += 1
            while lt(pivot, ws[j]):
                j -= 1
            if i <= j:
                ws[i], ws[j] = ws[j], ws[i]
                i += 1
                j -= 1
        if j - lo < hi - i:
            _sort_ws_ref(ws, lo, j)
            lo = i
        else:
            _sort_ws_ref(ws, i, hi)
            hi = j
    for kk in range(lo + 1, hi + 1):
        v = ws[kk]
        m = kk - 1
        while m >= lo and lt(v, ws[m]):
            ws[m + 1] = ws[m]
            m -= 1
        ws[m + 1] = v
    return ws


def product(A, B, op: int, conj_a: bool = False, conj_b: bool = False, order: str = "stable"):
    """The reference's product / commutator / anticommutator of two operands.
    Real coefficients: a product with a nonzero imaginary part raises
    (InexactError in the reference)."""
    z, x, re, im, cplx = pairs(A, B, op, conj_a, conj_b)
    if not cplx and op == MUL and np.any(im != 0.0):
        raise ValueError("InexactError: a product coefficient is not real")
    return reduce_runs(z, x, re, im, cplx, order)
