"""Independent pure-Python restatement of the reference's Dict-backed path.

Test-only.  This follows the PauliSum (Dict) semantics that the reference's
SparsePauliVector path must reproduce at window = 1 (spv_evolve.jl:1-7):
  * evolve!(::PauliSum, G, θ)            evolve.jl:32-46
  * evolve(O, gens, angs; truncation)    evolve.jl:111-121
  * coeff/weight/... clips               clip.jl:6-117
  * channels                             channels.jl:61-191
The product phase of G*P is composed site by site from the single-qubit
multiplication table (XY = iZ, YZ = iX, ZX = iY), independently of the fused
popcount identity used by the kernels, so it pins that identity.

Also dense-matrix and state-vector helpers (numpy) for brute-force checks.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np

Key = Tuple[int, int]  # (z, x) Python ints

# single-qubit letter codes: 0=I, 1=X, 2=Z, 3=Y  (z*2 + x)
_MUL = {}
for a in range(4):
    for b in range(4):
        _MUL[(a, b)] = None
_LET = {0: "I", 1: "X", 2: "Z", 3: "Y"}
_CODE = {v: k for k, v in _LET.items()}
_TABLE = {
    ("I", "I"): (1, "I"), ("I", "X"): (1, "X"), ("I", "Y"): (1, "Y"), ("I", "Z"): (1, "Z"),
    ("X", "I"): (1, "X"), ("X", "X"): (1, "I"), ("X", "Y"): (1j, "Z"), ("X", "Z"): (-1j, "Y"),
    ("Y", "I"): (1, "Y"), ("Y", "X"): (-1j, "Z"), ("Y", "Y"): (1, "I"), ("Y", "Z"): (1j, "X"),
    ("Z", "I"): (1, "Z"), ("Z", "X"): (1j, "Y"), ("Z", "Y"): (-1j, "X"), ("Z", "Z"): (1, "I"),
}


def letter(z: int, x: int, j: int) -> str:
    return _LET[((z >> j) & 1) * 2 + ((x >> j) & 1)]


def pauli_product(N: int, a: Key, b: Key) -> Tuple[complex, Key]:
    """(phase, key) with P_a P_b = phase * P_c, composed site by site."""
    phase = 1 + 0j
    zc = xc = 0
    for j in range(N):
        ph, l = _TABLE[(letter(a[0], a[1], j), letter(b[0], b[1], j))]
        phase *= ph
        code = _CODE[l]
        zc |= (code >> 1) << j
        xc |= (code & 1) << j
    return phase, (zc, xc)


def anticommute(a: Key, b: Key) -> bool:
    return (bin(a[1] & b[0]).count("1") + bin(a[0] & b[1]).count("1")) % 2 == 1


def dict_evolve_single(N: int, O: Dict[Key, float], G: Key, theta: float, cosv=None, sinv=None) -> None:
    """evolve.jl:32-46 on a dict of real or complex coefficients (in place)."""
    c_ = math.cos(theta) if cosv is None else cosv
    s_ = math.sin(theta) if sinv is None else sinv
    sin_branch: Dict[Key, float] = {}
    for p, c in list(O.items()):
        if anticommute(p, G):
            phase, k = pauli_product(N, G, p)
            # c * (i sinθ) * phase; for anticommuting pairs i*phase = ±1 exactly
            f = 1j * phase
            assert f.imag == 0 and abs(f.real) == 1
            v = c * s_ if f.real > 0 else -(c * s_)
            sin_branch[k] = sin_branch.get(k, 0.0) + v
            O[p] = c * c_
    for k, v in sin_branch.items():  # mergewith!(+, O, sin_branch)
        if k in O:
            O[k] = O[k] + v
        else:
            O[k] = v


def _dyadic(v: float):
    """v = n / 2^k exactly (k may be negative)."""
    n, d = float(v).as_integer_ratio()
    return n, d.bit_length() - 1


def _dy_add_half(a, b):
    """(a + b) / 2 of two dyadic rationals (n, k)."""
    (an, ak), (bn, bk) = a, b
    m = max(ak, bk)
    return (an << (m - ak)) + (bn << (m - bk)), m + 1


def _dy_sq_cmp(a, s):
    """Sign of a^2 - s for dyadic a = (n, k) and s = (sn, sk) (s already a square sum)."""
    (an, ak), (sn, sk) = a, s
    an2, ak2 = an * an, 2 * ak
    m = max(ak2, sk)
    l, r = an2 << (m - ak2), sn << (m - sk)
    return (l > r) - (l < r)


def exact_hypot(x: float, y: float) -> float:
    """The correctly rounded sqrt(x^2 + y^2), decided in exact (integer,
    dyadic) arithmetic: the double whose rounding interval holds the exact
    value (candidates around math.hypot, which is within one ulp)."""
    if math.isinf(x) or math.isinf(y):
        return math.inf
    if x == 0 and y == 0:
        return 0.0
    (xn, xk), (yn, yk) = _dyadic(x), _dyadic(y)
    m = max(xk, yk)
    S = ((xn * xn) << (2 * (m - xk))) + ((yn * yn) << (2 * (m - yk))), 2 * m
    big = 1.7976931348623157e308
    top = _dy_add_half(_dyadic(big), (1 << 1024, 0))  # rounds to inf from here
    if _dy_sq_cmp(top, S) <= 0:
        return math.inf
    h = math.hypot(x, y)
    if math.isinf(h):
        h = big
    cands = [h]
    for _ in range(3):
        cands = [math.nextafter(cands[0], 0.0)] + cands + [math.nextafter(cands[-1], math.inf)]

    def up(c):  # the next double above c (2^1024 past the largest finite one)
        n = math.nextafter(c, math.inf)
        return (1 << 1024, 0) if math.isinf(n) else _dyadic(n)

    for c in cands:
        if math.isinf(c):
            continue
        lo = _dy_add_half(_dyadic(math.nextafter(c, 0.0)), _dyadic(c))
        hi = _dy_add_half(_dyadic(c), up(c))
        if _dy_sq_cmp(lo, S) <= 0 <= _dy_sq_cmp(hi, S):
            return c
    raise AssertionError(f"no correctly rounded candidate for hypot({x!r}, {y!r})")


def cabs(c) -> float:
    """abs(c): |c| for Float64, the correctly rounded hypot for ComplexF64
    (the rounding Julia's hypot gives on FMA hardware)."""
    if isinstance(c, complex):
        return exact_hypot(c.real, c.imag)
    return abs(c)


def weight(k: Key) -> int:
    return bin(k[0] | k[1]).count("1")


def x_weight(k: Key) -> int:
    return bin(k[1]).count("1")


def majorana_weight(k: Key) -> int:
    """clip.jl:88-117 via an explicit per-site scan from the top site downward:
    control starts true and flips at every X/Y site; X/Y add 1, Z-only sites
    add 2 while control is true, I sites add 2 while control is false."""
    z, x = k
    w = 0
    control = True
    for j in range(127, -1, -1):
        zb = (z >> j) & 1
        xb = (x >> j) & 1
        if xb:
            w += 1
            control = not control
        elif zb:
            if control:
                w += 2
        else:
            if not control:
                w += 2
    return w


def dict_truncate(O: Dict[Key, float], strategy) -> None:
    """truncation.jl:166-243 (_apply! on PauliSum) for deterministic strategies."""
    from pauliprop import truncation as T
    s = strategy
    if isinstance(s, T.NoTruncation):
        return
    if isinstance(s, T.CoeffTruncation):
        for k in [k for k, c in O.items() if not cabs(c) > s.thresh]:
            del O[k]
    elif isinstance(s, T.WeightTruncation):
        for k in [k for k in O if not weight(k) <= s.max_weight]:
            del O[k]
    elif isinstance(s, T.XWeightTruncation):
        for k in [k for k in O if not x_weight(k) <= s.max_weight]:
            del O[k]
    elif isinstance(s, T.MajoranaWeightTruncation):
        for k in [k for k in O if not majorana_weight(k) <= s.max_weight]:
            del O[k]
    elif isinstance(s, T.WeightDampedTruncation):
        for k in [k for k, c in O.items() if not cabs(c) * math.exp(-s.alpha * weight(k)) > s.thresh]:
            del O[k]
    elif isinstance(s, T.XWeightDampedTruncation):
        for k in [k for k, c in O.items() if not cabs(c) * math.exp(-s.alpha * x_weight(k)) > s.thresh]:
            del O[k]
    elif isinstance(s, T.AdaptiveTruncation):
        if len(O) > s.max_terms:
            coeffs = sorted(cabs(c) for c in O.values())
            thresh = coeffs[len(coeffs) - 1 - s.max_terms]
            dict_truncate(O, T.CoeffTruncation(thresh))
        else:
            dict_truncate(O, T.CoeffTruncation(s.min_thresh))
    elif isinstance(s, T.CompositeTruncation):
        for t in s.strategies:
            dict_truncate(O, t)
    else:
        raise NotImplementedError(type(s))


def dict_evolve(N: int, O: Dict[Key, float], gens: Sequence[Key], angs: Sequence[float], strategy) -> Dict[Key, float]:
    """evolve.jl:111-121: evolve! then truncate! per rotation."""
    out = dict(O)
    for g, a in zip(gens, angs):
        dict_evolve_single(N, out, g, a)
        dict_truncate(out, strategy)
    return out


# --- dense helpers -----------------------------------------------------------

_I2 = np.eye(2, dtype=complex)
_X = np.array([[0, 1], [1, 0]], dtype=complex)
_Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
_Z = np.array([[1, 0], [0, -1]], dtype=complex)
_MAT = {"I": _I2, "X": _X, "Y": _Y, "Z": _Z}


def pauli_matrix(N: int, k: Key) -> np.ndarray:
    """Matrix(PauliBasis) convention (type_PauliBasis.jl:73-95): site 1 is the
    rightmost kron factor (least significant index bit)."""
    m = np.ones((1, 1), dtype=complex)
    for j in range(N):
        m = np.kron(_MAT[letter(k[0], k[1], j)], m)
    return m


def sum_matrix(N: int, O: Dict[Key, float]) -> np.ndarray:
    out = np.zeros((2 ** N, 2 ** N), dtype=complex)
    for k, c in O.items():
        out += c * pauli_matrix(N, k)
    return out


def dense_heisenberg(N: int, M: np.ndarray, G: Key, theta: float) -> np.ndarray:
    """O -> exp(iθG/2) O exp(-iθG/2)."""
    g = pauli_matrix(N, G)
    U = math.cos(theta / 2) * np.eye(2 ** N) - 1j * math.sin(theta / 2) * g
    return U.conj().T @ M @ U


def matrix_to_paulis(N: int, M: np.ndarray, tol: float = 1e-14) -> Dict[Key, complex]:
    out = {}
    d = 2 ** N
    for z in range(2 ** N):
        for x in range(2 ** N):
            c = np.trace(pauli_matrix(N, (z, x)).conj().T @ M) / d
            if abs(c) > tol:
                out[(z, x)] = c
    return out


def apply_pauli_state(N: int, psi: np.ndarray, k: Key) -> np.ndarray:
    """P|ψ> with P = i^{-n_Y} Z^z X^x (representation.md:30-40)."""
    z, x = k
    ny = bin(z & x).count("1")
    idx = np.arange(2 ** N, dtype=np.int64)
    tgt = idx ^ x
    par = np.zeros(2 ** N, dtype=np.int64)
    t = tgt.copy()
    zz = z
    while zz:
        low = zz & -zz
        par ^= ((t & low) != 0).astype(np.int64)
        zz ^= low
    out = np.zeros_like(psi)
    out[tgt] = psi[idx] * np.where(par == 1, -1.0, 1.0)
    return out * (1j) ** (-ny)


def statevector_expectation(N: int, O: Dict[Key, float], gens: Sequence[Key], angs: Sequence[float]) -> float:
    """<0| U_n^† ... U_1^† O U_1 ... U_n |0>, U_k = exp(-iθ_k G_k / 2)."""
    psi = np.zeros(2 ** N, dtype=complex)
    psi[0] = 1.0
    for g, a in reversed(list(zip(gens, angs))):
        psi = math.cos(a / 2) * psi - 1j * math.sin(a / 2) * apply_pauli_state(N, psi, g)
    acc = 0.0 + 0j
    for k, c in O.items():
        acc += c * np.vdot(psi, apply_pauli_state(N, psi, k))
    return acc.real


# --- windowed flat-storage restatement (spv_evolve.jl:312-387), stable sort ----

def window_evolve(N: int, terms: Sequence[Tuple[Key, float]], gens: Sequence[Key], angs: Sequence[float],
                  window: int, drop_strict, drop_local, energy_ket=None):
    """Pure-Python restatement of the windowed SparsePauliVector driver with
    unlimited append capacity (no early merges) and a *stable* boundary sort:
    rotations sweep live terms then the pending appends present at entry
    (spv_evolve.jl:207-218), appending children under ``drop_local``; every
    ``window`` rotations the appends are sorted (equal keys in append order)
    and merged, live first on ties, equal appends summed left to right, with
    ``drop_strict`` on every output (spv_kernels.jl:326-363).
    drop_*(key, |c|) -> bool.  Returns (sorted [(key, c)], accumulated energy)."""
    live = sorted(terms, key=lambda t: (t[0][0], t[0][1]))
    app: List[Tuple[Key, float]] = []
    acc_e = 0.0

    def expect(items):
        s = 0.0
        for (z, x), c in items:
            if x == 0:
                s = s + (1 - 2 * (bin(z & energy_ket).count("1") & 1)) * c
        return s

    def boundary():
        nonlocal live, app, acc_e
        before = expect(live + app) if energy_ket is not None else 0.0
        srt = sorted(app, key=lambda t: (t[0][0], t[0][1]))  # stable
        out = []
        i = j = 0
        while i < len(live) or j < len(srt):
            if j >= len(srt) or (i < len(live) and not (srt[j][0] < live[i][0])):
                k, a = live[i]
                i += 1
            else:
                k, a = srt[j]
                j += 1
            while j < len(srt) and srt[j][0] == k:
                a = a + srt[j][1]
                j += 1
            if not drop_strict(k, cabs(a)):
                out.append((k, a))
        live, app = out, []
        if energy_ket is not None:
            acc_e += (expect(live) - before).real

    for i, (g, th) in enumerate(zip(gens, angs)):
        c_, s_ = math.cos(th), math.sin(th)
        new_children = []
        for seq in (live, app):
            for idx in range(len(seq)):
                p, c = seq[idx]
                if not anticommute(p, g):
                    continue
                phase, k = pauli_product(N, g, p)
                f = 1j * phase
                v = c * s_ if f.real > 0 else -(c * s_)
                seq[idx] = (p, c * c_)
                if not drop_local(k, cabs(v)):
                    new_children.append((k, v))
        app.extend(new_children)
        if (i + 1) % window == 0 or i + 1 == len(gens):
            boundary()
    return live, acc_e
