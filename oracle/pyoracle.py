"""ctypes front-end of the CPU oracle (oracle/spv_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  The methods
mirror the reference's SparsePauliVector entry points
(src/type_SparsePauliVector.jl:167-188, src/spv_evolve.jl:13-26,133-134,
228-236,312-387, src/channels.jl:203-265).
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "_build", "libpp_oracle.so")
OR_EXPTAB = 257


class OrFilter(ctypes.Structure):
    _fields_ = [
        ("wmax", ctypes.c_int64), ("xwmax", ctypes.c_int64), ("mwmax", ctypes.c_int64),
        ("thresh", ctypes.c_double), ("alpha_w", ctypes.c_double), ("dthresh_w", ctypes.c_double),
        ("alpha_xw", ctypes.c_double), ("dthresh_xw", ctypes.c_double),
        ("exptab_w", ctypes.c_double * OR_EXPTAB), ("exptab_xw", ctypes.c_double * OR_EXPTAB),
    ]


class OrCounters(ctypes.Structure):
    _fields_ = [("rotations", ctypes.c_int64), ("terms_swept", ctypes.c_int64),
                ("terms_created", ctypes.c_int64), ("merge_in", ctypes.c_int64),
                ("merge_out", ctypes.c_int64), ("early_merges", ctypes.c_int64)]


_lib = None


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = ctypes.CDLL(_LIBPATH)
        P = ctypes.POINTER
        u64p = P(ctypes.c_uint64)
        dp = P(ctypes.c_double)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64]
        L.or_create_typed.restype = ctypes.c_void_p
        L.or_create_typed.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        L.or_expectation_c.argtypes = [ctypes.c_void_p, u64p, dp]
        L.or_hypot.restype = ctypes.c_double
        L.or_hypot.argtypes = [ctypes.c_double, ctypes.c_double]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_load.argtypes = [ctypes.c_void_p, u64p, u64p, dp, ctypes.c_int64]
        L.or_size.restype = ctypes.c_int64
        L.or_size.argtypes = [ctypes.c_void_p]
        L.or_pending.restype = ctypes.c_int64
        L.or_pending.argtypes = [ctypes.c_void_p]
        L.or_capacities.argtypes = [ctypes.c_void_p, P(ctypes.c_int64), P(ctypes.c_int64)]
        L.or_get.argtypes = [ctypes.c_void_p, u64p, u64p, dp]
        L.or_evolve_single.argtypes = [ctypes.c_void_p, u64p, u64p, ctypes.c_double, ctypes.c_double]
        L.or_evolve.argtypes = [ctypes.c_void_p, u64p, u64p, dp, dp, ctypes.c_int64, ctypes.c_int64,
                                P(OrFilter), ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                                P(OrFilter), P(OrCounters)]
        L.or_evolve_ex.argtypes = [ctypes.c_void_p, u64p, u64p, dp, dp, ctypes.c_int64, ctypes.c_int64,
                                   P(OrFilter), ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                                   P(OrFilter), u64p, P(ctypes.c_double), P(ctypes.c_double), P(OrCounters)]
        L.or_variance.restype = ctypes.c_double
        L.or_variance.argtypes = [ctypes.c_void_p, u64p]
        L.or_set_sort_mode.argtypes = [ctypes.c_int]
        L.or_compact.argtypes = [ctypes.c_void_p, P(OrFilter)]
        L.or_adaptive.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double]
        L.or_channel.argtypes = [ctypes.c_void_p, ctypes.c_int, u64p, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double]
        L.or_expectation.restype = ctypes.c_double
        L.or_expectation.argtypes = [ctypes.c_void_p, u64p]
        L.or_majorana_weight.restype = ctypes.c_int
        L.or_majorana_weight.argtypes = [ctypes.c_int, u64p, u64p]
        L.or_should_drop.restype = ctypes.c_int
        L.or_should_drop.argtypes = [ctypes.c_int, P(OrFilter), u64p, u64p, ctypes.c_double]
        L.or_check.restype = ctypes.c_int
        L.or_check.argtypes = [ctypes.c_void_p]
        L.or_julia_powi.restype = ctypes.c_double
        L.or_julia_powi.argtypes = [ctypes.c_double, ctypes.c_int64]
        L.or_filter_make.argtypes = [P(OrFilter), ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double]
        _lib = L
    return _lib


def set_sort_mode(stable: bool):
    """Boundary sort with more than one contributing rotation: the reference's
    quicksort (False) or a stable merge sort (True, the device's order)."""
    lib().or_set_sort_mode(1 if stable else 0)


def _u64p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def words(N: int) -> int:
    return 1 if N <= 64 else 2


def pack_int(v: int, W: int) -> np.ndarray:
    return np.array([(v >> (64 * k)) & ((1 << 64) - 1) for k in range(W)], dtype=np.uint64)


def pack_terms(N: int, terms: Sequence[Tuple[int, int, float]], dtype=np.float64):
    """(z, x, c) triples with Python-int bitstrings -> word arrays; c as
    float64, or complex128 (interleaved re, im) for ComplexF64."""
    W = words(N)
    n = len(terms)
    z = np.zeros((max(n, 1), W), dtype=np.uint64)
    x = np.zeros((max(n, 1), W), dtype=np.uint64)
    c = np.zeros(max(n, 1), dtype=dtype)
    m = (1 << 64) - 1
    for i, (zi, xi, ci) in enumerate(terms):
        for k in range(W):
            z[i, k] = (zi >> (64 * k)) & m
            x[i, k] = (xi >> (64 * k)) & m
        c[i] = ci
    return z, x, c


def unpack_words(a: np.ndarray) -> List[int]:
    W = a.shape[1]
    out = []
    for row in a:
        v = 0
        for k in range(W):
            v |= int(row[k]) << (64 * k)
        out.append(v)
    return out


def filter_from(f) -> OrFilter:
    """Build the oracle filter (with its exp tables) from a MergeFilter-layout struct."""
    of = OrFilter()
    lib().or_filter_make(ctypes.byref(of), f.wmax, f.xwmax, f.mwmax, f.thresh, f.alpha_w,
                         f.dthresh_w, f.alpha_xw, f.dthresh_xw)
    return of


def julia_powi(x: float, n: int) -> float:
    return lib().or_julia_powi(float(x), int(n))


def julia_hypot(x: float, y: float) -> float:
    """The oracle's restatement of Julia's hypot (abs(::ComplexF64))."""
    return lib().or_hypot(float(x), float(y))


class OracleSPV:
    """CPU restatement of SparsePauliVector{N,W,T}, T = Float64 (``T=float``)
    or ComplexF64 (``T=complex``)."""

    def __init__(self, N: int, terms: Iterable[Tuple[int, int, float]] = (),
                 capacity_factor: float = 2.0, append_factor: float = 1.0, min_capacity: int = 16,
                 T=float):
        terms = list(terms)
        self.N = N
        self.W = words(N)
        self.complex = T is complex
        self.dtype = np.complex128 if self.complex else np.float64
        # type_SparsePauliVector.jl:175-176
        live_cap = max(min_capacity, math.ceil(capacity_factor * max(len(terms), 1)))
        app_cap = max(min_capacity, math.ceil(append_factor * live_cap))
        self.h = lib().or_create_typed(N, live_cap, app_cap, 1 if self.complex else 0)
        if not self.h:
            raise ValueError("oracle create failed")
        if terms:
            z, x, c = pack_terms(N, terms, self.dtype)
            lib().or_load(self.h, _u64p(z), _u64p(x), _dp(c), len(terms))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().or_destroy(h)
            self.h = None

    def __len__(self):
        return lib().or_size(self.h)

    def capacities(self):
        a = ctypes.c_int64()
        b = ctypes.c_int64()
        lib().or_capacities(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def arrays(self):
        n = len(self)
        z = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        x = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        c = np.zeros(max(n, 1), dtype=self.dtype)
        lib().or_get(self.h, _u64p(z), _u64p(x), _dp(c))
        return z[:n], x[:n], c[:n]

    def terms(self):
        z, x, c = self.arrays()
        return list(zip(unpack_words(z), unpack_words(x), c.tolist()))

    def check(self) -> bool:
        return bool(lib().or_check(self.h))

    def evolve_single(self, gz: int, gx: int, theta: float):
        rc = lib().or_evolve_single(self.h, _u64p(pack_int(gz, self.W)), _u64p(pack_int(gx, self.W)),
                                    math.cos(theta), math.sin(theta))
        if rc:
            raise RuntimeError(f"oracle evolve_single rc={rc}")

    def evolve(self, gens: Sequence[Tuple[int, int]], angles: Sequence[float], window: int = 1,
               strict=None, adaptive: Optional[Tuple[int, float]] = None, local=None,
               cosv=None, sinv=None, energy: Optional[list] = None,
               variance: Optional[list] = None) -> OrCounters:
        """spv_evolve.jl:312-387.  ``energy``: a one-element list [ket] or
        [ket, accumulated]; on return energy[1] holds EnergyCorrection's
        accumulated_energy (truncation.jl:262-310).  With ``variance`` (a list,
        [] or [accumulated]) as well it is EnergyVarianceCorrection: on return
        variance[0] holds accumulated_variance."""
        from_struct = lambda f: filter_from(f) if f is not None else None
        L = len(gens)
        gz = np.zeros((max(L, 1), self.W), dtype=np.uint64)
        gx = np.zeros((max(L, 1), self.W), dtype=np.uint64)
        for i, (z, x) in enumerate(gens):
            gz[i] = pack_int(z, self.W)
            gx[i] = pack_int(x, self.W)
        cv = np.array(cosv if cosv is not None else [math.cos(a) for a in angles] or [0.0], dtype=np.float64)
        sv = np.array(sinv if sinv is not None else [math.sin(a) for a in angles] or [0.0], dtype=np.float64)
        nof = OrFilter()
        lib().or_filter_make(ctypes.byref(nof), (1 << 63) - 1, (1 << 63) - 1, (1 << 63) - 1, -1.0, 0.0, -1.0, 0.0, -1.0)
        fs = from_struct(strict) or nof
        fl = from_struct(local) or nof
        cnt = OrCounters()
        kind = 1 if adaptive is not None else 0
        mt, mth = adaptive if adaptive is not None else (0, 0.0)
        if energy is None:
            rc = lib().or_evolve(self.h, _u64p(gz), _u64p(gx), _dp(cv), _dp(sv), L, window,
                                 ctypes.byref(fs), kind, mt, mth, ctypes.byref(fl), ctypes.byref(cnt))
        else:
            ket = pack_int(int(energy[0]), 2)
            acc = ctypes.c_double(energy[1] if len(energy) > 1 else 0.0)
            vacc = ctypes.c_double(variance[0] if variance else 0.0)
            rc = lib().or_evolve_ex(self.h, _u64p(gz), _u64p(gx), _dp(cv), _dp(sv), L, window,
                                    ctypes.byref(fs), kind, mt, mth, ctypes.byref(fl), _u64p(ket),
                                    ctypes.byref(acc), ctypes.byref(vacc) if variance is not None else None,
                                    ctypes.byref(cnt))
            del energy[1:]
            energy.append(acc.value)
            if variance is not None:
                variance[:] = [vacc.value]
        if rc:
            raise RuntimeError(f"oracle evolve rc={rc}")
        return cnt

    def compact(self, f):
        of = filter_from(f)
        lib().or_compact(self.h, ctypes.byref(of))

    def adaptive(self, max_terms: int, min_thresh: float):
        lib().or_adaptive(self.h, max_terms, min_thresh)

    _KINDS = {"depolarizing": 0, "dephasing": 1, "bit_flip": 2, "bit_phase_flip": 3, "pauli": 4}

    def channel(self, kind: str, p=None, pX=0.0, pY=0.0, pZ=0.0, mask: Optional[int] = None):
        """channels.jl:203-265; λ formulas exactly as written there."""
        M = mask if mask is not None else (1 << self.N) - 1
        lam = 0.0
        lamX = lamY = lamZ = 0.0
        if kind == "depolarizing":
            lam = 1 - 4 * p / 3
        elif kind in ("dephasing", "bit_flip", "bit_phase_flip"):
            lam = 1 - 2 * p
        elif kind == "pauli":
            lamX = 1 - 2 * (pY + pZ)
            lamY = 1 - 2 * (pX + pZ)
            lamZ = 1 - 2 * (pX + pY)
        lib().or_channel(self.h, self._KINDS[kind], _u64p(pack_int(M, self.W)), lam, lamX, lamY, lamZ)

    def expectation(self, ket: int = 0):
        """⟨ket|O|ket⟩: a float for Float64, a complex for ComplexF64."""
        if not self.complex:
            return lib().or_expectation(self.h, _u64p(pack_int(ket, self.W)))
        out = np.zeros(2, dtype=np.float64)
        lib().or_expectation_c(self.h, _u64p(pack_int(ket, self.W)), _dp(out))
        return complex(out[0], out[1])

    def variance(self, ket: int = 0) -> float:
        """statistics.jl:6-22 on the merged live set."""
        return lib().or_variance(self.h, _u64p(pack_int(ket, self.W)))


def majorana_weight_bits(N: int, z: int, x: int) -> int:
    W = words(N)
    return lib().or_majorana_weight(W, _u64p(pack_int(z, W)), _u64p(pack_int(x, W)))


def should_drop(N: int, f, z: int, x: int, absc: float) -> bool:
    W = words(N)
    of = filter_from(f)
    return bool(lib().or_should_drop(W, ctypes.byref(of), _u64p(pack_int(z, W)), _u64p(pack_int(x, W)), absc))
