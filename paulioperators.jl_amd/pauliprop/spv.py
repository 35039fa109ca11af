"""SparsePauliVector on an MI355X: the drop-in for PauliOperators.jl's SPV path.

Each method mirrors a reference method (cited) with the same argument meaning
and error behaviour; Julia's ``f!`` is spelled ``f_``.  Coefficients are
Float64 (``T=float``, what the workloads use) or ComplexF64 (``T=complex``,
the reference's default element type, type_SparsePauliVector.jl:138,167).
The live terms stay in device memory between calls.
"""
from __future__ import annotations

import ctypes
import math
from typing import Iterable, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, PPCounters, check, dp, u64p
from .basis import Ket, PauliBasis, PauliSum, word_count
from .truncation import (AdaptiveTruncation, CoeffTruncation, CompositeTruncation, CorrectionAccumulator,
                         EnergyCorrection, EnergyVarianceCorrection, MajoranaWeightTruncation, NoCorrection,
                         NoTruncation,
                         StochasticCoeffTruncation, strategy_steps,
                         StochasticSamplingTruncation, TruncationStrategy, WeightDampedTruncation,
                         WeightTruncation, XWeightDampedTruncation, XWeightTruncation, compile_filter,
                         is_compilable, NOFILTER)

M64 = (1 << 64) - 1


def _rng_seed(rng) -> int:
    """A 64-bit key for the device's counter-based generator: an int is used
    as given, a numpy Generator / random.Random draws one, None draws from
    the OS (the reference's Random.default_rng() default)."""
    if rng is None:
        return int.from_bytes(__import__("os").urandom(8), "little")
    if isinstance(rng, (int, np.integer)):
        return int(rng) & M64
    if hasattr(rng, "integers"):
        return int(rng.integers(0, 1 << 63)) | (int(rng.integers(0, 2)) << 63)
    return int(rng.getrandbits(64))


def _strategy_ops(strategy):
    """pp_strategy_op array for a strategy's _apply! steps."""
    steps = strategy_steps(strategy)
    ops = (_lib.PPStrategyOp * max(len(steps), 1))()
    for i, st in enumerate(steps):
        if st[0] == "filter":
            ops[i].kind = _lib.PP_OP_FILTER
            ops[i].filter = st[1]
        elif st[0] == "adaptive":
            ops[i].kind = _lib.PP_OP_ADAPTIVE
            ops[i].max_terms = st[1]
            ops[i].value = st[2]
        elif st[0] == "stochastic_coeff":
            ops[i].kind = _lib.PP_OP_STOCHASTIC_COEFF
            ops[i].value = st[1]
            ops[i].seed = _rng_seed(st[2])
        else:
            ops[i].kind = _lib.PP_OP_STOCHASTIC_SAMPLING
            ops[i].max_terms = st[1]
            ops[i].seed = _rng_seed(st[2])
    return ops, len(steps)


def _energy_args(correction, W):
    """(ket words, energy accumulator, variance accumulator or None)."""
    if isinstance(correction, NoCorrection):
        return None, None, None
    if isinstance(correction, (EnergyCorrection, EnergyVarianceCorrection)):
        ket = _pack(correction.ket_int() & ((1 << (64 * W)) - 1), W)
        var = ctypes.c_double(correction.accumulated_variance) if isinstance(correction, EnergyVarianceCorrection) \
            else None
        return ket, ctypes.c_double(correction.accumulated_energy), var
    raise ArgumentError(f"{type(correction).__name__} has no device implementation")


def _pack(v: int, W: int) -> np.ndarray:
    return np.array([(v >> (64 * k)) & M64 for k in range(W)], dtype=np.uint64)


def _unpack_rows(a: np.ndarray) -> List[int]:
    if a.shape[1] == 1:
        return [int(v) for v in a[:, 0]]
    return [int(lo) | (int(hi) << 64) for lo, hi in a]


_POW_TABLES = {}  # (lambda, N) -> [lambda^0 .. lambda^N] by Julia's Float64^Int (pp_julia_powi)


def _pow_table(lam: float, N: int) -> np.ndarray:
    """The per-weight channel factors, computed once per (lambda, N); the
    caller must not modify the returned array."""
    key = (float(lam), int(N))
    t = _POW_TABLES.get(key)
    if t is None:
        if len(_POW_TABLES) > 4096:
            _POW_TABLES.clear()
        L = _lib.lib()
        t = np.array([L.pp_julia_powi(lam, j) for j in range(N + 1)], dtype=np.float64)
        t.setflags(write=False)
        _POW_TABLES[key] = t
    return t


class SparsePauliVector:
    """SparsePauliVector{N,W,T} (type_SparsePauliVector.jl:69-84), T = Float64
    or ComplexF64, device-resident."""

    def __init__(self, N: int, terms=None, *, T=float, capacity_factor: float = 2.0, append_factor: float = 1.0,
                 min_capacity: int = 16, imag_tol: float = 1e-10, _caps: Optional[Tuple[int, int]] = None):
        """``SparsePauliVector(O::PauliSum; T, capacity_factor, append_factor, min_capacity, imag_tol)``
        (type_SparsePauliVector.jl:167-188).  ``terms`` is a PauliSum/dict
        {PauliBasis: coeff}, a list of (PauliBasis, coeff), or a tuple of word
        arrays (z, x, c) already in key order.  ``T`` is ``float`` (Float64) or
        ``complex`` (ComplexF64)."""
        if capacity_factor < 1:
            raise ArgumentError("capacity_factor must be >= 1")
        if T not in (float, complex):
            raise ArgumentError("T must be float (Float64) or complex (ComplexF64)")
        self.N = int(N)
        self.W = word_count(self.N)
        self.T = T
        self.complex = T is complex
        self.dtype = np.complex128 if self.complex else np.float64
        arrays = None
        if isinstance(terms, tuple) and len(terms) == 3 and isinstance(terms[0], np.ndarray):
            arrays = terms
            nterms = len(terms[2])
            presorted = True
        else:
            items = list(terms.items()) if isinstance(terms, dict) else list(terms or [])
            nterms = len(items)
            presorted = False
        live_cap = max(min_capacity, math.ceil(capacity_factor * max(nterms, 1)))
        app_cap = max(min_capacity, math.ceil(append_factor * live_cap))
        if _caps is not None:
            live_cap, app_cap = _caps
        h = ctypes.c_void_p()
        check(_lib.lib().pp_create_typed(self.N, live_cap, app_cap, 1 if self.complex else 0, ctypes.byref(h)))
        self._h = h
        if arrays is None:
            z = np.zeros((max(nterms, 1), self.W), dtype=np.uint64)
            x = np.zeros((max(nterms, 1), self.W), dtype=np.uint64)
            c = np.zeros(max(nterms, 1), dtype=self.dtype)
            for i, (p, coef) in enumerate(items):
                if isinstance(p, PauliBasis):
                    if p.N != self.N:
                        raise ArgumentError("qubit count mismatch")
                    pz, px = p.z, p.x
                else:
                    pz, px = p
                cc = complex(coef)
                # _engine_coeff (type_SparsePauliVector.jl:143-149)
                if not self.complex and abs(cc.imag) > imag_tol * max(1.0, abs(cc)):
                    raise ValueError(f"coefficient {coef} has an imaginary part; Float64 storage needs real input")
                z[i] = _pack(pz, self.W)
                x[i] = _pack(px, self.W)
                c[i] = cc if self.complex else cc.real
            arrays = (z[:nterms], x[:nterms], c[:nterms])
        if nterms:
            z, x = (np.ascontiguousarray(a) for a in arrays[:2])
            c = np.ascontiguousarray(arrays[2], dtype=self.dtype)
            check(_lib.lib().pp_upload(self._h, u64p(z), u64p(x), dp(c), nterms, 1 if presorted else 0))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().pp_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_resident(self, mode: int = 1, cap_limit: int = 0):
        """Resident layer launches for evolve_ at window = 1 (see
        SparsePauliBatch.set_resident)."""
        check(_lib.lib().pp_set_resident(self._h, int(mode), int(cap_limit)))
        return self

    def set_epoch(self, epoch: int):
        """Test hook: move the merge look-back tag base (sequence and windowed
        merges) forward to `epoch` < 2^22 - 1, to reach the tag wrap."""
        check(_lib.lib().pp_set_epoch(self._h, int(epoch)))
        return self

    # --- size / conversion ------------------------------------------------------------
    def __len__(self) -> int:
        n = ctypes.c_int64()
        check(_lib.lib().pp_size(self._h, ctypes.byref(n)))
        return n.value

    def capacities(self) -> Tuple[int, int]:
        a = ctypes.c_int64()
        b = ctypes.c_int64()
        check(_lib.lib().pp_capacities(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def arrays(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Live (z, x, c) in key order: z, x as (n, W) uint64, c as float64
        (complex128 for ComplexF64)."""
        n = len(self)
        z = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        x = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        c = np.zeros(max(n, 1), dtype=self.dtype)
        got = ctypes.c_int64()
        check(_lib.lib().pp_download(self._h, u64p(z), u64p(x), dp(c), max(n, 1), ctypes.byref(got)))
        return z[:n], x[:n], c[:n]

    def terms(self) -> List[Tuple[int, int, float]]:
        z, x, c = self.arrays()
        return list(zip(_unpack_rows(z), _unpack_rows(x), c.tolist()))

    def to_paulisum(self) -> PauliSum:
        """``PauliSum(v)`` (type_SparsePauliVector.jl:212-221)."""
        out = PauliSum(self.N)
        for z, x, c in self.terms():
            out[PauliBasis(self.N, z, x)] = c
        return out

    def copy(self) -> "SparsePauliVector":
        """``copy(v)`` (type_SparsePauliVector.jl:249-254), device to device."""
        out = SparsePauliVector(self.N, None, T=self.T, _caps=self.capacities())
        check(_lib.lib().pp_copy(out._h, self._h))
        return out

    # --- spv_ops.jl:520-587 ------------------------------------------------------------
    def largest(self) -> "SparsePauliVector":
        """``largest(v)`` (spv_ops.jl:520-536): the term with the largest |c| (the
        first such in key order) as a one-term vector."""
        z, x, c, _ = self._largest(0)
        out = SparsePauliVector(self.N, None, T=self.T, _caps=(16, 16))
        out[PauliBasis(self.N, z, x)] = c
        return out

    def largest_diag(self):
        """``largest_diag(v)`` (spv_ops.jl:538-553): (PauliBasis, c) of the
        diagonal (x == 0) term with the largest |c|."""
        z, x, c, _ = self._largest(1)
        return PauliBasis(self.N, z, x), c

    def _largest(self, diag_only: int):
        zw = np.zeros(self.W, dtype=np.uint64)
        xw = np.zeros(self.W, dtype=np.uint64)
        cw = np.zeros(1, dtype=self.dtype)
        idx = ctypes.c_int64()
        check(_lib.lib().pp_largest(self._h, diag_only, u64p(zw), u64p(xw), dp(cw), ctypes.byref(idx)))
        return _unpack_rows(zw[None, :])[0], _unpack_rows(xw[None, :])[0], self.T(cw[0]), idx.value

    def diag(self) -> "SparsePauliVector":
        """``diag(v)`` (spv_ops.jl:571-587): a copy with the x == 0 terms."""
        out = self.copy()
        check(_lib.lib().pp_select_diag(out._h, 0))
        return out

    def offdiag(self) -> "SparsePauliVector":
        """``offdiag(v)`` (spv_ops.jl:555-569): a copy with the x != 0 terms."""
        out = self.copy()
        check(_lib.lib().pp_select_diag(out._h, 1))
        return out

    def matrix(self) -> np.ndarray:
        """``Matrix(v)`` (spv_ops.jl:109-115) for N <= 13: the dense 2^N x 2^N
        matrix (site 1 the least significant index bit), float64 for Float64
        coefficients (an imaginary entry raises, as Julia's InexactError),
        complex128 for ComplexF64."""
        if self.N > 13:
            raise ArgumentError("Matrix(v) is limited to N <= 13")
        D = 1 << self.N
        buf = np.zeros(D * D, dtype=self.dtype)
        check(_lib.lib().pp_matrix(self._h, dp(buf)))
        return buf.reshape(D, D).T.copy()  # column-major entries

    def check(self) -> bool:
        """``check_spv`` key-order invariant (type_SparsePauliVector.jl:263-278)."""
        ok = ctypes.c_int()
        check(_lib.lib().pp_check(self._h, ctypes.byref(ok)))
        if not ok.value:
            raise RuntimeError("live keys not strictly increasing")
        return True

    def _key_arrays(self, keys):
        z = np.zeros((max(len(keys), 1), self.W), dtype=np.uint64)
        x = np.zeros((max(len(keys), 1), self.W), dtype=np.uint64)
        for i, p in enumerate(keys):
            pz, px = (p.z, p.x) if isinstance(p, PauliBasis) else p
            z[i] = _pack(pz, self.W)
            x[i] = _pack(px, self.W)
        return z, x

    def lookup(self, keys):
        """Coefficients of many keys at once by device binary search
        (type_SparsePauliVector.jl:286-299 _find): (values, found) arrays."""
        keys = list(keys)
        z, x = self._key_arrays(keys)
        vals = np.zeros(max(len(keys), 1), dtype=self.dtype)
        found = np.zeros(max(len(keys), 1), dtype=np.int32)
        check(_lib.lib().pp_lookup(self._h, u64p(z), u64p(x), len(keys), dp(vals),
                                   found.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return vals[:len(keys)], found[:len(keys)].astype(bool)

    def __getitem__(self, p: PauliBasis) -> float:
        """``getindex`` (type_SparsePauliVector.jl:307-311): KeyError when absent."""
        if isinstance(p, str):
            p = PauliBasis.from_string(p)
        v, f = self.lookup([p])
        if not f[0]:
            raise KeyError(p)
        return self.T(v[0])

    def get(self, p: PauliBasis, default=None):
        """``get(v, p, default)`` (type_SparsePauliVector.jl:301-305)."""
        v, f = self.lookup([p])
        return self.T(v[0]) if f[0] else default

    def haskey(self, p: PauliBasis) -> bool:
        return bool(self.lookup([p])[1][0])

    def __contains__(self, p: PauliBasis) -> bool:
        return self.haskey(p)

    def setindex_(self, items):
        """``v[p] = c`` for many (key, c) pairs (type_SparsePauliVector.jl:315-335):
        present keys take the new value, absent keys are inserted; for a key
        given twice the last value wins."""
        items = list(items.items() if isinstance(items, dict) else items)
        z, x = self._key_arrays([k for k, _ in items])
        c = np.array([self._coef(v) for _, v in items] or [0.0], dtype=self.dtype)
        check(_lib.lib().pp_setindex(self._h, u64p(z), u64p(x), dp(c), len(items)))
        return self

    def __setitem__(self, p: PauliBasis, val: float):
        self.setindex_([(p, val)])

    def delete_(self, keys):
        """``delete!(v, p)`` for many keys (type_SparsePauliVector.jl:345-358)."""
        keys = list(keys)
        z, x = self._key_arrays(keys)
        check(_lib.lib().pp_delete(self._h, u64p(z), u64p(x), len(keys)))
        return self

    def __delitem__(self, p: PauliBasis):
        self.delete_([p])

    def filter_(self, f):
        """``filter!(f, v)`` (type_SparsePauliVector.jl:389-404): keeps the
        pairs (PauliBasis, c) for which the host predicate f is true, in key
        order; the rejected keys are removed on the device."""
        drop = [(z, x) for z, x, c in self.terms() if not f((PauliBasis(self.N, z, x), c))]
        return self.delete_(drop) if drop else self

    def filter(self, f) -> "SparsePauliVector":
        return self.copy().filter_(f)

    def _coef(self, v):
        """convert(T, v): a Float64 vector rejects a nonzero imaginary part
        (Julia's InexactError)."""
        if self.complex:
            return complex(v)
        cv = complex(v)
        if cv.imag != 0:
            raise ArgumentError(f"InexactError: {v} is not a Float64")
        return cv.real

    def mul_(self, a):
        """``mul!(v, a)`` (spv_ops.jl:12-17): a real a scales componentwise; a
        complex a multiplies as Complex * Complex (ComplexF64 vectors)."""
        if isinstance(a, complex):
            if not self.complex:
                self._coef(a)
                check(_lib.lib().pp_scale(self._h, float(a.real)))
            else:
                check(_lib.lib().pp_scale_complex(self._h, float(a.real), float(a.imag)))
        else:
            check(_lib.lib().pp_scale(self._h, float(a)))
        return self

    def sum_(self, other: "SparsePauliVector"):
        """``sum!(v, other)`` (spv_ops.jl:182-193): v += other."""
        if other.N != self.N:
            raise DimensionMismatch("qubit counts differ")
        if other.T is not self.T:
            raise ArgumentError("coefficient types differ")
        if other is self:
            return self.mul_(2.0)
        check(_lib.lib().pp_sum(self._h, other._h))
        return self

    def __mul__(self, a: float) -> "SparsePauliVector":
        return self.copy().mul_(a)

    __rmul__ = __mul__

    def __neg__(self) -> "SparsePauliVector":
        return self.copy().mul_(-1.0)

    def __add__(self, other: "SparsePauliVector") -> "SparsePauliVector":
        return self.copy().sum_(other)

    def __eq__(self, other) -> bool:
        """``==`` (type_SparsePauliVector.jl:380-388): same keys and values."""
        if not isinstance(other, SparsePauliVector) or other.N != self.N:
            return NotImplemented
        a, b = self.arrays(), other.arrays()
        return all(np.array_equal(p, q) for p, q in zip(a, b))

    __hash__ = None

    def norm(self, p: float = 2) -> float:
        """``norm(v, p)`` (spv_ops.jl:57-84)."""
        out = ctypes.c_double()
        check(_lib.lib().pp_norm(self._h, float(p), ctypes.byref(out)))
        return out.value

    def tr(self):
        """``tr(v)`` (spv_ops.jl:52-55): the identity coefficient times 2^N."""
        return (self.get((0, 0), self.T(0.0))) * 2.0 ** self.N

    # --- evolution -----------------------------------------------------------------------
    def _gen_arrays(self, gens: Sequence[PauliBasis]):
        L = len(gens)
        gz = np.zeros((max(L, 1), self.W), dtype=np.uint64)
        gx = np.zeros((max(L, 1), self.W), dtype=np.uint64)
        for i, g in enumerate(gens):
            if isinstance(g, PauliBasis):
                if g.N != self.N:
                    raise ArgumentError("generator qubit count mismatch")
                z, x = g.z, g.x
            else:
                z, x = g
            gz[i] = _pack(z, self.W)
            gx[i] = _pack(x, self.W)
        return gz, gx

    def evolve_(self, generators, angles, *, window: int = 1, truncation: TruncationStrategy = NoTruncation(),
                local_truncation: TruncationStrategy = NoTruncation(),
                correction: CorrectionAccumulator = NoCorrection(),
                counters: Optional[PPCounters] = None) -> "SparsePauliVector":
        """``evolve!(O, G, θ)`` (spv_evolve.jl:228-236) when given one generator and
        angle, else ``evolve!(O, generators, angles; window, truncation,
        local_truncation, correction, counters)`` (spv_evolve.jl:312-387)."""
        if isinstance(generators, PauliBasis) or (isinstance(generators, tuple) and len(generators) == 2
                                                  and isinstance(generators[0], int)):
            gz, gx = self._gen_arrays([generators])
            th = float(angles)
            check(_lib.lib().pp_evolve_single(self._h, u64p(gz), u64p(gx), math.cos(th), math.sin(th)))
            return self
        gens = list(generators)
        angs = [float(a) for a in angles]
        if len(gens) != len(angs):
            raise DimensionMismatch("generators and angles must have same length")
        if window < 1:
            raise ArgumentError("window must be >= 1")
        if not is_compilable(local_truncation):
            raise ArgumentError("local_truncation must be a deterministic (weight/coefficient) strategy "
                                "— it runs per append")
        ket, acc, vacc = _energy_args(correction, self.W)
        compiled = is_compilable(truncation)
        ops, nops = (None, 0) if compiled else _strategy_ops(truncation)
        if not gens:
            return self
        gz, gx = self._gen_arrays(gens)
        cv = np.array([math.cos(a) for a in angs], dtype=np.float64)
        sv = np.array([math.sin(a) for a in angs], dtype=np.float64)
        fs = compile_filter(truncation) if compiled else NOFILTER()
        fl = compile_filter(local_truncation)
        cnt = counters if counters is not None else PPCounters()
        if compiled and ket is None:
            check(_lib.lib().pp_evolve(self._h, u64p(gz), u64p(gx), dp(cv), dp(sv), len(gens), int(window),
                                       ctypes.byref(fs), ctypes.byref(fl), ctypes.byref(cnt)))
        else:
            check(_lib.lib().pp_evolve_ex(self._h, u64p(gz), u64p(gx), dp(cv), dp(sv), len(gens), int(window),
                                          ctypes.byref(fs), ops, nops, ctypes.byref(fl),
                                          u64p(ket) if ket is not None else None,
                                          ctypes.byref(acc) if acc is not None else None,
                                          ctypes.byref(vacc) if vacc is not None else None, ctypes.byref(cnt)))
            if acc is not None:
                correction.accumulated_energy = acc.value
            if vacc is not None:
                correction.accumulated_variance = vacc.value
        return self

    def evolve(self, generators, angles, **kw) -> "SparsePauliVector":
        """Non-mutating ``evolve`` (spv_evolve.jl:243-244, 394-396)."""
        return self.copy().evolve_(generators, angles, **kw)

    # --- truncation / clips ------------------------------------------------------------------
    def _compact(self, f):
        check(_lib.lib().pp_compact(self._h, ctypes.byref(f)))
        return self

    def coeff_clip_(self, thresh: float):
        """spv_evolve.jl:13-14 — drop |c| <= thresh."""
        return self._compact(compile_filter(CoeffTruncation(float(thresh))))

    def clip_(self, thresh: float = 1e-16):
        return self.coeff_clip_(thresh)

    def weight_clip_(self, max_weight: int):
        return self._compact(compile_filter(WeightTruncation(int(max_weight))))

    def weight_damped_clip_(self, alpha: float, thresh: float):
        return self._compact(compile_filter(WeightDampedTruncation(float(alpha), float(thresh))))

    def x_weight_clip_(self, max_weight: int):
        return self._compact(compile_filter(XWeightTruncation(int(max_weight))))

    def x_weight_damped_clip_(self, alpha: float, thresh: float):
        return self._compact(compile_filter(XWeightDampedTruncation(float(alpha), float(thresh))))

    def majorana_weight_clip_(self, max_weight: int):
        return self._compact(compile_filter(MajoranaWeightTruncation(int(max_weight))))

    def truncate_(self, strategy: TruncationStrategy, corr: CorrectionAccumulator = NoCorrection()):
        """``truncate!(v, strategy, corr)`` (truncation.jl:330-337)."""
        ket, acc, vacc = _energy_args(corr, self.W)
        if is_compilable(strategy) and ket is None:
            return self._compact(compile_filter(strategy))
        ops, nops = _strategy_ops(strategy)
        check(_lib.lib().pp_truncate(self._h, ops, nops, u64p(ket) if ket is not None else None,
                                     ctypes.byref(acc) if acc is not None else None,
                                     ctypes.byref(vacc) if vacc is not None else None))
        if acc is not None:
            corr.accumulated_energy = acc.value
        if vacc is not None:
            corr.accumulated_variance = vacc.value
        return self

    def stochastic_clip_(self, epsilon: float, rng=None):
        """``stochastic_clip!(v, ε; rng)`` (spv_evolve.jl:35-52): |c| < ε survives
        with probability |c|/ε as ε·sign(c).  The uniform variates come from a
        counter-based generator keyed by ``rng`` (an int seed, a numpy
        Generator or random.Random, or None), not from Julia's stream."""
        return self.truncate_(StochasticCoeffTruncation(float(epsilon), rng))

    # --- channels (channels.jl:203-265) ------------------------------------------------------
    def _mask(self, qubits) -> np.ndarray:
        """_qubit_mask (channels.jl:14-28)."""
        if qubits is None:
            M = (1 << self.N) - 1
        else:
            qs = [qubits] if isinstance(qubits, int) else list(qubits)
            M = 0
            for q in qs:
                if not (1 <= q <= self.N):
                    raise ArgumentError(f"qubit index {q} out of range 1:{self.N}")
                M |= 1 << (q - 1)
        return _pack(M, self.W)

    def _table(self, lam: float) -> np.ndarray:
        return _pow_table(lam, self.N)

    def _scale(self, kind: int, qubits, t0, t1=None, t2=None):
        M = self._mask(qubits)
        t1 = t1 if t1 is not None else t0
        t2 = t2 if t2 is not None else t0
        check(_lib.lib().pp_scale_by_popcount(self._h, kind, u64p(M), dp(t0), dp(t1), dp(t2)))
        return self

    def depolarizing_channel_(self, p: float, qubits=None):
        if not (0 <= p <= 1):
            raise ArgumentError(f"depolarizing parameter p={p} must be in [0,1]")
        return self._scale(0, qubits, self._table(1 - 4 * p / 3))

    def dephasing_channel_(self, p: float, qubits=None):
        if not (0 <= p <= 1):
            raise ArgumentError(f"dephasing parameter p={p} must be in [0,1]")
        return self._scale(1, qubits, self._table(1 - 2 * p))

    phase_flip_channel_ = dephasing_channel_

    def bit_flip_channel_(self, p: float, qubits=None):
        if not (0 <= p <= 1):
            raise ArgumentError(f"bit_flip parameter p={p} must be in [0,1]")
        return self._scale(2, qubits, self._table(1 - 2 * p))

    def bit_phase_flip_channel_(self, p: float, qubits=None):
        if not (0 <= p <= 1):
            raise ArgumentError(f"bit_phase_flip parameter p={p} must be in [0,1]")
        return self._scale(3, qubits, self._table(1 - 2 * p))

    def pauli_channel_(self, pX: float, pY: float, pZ: float, qubits=None):
        if not (pX >= 0 and pY >= 0 and pZ >= 0):
            raise ArgumentError(f"probabilities must be non-negative (got pX={pX}, pY={pY}, pZ={pZ})")
        if not (pX + pY + pZ <= 1 + 4 * 2.220446049250313e-16):
            raise ArgumentError(f"pX+pY+pZ must be <= 1 (got {pX + pY + pZ})")
        lX = 1 - 2 * (pY + pZ)
        lY = 1 - 2 * (pX + pZ)
        lZ = 1 - 2 * (pX + pY)
        return self._scale(4, qubits, self._table(lX), self._table(lY), self._table(lZ))

    # --- observables ------------------------------------------------------------------------------
    def expectation_value(self, psi: Union[Ket, int] = 0):
        """``expectation_value(v, ψ::Ket)`` (spv_evolve.jl:133-134): a float
        (a complex for ComplexF64)."""
        k = psi.v if isinstance(psi, Ket) else int(psi)
        ket = _pack(k & ((1 << (64 * self.W)) - 1), self.W)
        out = np.zeros(2, dtype=np.float64)
        check(_lib.lib().pp_expectation_ket(self._h, u64p(ket), dp(out)))
        return complex(out[0], out[1]) if self.complex else float(out[0])

    def variance(self, psi: Union[Ket, int] = 0) -> float:
        """``variance(v, ψ::Ket)`` (statistics.jl:6-22): ⟨ψ|v²|ψ⟩ - ⟨ψ|v|ψ⟩²."""
        k = psi.v if isinstance(psi, Ket) else int(psi)
        ket = _pack(k & ((1 << (64 * self.W)) - 1), self.W)
        out = ctypes.c_double()
        check(_lib.lib().pp_variance_ket(self._h, u64p(ket), ctypes.byref(out)))
        return out.value


# module-level spellings of the reference's free functions --------------------------------------

def evolve_(v: SparsePauliVector, generators, angles, **kw):
    return v.evolve_(generators, angles, **kw)


def evolve(v: SparsePauliVector, generators, angles, **kw):
    return v.evolve(generators, angles, **kw)


def truncate_(v: SparsePauliVector, strategy, corr=NoCorrection()):
    return v.truncate_(strategy, corr)


def norm(v: SparsePauliVector, p: float = 2) -> float:
    return v.norm(p)


def inner_product(v1: SparsePauliVector, v2: SparsePauliVector):
    """``inner_product(v1, v2)`` (spv_ops.jl:378-396): Σ over shared keys of
    conj(c1) c2 (a complex for ComplexF64)."""
    if v1.N != v2.N:
        raise DimensionMismatch("qubit counts differ")
    if v1.T is not v2.T:
        raise ArgumentError("coefficient types differ")
    out = np.zeros(2, dtype=np.float64)
    check(_lib.lib().pp_inner_product(v1._h, v2._h, dp(out)))
    return complex(out[0], out[1]) if v1.complex else float(out[0])


def otimes(v1: SparsePauliVector, v2: SparsePauliVector) -> SparsePauliVector:
    """``otimes(v1, v2)`` = v1 ⊗ v2 (spv_ops.jl:342-358): N1 + N2 qubits, v2 on
    the high sites, coefficients c1 c2."""
    if v1.T is not v2.T:
        raise ArgumentError("coefficient types differ")
    if v1.N + v2.N > 128:
        raise ArgumentError("N1 + N2 must be <= 128")
    out = SparsePauliVector(v1.N + v2.N, None, T=v1.T, _caps=(16, 16))
    check(_lib.lib().pp_otimes(out._h, v1._h, v2._h))
    return out


def osum(v1: SparsePauliVector, v2: SparsePauliVector) -> SparsePauliVector:
    """``osum(v1, v2)`` = v1 ⊗ I + I ⊗ v2 (spv_ops.jl:360-366)."""
    one = v1.T(1.0)
    I_N = SparsePauliVector(v1.N, [((0, 0), one)], T=v1.T, _caps=(16, 16))
    I_M = SparsePauliVector(v2.N, [((0, 0), one)], T=v2.T, _caps=(16, 16))
    return otimes(v1, I_M) + otimes(I_N, v2)


def largest(v: SparsePauliVector) -> SparsePauliVector:
    return v.largest()


def largest_diag(v: SparsePauliVector):
    return v.largest_diag()


def diag(v: SparsePauliVector) -> SparsePauliVector:
    return v.diag()


def offdiag(v: SparsePauliVector) -> SparsePauliVector:
    return v.offdiag()


def Matrix(v: SparsePauliVector) -> np.ndarray:
    return v.matrix()


def isapprox(v1: SparsePauliVector, v2: SparsePauliVector, atol: float = 1e-14, rtol: float = 0.0) -> bool:
    """``isapprox(v1, v2; atol, rtol)`` (spv_ops.jl:88-107): the L2 norm of the
    difference over the union of keys against atol + rtol max(norm)."""
    if v1.N != v2.N:
        raise DimensionMismatch("qubit counts differ")
    if v1.T is not v2.T:
        raise ArgumentError("coefficient types differ")
    out = ctypes.c_double()
    check(_lib.lib().pp_diff_norm(v1._h, v2._h, ctypes.byref(out)))
    return out.value <= atol + rtol * max(v1.norm(), v2.norm())


def stochastic_clip_(v: SparsePauliVector, epsilon: float, rng=None):
    return v.stochastic_clip_(epsilon, rng)


def expectation_value(v: SparsePauliVector, psi) -> float:
    return v.expectation_value(psi)


def variance(v: SparsePauliVector, psi) -> float:
    return v.variance(psi)


def depolarizing_channel_(v, p, qubits=None):
    return v.depolarizing_channel_(p, qubits)


def dephasing_channel_(v, p, qubits=None):
    return v.dephasing_channel_(p, qubits)


phase_flip_channel_ = dephasing_channel_


def bit_flip_channel_(v, p, qubits=None):
    return v.bit_flip_channel_(p, qubits)


def bit_phase_flip_channel_(v, p, qubits=None):
    return v.bit_phase_flip_channel_(p, qubits)


def pauli_channel_(v, pX, pY, pZ, qubits=None):
    return v.pauli_channel_(pX, pY, pZ, qubits)


def depolarizing_channel(v, p, qubits=None):
    return v.copy().depolarizing_channel_(p, qubits)


def dephasing_channel(v, p, qubits=None):
    return v.copy().dephasing_channel_(p, qubits)


def bit_flip_channel(v, p, qubits=None):
    return v.copy().bit_flip_channel_(p, qubits)


def bit_phase_flip_channel(v, p, qubits=None):
    return v.copy().bit_phase_flip_channel_(p, qubits)


def pauli_channel(v, pX, pY, pZ, qubits=None):
    return v.copy().pauli_channel_(pX, pY, pZ, qubits)


def depolarizing_p_for_weight_decay(rate: float, dt: float) -> float:
    """channels.jl:281-282."""
    return (3 / 4) * (1 - math.exp(-rate * dt))
