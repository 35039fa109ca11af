"""SparsePauliBatch: many independent SparsePauliVectors evolved together.

The reference evolves one vector per ``evolve!`` call (spv_evolve.jl:312-387).
A batch applies, per call, one ``evolve!`` to every vector it holds (each with
its own generators, angles and truncation), with one kernel sequence per
rotation step for all vectors, so hundreds of small circuits share the GPU.
Per vector the results are exactly those of the single-vector calls.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, PPCounters, check, dp, u64p
from .basis import PauliBasis, word_count
from .spv import _pack, _pow_table, _unpack_rows
from .truncation import NoTruncation, PPFilter, compile_filter, is_compilable

_KIND = {"depolarizing": 0, "dephasing": 1, "bit_flip": 2, "bit_phase_flip": 3, "pauli": 4}
_FILTERS = {}  # tuple of strategies (frozen, hashable) -> (all compilable, PPFilter array)


def _filter_array(strategies):
    """(every strategy compilable, the PPFilter array of compile_filter per
    vector); strategies are frozen dataclasses, so equal tuples compile to
    equal arrays and the result is kept."""
    key = tuple(strategies)
    try:
        hit = _FILTERS.get(key)
    except TypeError:  # an unhashable strategy: no caching
        hit, key = None, None
    if hit is None:
        ok = all(is_compilable(s) for s in strategies)
        arr = (PPFilter * len(strategies))(*[compile_filter(s) for s in strategies]) if ok else None
        hit = (ok, arr)
        if key is not None:
            if len(_FILTERS) > 256:
                _FILTERS.clear()
            _FILTERS[key] = hit
    return hit


class SparsePauliBatch:
    def __init__(self, Ns: Sequence[int], terms: Sequence, *, capacity_factor: float = 2.0,
                 append_factor: float = 1.0, min_capacity: int = 16):
        """``terms[b]``: list of ((z, x) or PauliBasis, coefficient) for vector b;
        capacities per vector as SparsePauliVector(::PauliSum) computes them
        (type_SparsePauliVector.jl:175-176)."""
        self.Ns = [int(n) for n in Ns]
        self.B = len(self.Ns)
        if len(terms) != self.B:
            raise DimensionMismatch("one term list per vector")
        live = []
        app = []
        for t in terms:
            lc = max(min_capacity, math.ceil(capacity_factor * max(len(t), 1)))
            live.append(lc)
            app.append(max(min_capacity, math.ceil(append_factor * lc)))
        N_arr = (ctypes.c_int * self.B)(*self.Ns)
        lc_arr = (ctypes.c_int64 * self.B)(*live)
        ac_arr = (ctypes.c_int64 * self.B)(*app)
        h = ctypes.c_void_p()
        check(_lib.lib().pp_batch_create(self.B, N_arr, lc_arr, ac_arr, ctypes.byref(h)))
        self._h = h
        w = ctypes.c_int()
        check(_lib.lib().pp_batch_words(self._h, ctypes.byref(w)))
        self.W = w.value
        for b, t in enumerate(terms):
            self._upload(b, t)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().pp_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def __len__(self):
        return self.B

    def _upload(self, b: int, terms):
        n = len(terms)
        z = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        x = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        c = np.zeros(max(n, 1), dtype=np.float64)
        for i, (p, coef) in enumerate(terms):
            pz, px = (p.z, p.x) if isinstance(p, PauliBasis) else p
            z[i] = _pack(pz, self.W)
            x[i] = _pack(px, self.W)
            c[i] = float(coef)
        check(_lib.lib().pp_batch_upload(self._h, b, u64p(z), u64p(x), dp(c), n, 0))

    def upload_arrays(self, b: int, z: np.ndarray, x: np.ndarray, c: np.ndarray, sorted_keys: bool = True):
        """Replace vector b's live set by word arrays z[n][W], x[n][W], c[n]
        (keys strictly increasing when sorted_keys)."""
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, self.W)
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, self.W)
        c = np.ascontiguousarray(c, dtype=np.float64)
        check(_lib.lib().pp_batch_upload(self._h, b, u64p(z), u64p(x), dp(c), len(c), 1 if sorted_keys else 0))
        return self

    def use_tuning(self, table):
        """Pick tile shapes before every evolve call from an autotuner table
        (pauliprop.autotune) by the current largest live size."""
        self._tuning = table
        return self

    def sizes(self) -> List[int]:
        n = (ctypes.c_int64 * self.B)()
        check(_lib.lib().pp_batch_sizes(self._h, n))
        return list(n)

    def arrays(self, b: int):
        n = self.sizes()[b]
        z = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        x = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        c = np.zeros(max(n, 1), dtype=np.float64)
        got = ctypes.c_int64()
        check(_lib.lib().pp_batch_download(self._h, b, u64p(z), u64p(x), dp(c), max(n, 1), ctypes.byref(got)))
        return z[:n], x[:n], c[:n]

    def terms(self, b: int):
        z, x, c = self.arrays(b)
        return list(zip(_unpack_rows(z), _unpack_rows(x), c.tolist()))

    def pack(self, generators: Sequence[Sequence], angles: Sequence[Sequence[float]]):
        """Step arrays for ``evolve_packed``: generator s of vector b at [s, b]."""
        if len(generators) != self.B or len(angles) != self.B:
            raise DimensionMismatch("one generator list and one angle list per vector")
        for g, a in zip(generators, angles):
            if len(g) != len(a):
                raise DimensionMismatch("generators and angles must have same length")
        nsteps = max((len(g) for g in generators), default=0)
        B, W = self.B, self.W
        gz = np.zeros((nsteps, B, W), dtype=np.uint64)
        gx = np.zeros((nsteps, B, W), dtype=np.uint64)
        cv = np.zeros((nsteps, B), dtype=np.float64)
        sv = np.zeros((nsteps, B), dtype=np.float64)
        act = np.zeros((nsteps, B), dtype=np.uint8)
        for b in range(B):
            for s, (g, a) in enumerate(zip(generators[b], angles[b])):
                z, x = (g.z, g.x) if isinstance(g, PauliBasis) else g
                gz[s, b] = _pack(z, W)
                gx[s, b] = _pack(x, W)
                cv[s, b] = math.cos(a)
                sv[s, b] = math.sin(a)
                act[s, b] = 1
        return gz, gx, cv, sv, act

    def evolve_(self, generators: Sequence[Sequence], angles: Sequence[Sequence[float]], *, truncation=NoTruncation(),
                local_truncation=NoTruncation(), counters: bool = False):
        """One ``evolve!(v_b, generators[b], angles[b]; truncation, local_truncation)``
        per vector b (window = 1).  ``truncation`` may be one strategy or one per vector."""
        return self.evolve_packed(self.pack(generators, angles), truncation=truncation,
                                  local_truncation=local_truncation, counters=counters)

    def evolve_packed(self, packed, *, truncation=NoTruncation(), local_truncation=NoTruncation(),
                      counters: bool = False):
        """``evolve_`` on arrays from ``pack`` (cos/sin already applied)."""
        strat = list(truncation) if isinstance(truncation, (list, tuple)) else [truncation] * self.B
        loc = list(local_truncation) if isinstance(local_truncation, (list, tuple)) else [local_truncation] * self.B
        ok_s, fs = _filter_array(strat)
        if not ok_s:
            s = next(s for s in strat if not is_compilable(s))
            raise NotImplementedError(f"{type(s).__name__} at merge boundaries is not available in this build")
        ok_l, fl = _filter_array(loc)
        if not ok_l:
            raise ArgumentError("local_truncation must be a deterministic (weight/coefficient) strategy "
                                "— it runs per append")
        gz, gx, cv, sv, act = packed
        nsteps = gz.shape[0]
        if nsteps == 0:
            return [PPCounters() for _ in range(self.B)] if counters else None
        B = self.B
        if getattr(self, "_tuning", None):
            from .autotune import lookup
            self.set_tiles(*lookup(self._tuning, self.W, max(self.sizes())))
        cnt = (PPCounters * B)()
        check(_lib.lib().pp_batch_evolve(self._h, nsteps, u64p(gz), u64p(gx), dp(cv), dp(sv),
                                         act.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), fs, fl, cnt))
        return list(cnt) if counters else None

    def channel_(self, kind, p=None, *, pX=0.0, pY=0.0, pZ=0.0, qubits=None):
        """The SPV channels of channels.jl:203-265 on every vector (kind is a
        channel name, or a list of names/None per vector)."""
        kinds = list(kind) if isinstance(kind, (list, tuple)) else [kind] * self.B
        ps = list(p) if isinstance(p, (list, tuple)) else [p] * self.B
        qk = qubits if qubits is None or isinstance(qubits, int) else tuple(qubits)
        key = (tuple(kinds), tuple(ps), pX, pY, pZ, qk)
        cache = self.__dict__.setdefault("_chan_cache", {})
        hit = cache.get(key)
        if hit is not None:
            karr, masks, tabs = hit
            check(_lib.lib().pp_batch_scale_by_popcount(self._h, karr, u64p(masks), dp(tabs)))
            return self
        L = _lib.lib()
        karr = (ctypes.c_int * self.B)()
        masks = np.zeros((self.B, self.W), dtype=np.uint64)
        tabs = np.zeros((self.B, 3, 129), dtype=np.float64)
        for b in range(self.B):
            k = kinds[b]
            karr[b] = -1 if k is None else _KIND[k]
            if k is None:
                continue
            N = self.Ns[b]
            if qubits is None:
                M = (1 << N) - 1
            else:
                M = 0
                for q in ([qubits] if isinstance(qubits, int) else qubits):
                    if not (1 <= q <= N):
                        raise ArgumentError(f"qubit index {q} out of range 1:{N}")
                    M |= 1 << (q - 1)
            masks[b] = _pack(M, self.W)
            if k == "pauli":
                lams = (1 - 2 * (pY + pZ), 1 - 2 * (pX + pZ), 1 - 2 * (pX + pY))
            else:
                pb = ps[b]
                if not (0 <= pb <= 1):
                    raise ArgumentError(f"{k} parameter p={pb} must be in [0,1]")
                lams = (1 - 4 * pb / 3,) if k == "depolarizing" else (1 - 2 * pb,)
            for t, lam in enumerate(lams):
                tabs[b, t, :N + 1] = _pow_table(lam, N)
        if len(cache) > 64:
            cache.clear()
        cache[key] = (karr, masks, tabs)
        check(L.pp_batch_scale_by_popcount(self._h, karr, u64p(masks), dp(tabs)))
        return self

    def expectation_values(self, kets: Optional[Sequence[int]] = None) -> List[float]:
        kets = kets or [0] * self.B
        k = np.zeros((self.B, self.W), dtype=np.uint64)
        for b, v in enumerate(kets):
            k[b] = _pack(int(v) & ((1 << (64 * self.W)) - 1), self.W)
        out = np.zeros(self.B, dtype=np.float64)
        check(_lib.lib().pp_batch_expectation_ket(self._h, u64p(k), dp(out)))
        return out.tolist()

    def reserve(self, cap):
        """Pre-size device buffers to hold ``cap`` (int or one per vector) live terms."""
        caps = list(cap) if isinstance(cap, (list, tuple)) else [int(cap)] * self.B
        arr = (ctypes.c_int64 * self.B)(*caps)
        check(_lib.lib().pp_batch_reserve(self._h, arr))
        return self

    def set_tiles(self, tile: int = 0, merge_tile: int = 0):
        """Sweep tile (512/2048) and merge tile (512/1024/2048); 0 = automatic."""
        check(_lib.lib().pp_batch_set_tiles(self._h, int(tile), int(merge_tile)))
        return self

    def set_graphs(self, on: bool):
        """Capture each evolve call's launch sequence in a hipGraph and replay it
        on later calls of the same shape."""
        check(_lib.lib().pp_batch_set_graphs(self._h, 1 if on else 0))
        return self

    def graph_stats(self):
        c = ctypes.c_int64()
        n = ctypes.c_int64()
        check(_lib.lib().pp_batch_graph_stats(self._h, ctypes.byref(c), ctypes.byref(n)))
        return c.value, n.value

    def set_resident(self, mode: int = 1, cap_limit: int = 0):
        """Resident layer launches (one kernel launch per evolve call with the
        terms in LDS): 0 off, 1 where the plan fits (default), 2 every call the
        kernel can take; cap_limit > 0 lowers the chunk capacity per
        workgroup (exercises the fallback to the per-step kernels)."""
        check(_lib.lib().pp_batch_set_resident(self._h, int(mode), int(cap_limit)))
        return self

    def set_epoch(self, epoch: int):
        """Test hook: move the merge look-back tag base forward to `epoch`
        < 2^22 - 1, to reach the tag wrap."""
        check(_lib.lib().pp_batch_set_epoch(self._h, int(epoch)))
        return self

    def resident_stats(self):
        """(resident launches, vectors finished by the per-step kernels after a
        resident stop, aborted launches)."""
        a = ctypes.c_int64()
        b = ctypes.c_int64()
        c = ctypes.c_int64()
        check(_lib.lib().pp_batch_resident_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def resident_time(self, reset: bool = False) -> float:
        """HIP-event seconds of the resident launches while timing is on."""
        s = ctypes.c_double()
        check(_lib.lib().pp_batch_resident_time(self._h, ctypes.byref(s), 1 if reset else 0))
        return s.value

    def resident_work(self, reset: bool = False):
        """(rotations, swept terms, merge outputs, chunk rebalances) of the
        resident launches."""
        r = ctypes.c_int64()
        s = ctypes.c_int64()
        o = ctypes.c_int64()
        q = ctypes.c_int64()
        check(_lib.lib().pp_batch_resident_work(self._h, ctypes.byref(r), ctypes.byref(s), ctypes.byref(o),
                                                ctypes.byref(q), 1 if reset else 0))
        return r.value, s.value, o.value, q.value

    def set_timing(self, on: bool):
        check(_lib.lib().pp_batch_set_timing(self._h, 1 if on else 0))

    def merge_time(self, reset: bool = False):
        s = ctypes.c_double()
        n = ctypes.c_int64()
        check(_lib.lib().pp_batch_merge_time(self._h, ctypes.byref(s), ctypes.byref(n), 1 if reset else 0))
        return s.value, n.value
