"""Truncation strategies and their compiled per-term filter.

Mirrors src/truncation.jl:37-159 (strategy types) and src/spv_kernels.jl:33-125
(MergeFilter, NOFILTER, _compile_filter, _combine_filters, _is_compilable).
The compiled filter crosses the C ABI as ``PPFilter``, whose first 64 bytes are
the byte layout of the reference's ``MergeFilter`` (3 x Int64, 5 x Float64).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Any, Tuple

INT_MAX = (1 << 63) - 1  # typemax(Int) sentinel


class PPFilter(ctypes.Structure):
    """Byte layout of MergeFilter (spv_kernels.jl:33-42)."""

    _fields_ = [
        ("wmax", ctypes.c_int64),
        ("xwmax", ctypes.c_int64),
        ("mwmax", ctypes.c_int64),
        ("thresh", ctypes.c_double),
        ("alpha_w", ctypes.c_double),
        ("dthresh_w", ctypes.c_double),
        ("alpha_xw", ctypes.c_double),
        ("dthresh_xw", ctypes.c_double),
    ]

    def astuple(self) -> Tuple:
        return tuple(getattr(self, f) for f, _ in self._fields_)

    def __repr__(self) -> str:
        return "PPFilter(" + ", ".join(f"{f}={getattr(self, f)!r}" for f, _ in self._fields_) + ")"


def make_filter(wmax=INT_MAX, xwmax=INT_MAX, mwmax=INT_MAX, thresh=-1.0, alpha_w=0.0,
                dthresh_w=-1.0, alpha_xw=0.0, dthresh_xw=-1.0) -> PPFilter:
    return PPFilter(int(wmax), int(xwmax), int(mwmax), float(thresh), float(alpha_w),
                    float(dthresh_w), float(alpha_xw), float(dthresh_xw))


def NOFILTER() -> PPFilter:
    """spv_kernels.jl:44-45."""
    return make_filter()


class TruncationStrategy:
    """truncation.jl:14 — abstract supertype."""


@dataclass(frozen=True)
class NoTruncation(TruncationStrategy):
    pass


@dataclass(frozen=True)
class CoeffTruncation(TruncationStrategy):
    """Drop terms with |c| <= thresh (truncation.jl:44-47; default 1e-6)."""
    thresh: float = 1e-6


@dataclass(frozen=True)
class WeightTruncation(TruncationStrategy):
    """Drop terms with Pauli weight > max_weight (truncation.jl:54-56)."""
    max_weight: int


@dataclass(frozen=True)
class XWeightTruncation(TruncationStrategy):
    max_weight: int


@dataclass(frozen=True)
class MajoranaWeightTruncation(TruncationStrategy):
    max_weight: int


@dataclass(frozen=True)
class WeightDampedTruncation(TruncationStrategy):
    """Drop |c| * exp(-alpha * weight) <= thresh (truncation.jl:84-88)."""
    alpha: float
    thresh: float = 1e-6


@dataclass(frozen=True)
class XWeightDampedTruncation(TruncationStrategy):
    alpha: float
    thresh: float = 1e-6


@dataclass(frozen=True)
class AdaptiveTruncation(TruncationStrategy):
    """truncation.jl:137-141: top-k coefficient threshold when over max_terms."""
    max_terms: int = 10000
    min_thresh: float = 1e-12


@dataclass(frozen=True)
class StochasticCoeffTruncation(TruncationStrategy):
    """truncation.jl:113-117 (Russian roulette): stochastic_clip! on the device."""
    epsilon: float
    rng: Any = None


@dataclass(frozen=True)
class StochasticSamplingTruncation(TruncationStrategy):
    """truncation.jl:125-129: exponential-key sampling of n_keep terms on the device."""
    n_keep: int
    rng: Any = None


@dataclass(frozen=True)
class CompositeTruncation(TruncationStrategy):
    """truncation.jl:155-159 — strategies applied in sequence."""
    strategies: Tuple[TruncationStrategy, ...] = field(default_factory=tuple)

    def __init__(self, *strategies):
        if len(strategies) == 1 and isinstance(strategies[0], (list, tuple)):
            strategies = tuple(strategies[0])
        object.__setattr__(self, "strategies", tuple(strategies))


_SIMPLE = (NoTruncation, CoeffTruncation, WeightTruncation, XWeightTruncation,
           MajoranaWeightTruncation, WeightDampedTruncation, XWeightDampedTruncation)


def is_compilable(s: TruncationStrategy) -> bool:
    """spv_kernels.jl:116-125."""
    if isinstance(s, _SIMPLE):
        return True
    if isinstance(s, CompositeTruncation):
        if not all(is_compilable(t) for t in s.strategies):
            return False
        if sum(isinstance(t, WeightDampedTruncation) for t in s.strategies) > 1:
            return False
        if sum(isinstance(t, XWeightDampedTruncation) for t in s.strategies) > 1:
            return False
        return True
    return False


def _combine(f1: PPFilter, f2: PPFilter) -> PPFilter:
    """spv_kernels.jl:98-108."""
    if f1.dthresh_w >= 0 and f2.dthresh_w >= 0:
        raise RuntimeError("cannot combine two weight-damped truncations in one composite")
    if f1.dthresh_xw >= 0 and f2.dthresh_xw >= 0:
        raise RuntimeError("cannot combine two x-weight-damped truncations in one composite")
    dw = f1 if f1.dthresh_w >= 0 else f2
    dxw = f1 if f1.dthresh_xw >= 0 else f2
    return make_filter(min(f1.wmax, f2.wmax), min(f1.xwmax, f2.xwmax), min(f1.mwmax, f2.mwmax),
                       max(f1.thresh, f2.thresh), dw.alpha_w, dw.dthresh_w, dxw.alpha_xw,
                       dxw.dthresh_xw)


def compile_filter(s: TruncationStrategy) -> PPFilter:
    """spv_kernels.jl:81-111."""
    if isinstance(s, NoTruncation):
        return NOFILTER()
    if isinstance(s, CoeffTruncation):
        return make_filter(thresh=float(s.thresh))
    if isinstance(s, WeightTruncation):
        return make_filter(wmax=int(s.max_weight))
    if isinstance(s, XWeightTruncation):
        return make_filter(xwmax=int(s.max_weight))
    if isinstance(s, MajoranaWeightTruncation):
        return make_filter(mwmax=int(s.max_weight))
    if isinstance(s, WeightDampedTruncation):
        return make_filter(alpha_w=float(s.alpha), dthresh_w=float(s.thresh))
    if isinstance(s, XWeightDampedTruncation):
        return make_filter(alpha_xw=float(s.alpha), dthresh_xw=float(s.thresh))
    if isinstance(s, CompositeTruncation):
        f = NOFILTER()
        for t in s.strategies:
            f = _combine(f, compile_filter(t))
        return f
    raise RuntimeError(f"{type(s).__name__} cannot be compiled to a MergeFilter")


class CorrectionAccumulator:
    """truncation.jl:25 — abstract supertype."""


class NoCorrection(CorrectionAccumulator):
    pass


class EnergyCorrection(CorrectionAccumulator):
    """truncation.jl:262-266: accumulates the change of ⟨ψ|O|ψ⟩ caused by
    every truncation (after - before, truncation.jl:304-306)."""

    def __init__(self, psi, accumulated_energy: float = 0.0):
        self.psi = psi
        self.accumulated_energy = float(accumulated_energy)

    def ket_int(self) -> int:
        v = getattr(self.psi, "v", self.psi)
        return int(v)


class EnergyVarianceCorrection(CorrectionAccumulator):
    """truncation.jl:273-278: accumulates the changes of ⟨ψ|O|ψ⟩ and of
    variance(O, ψ) caused by every truncation, both measured on the merged
    state (spv_evolve.jl:246-270)."""

    def __init__(self, psi, accumulated_energy: float = 0.0, accumulated_variance: float = 0.0):
        self.psi = psi
        self.accumulated_energy = float(accumulated_energy)
        self.accumulated_variance = float(accumulated_variance)

    def ket_int(self) -> int:
        v = getattr(self.psi, "v", self.psi)
        return int(v)


def strategy_steps(s: TruncationStrategy) -> list:
    """A strategy as the ordered steps of its _apply! (spv_evolve.jl:59-119;
    CompositeTruncation applies its members in order, truncation.jl:236-243):
    ("filter", PPFilter) | ("adaptive", max_terms, min_thresh) |
    ("stochastic_coeff", epsilon, rng) | ("stochastic_sampling", n_keep, rng)."""
    if is_compilable(s):
        return [("filter", compile_filter(s))]
    if isinstance(s, AdaptiveTruncation):
        return [("adaptive", int(s.max_terms), float(s.min_thresh))]
    if isinstance(s, StochasticCoeffTruncation):
        return [("stochastic_coeff", float(s.epsilon), s.rng)]
    if isinstance(s, StochasticSamplingTruncation):
        return [("stochastic_sampling", int(s.n_keep), s.rng)]
    if isinstance(s, CompositeTruncation):
        out = []
        for t in s.strategies:
            out.extend(strategy_steps(t))
        return out
    raise NotImplementedError(f"{type(s).__name__} is not available on the device path in this build")
