"""pauliprop — MI355X-native SparsePauliVector evolution (host side).

The Python mirror of PauliOperators.jl's SparsePauliVector API.  Julia's
mutating `f!` names become `f_` here (PyTorch's in-place convention).
"""
from .basis import (Ket, Pauli, PauliBasis, PauliSum, commute, majorana_weight, weight,
                    word_count, x_weight)
from .truncation import (AdaptiveTruncation, CoeffTruncation, CompositeTruncation,
                         EnergyCorrection, EnergyVarianceCorrection, MajoranaWeightTruncation,
                         NoCorrection, NoTruncation, PPFilter,
                         StochasticCoeffTruncation, StochasticSamplingTruncation,
                         TruncationStrategy, WeightDampedTruncation, WeightTruncation,
                         XWeightDampedTruncation, XWeightTruncation, compile_filter,
                         is_compilable, NOFILTER)
from ._lib import ArgumentError, DimensionMismatch, PPCounters, PPError, PPStrategyOp
from .spv import (SparsePauliVector, bit_flip_channel, bit_flip_channel_, bit_phase_flip_channel,
                  bit_phase_flip_channel_, dephasing_channel, dephasing_channel_, depolarizing_channel,
                  depolarizing_channel_, depolarizing_p_for_weight_decay, evolve, evolve_, expectation_value,
                  inner_product, isapprox, norm, otimes, osum, largest, largest_diag, diag, offdiag, Matrix,
                  pauli_channel, pauli_channel_, phase_flip_channel_, stochastic_clip_, truncate_, variance)
from .batch import SparsePauliBatch
from .algebra import anticommutator, commutator, multiply
