"""pauliprop — MI355X-native SparsePauliVector evolution (host side).

The Python mirror of PauliOperators.jl's SparsePauliVector API.  Julia's
mutating `f!` names become `f_` here (PyTorch's in-place convention).
"""
from .basis import (Ket, Pauli, PauliBasis, PauliSum, commute, majorana_weight, weight,
                    word_count, x_weight)
from .truncation import (AdaptiveTruncation, CoeffTruncation, CompositeTruncation,
                         MajoranaWeightTruncation, NoCorrection, NoTruncation, PPFilter,
                         StochasticCoeffTruncation, StochasticSamplingTruncation,
                         TruncationStrategy, WeightDampedTruncation, WeightTruncation,
                         XWeightDampedTruncation, XWeightTruncation, compile_filter,
                         is_compilable, NOFILTER)
