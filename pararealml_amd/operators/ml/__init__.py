"""Learned coarse propagator: a BF16 convolutional surrogate trained on the
device FDM solver's trajectories; the reference's supervised-learning data
protocol with batched device solves."""

from pararealml_amd.operators.ml.conv_surrogate import ConvSurrogate
from pararealml_amd.operators.ml.learned_coarse import (
    CorrectedCoarseOperator,
    LearnedCoarseOperator,
    corrected_pairs,
    gaussian_mixture_states,
    generate_pairs,
)
from pararealml_amd.operators.ml.supervised_data import SupervisedDataProtocol

__all__ = [
    "ConvSurrogate",
    "CorrectedCoarseOperator",
    "corrected_pairs",
    "LearnedCoarseOperator",
    "gaussian_mixture_states",
    "generate_pairs",
    "SupervisedDataProtocol",
]
