"""Finite-difference operator, integrators and differentiators
(pararealml/operators/fdm/__init__.py)."""

from pararealml_amd.operators.fdm.fdm_operator import FDMOperator
from pararealml_amd.operators.fdm.numerical_differentiator import (
    NumericalDifferentiator,
    ThreePointCentralDifferenceMethod,
)
from pararealml_amd.operators.fdm.numerical_integrator import (
    RK4,
    BackwardEulerMethod,
    CrankNicolsonMethod,
    ExplicitMidpointMethod,
    ForwardEulerMethod,
    ImplicitMethod,
    NumericalIntegrator,
)

__all__ = [
    "FDMOperator",
    "NumericalDifferentiator",
    "ThreePointCentralDifferenceMethod",
    "NumericalIntegrator",
    "ForwardEulerMethod",
    "ExplicitMidpointMethod",
    "RK4",
    "ImplicitMethod",
    "BackwardEulerMethod",
    "CrankNicolsonMethod",
]
