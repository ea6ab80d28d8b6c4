"""Parareal on one GPU (pararealml/operators/parareal/__init__.py)."""

from pararealml_amd.operators.parareal.parareal_operator import PararealOperator
from pararealml_amd.operators.parareal.propagators import CoarseGridFDMOperator

__all__ = ["PararealOperator", "CoarseGridFDMOperator"]
