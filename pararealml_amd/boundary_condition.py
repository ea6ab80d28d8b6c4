"""Boundary conditions (host-side problem setup).

Semantics follow pararealml/boundary_condition.py:11-287.  A condition is a
vectorised function of boundary coordinates (and optionally time) returning
an ``(n_points, y_dimension)`` array; NaN entries leave the corresponding
element unconstrained.  The derivative condition constrains the derivative
along the +axis direction.
"""

from abc import ABC, abstractmethod
from typing import Callable, Optional, Sequence

import numpy as np

VectorizedBoundaryConditionFunction = Callable[
    [np.ndarray, Optional[float]], np.ndarray
]


class BoundaryCondition(ABC):
    """Base class (pararealml/boundary_condition.py:11-76)."""

    def __init__(
        self, has_y_condition: bool, has_d_y_condition: bool, is_static: bool
    ):
        self._has_y_condition = has_y_condition
        self._has_d_y_condition = has_d_y_condition
        self._is_static = is_static

    @property
    def has_y_condition(self) -> bool:
        return self._has_y_condition

    @property
    def has_d_y_condition(self) -> bool:
        return self._has_d_y_condition

    @property
    def is_static(self) -> bool:
        return self._is_static

    @abstractmethod
    def y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        """Values of y at the boundary points x, shape (n, y_dimension)."""

    @abstractmethod
    def d_y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        """Values of the boundary-normal derivative of y at x."""


class DirichletBoundaryCondition(BoundaryCondition):
    """pararealml/boundary_condition.py:79-112."""

    def __init__(
        self,
        y_condition: VectorizedBoundaryConditionFunction,
        is_static: bool = False,
    ):
        self._y_condition = y_condition
        super().__init__(True, False, is_static)

    def y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        return self._y_condition(x, t)

    def d_y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        raise RuntimeError(
            "Dirichlet conditions do not constrain the normal derivative of y"
        )


class NeumannBoundaryCondition(BoundaryCondition):
    """pararealml/boundary_condition.py:115-143."""

    def __init__(
        self,
        d_y_condition: VectorizedBoundaryConditionFunction,
        is_static: bool = False,
    ):
        self._d_y_condition = d_y_condition
        super().__init__(False, True, is_static)

    def y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        raise RuntimeError("Neumann conditions do not constrain y")

    def d_y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        return self._d_y_condition(x, t)


class CauchyBoundaryCondition(BoundaryCondition):
    """pararealml/boundary_condition.py:146-180."""

    def __init__(
        self,
        y_condition: VectorizedBoundaryConditionFunction,
        d_y_condition: VectorizedBoundaryConditionFunction,
        is_static: bool = False,
    ):
        self._y_condition = y_condition
        self._d_y_condition = d_y_condition
        super().__init__(True, True, is_static)

    def y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        return self._y_condition(x, t)

    def d_y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        return self._d_y_condition(x, t)


class ConstantBoundaryCondition(BoundaryCondition):
    """Static, space-independent constants per y element
    (pararealml/boundary_condition.py:183-243)."""

    def __init__(
        self,
        constant_y_conditions: Optional[Sequence[Optional[float]]],
        constant_d_y_conditions: Optional[Sequence[Optional[float]]],
    ):
        if constant_y_conditions is None and constant_d_y_conditions is None:
            raise ValueError(
                "at least one type of constant conditions must not be None"
            )
        self._constant_y_conditions = constant_y_conditions
        self._constant_d_y_conditions = constant_d_y_conditions
        super().__init__(
            constant_y_conditions is not None,
            constant_d_y_conditions is not None,
            True,
        )

    @staticmethod
    def _tile(constants: Sequence[Optional[float]], n: int) -> np.ndarray:
        # A ``None`` constant yields an object column of ``None`` values,
        # exactly as the reference's column-wise construction does.
        return np.hstack([np.full((n, 1), c) for c in constants])

    def y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        if not self._constant_y_conditions:
            raise RuntimeError("no boundary conditions defined on y")
        return self._tile(self._constant_y_conditions, len(x))

    def d_y_condition(self, x: np.ndarray, t: Optional[float]) -> np.ndarray:
        if not self._constant_d_y_conditions:
            raise RuntimeError(
                "no boundary conditions defined on the normal derivative of y"
            )
        return self._tile(self._constant_d_y_conditions, len(x))


class ConstantValueBoundaryCondition(ConstantBoundaryCondition):
    """pararealml/boundary_condition.py:246-257."""

    def __init__(self, constant_y_conditions: Sequence[Optional[float]]):
        super().__init__(constant_y_conditions, None)


class ConstantFluxBoundaryCondition(ConstantBoundaryCondition):
    """pararealml/boundary_condition.py:260-271."""

    def __init__(self, constant_d_y_conditions: Sequence[Optional[float]]):
        super().__init__(None, constant_d_y_conditions)


def vectorize_bc_function(
    bc_function: Callable[
        [Sequence[float], Optional[float]], Sequence[Optional[float]]
    ]
) -> VectorizedBoundaryConditionFunction:
    """Row-by-row vectorisation of a pointwise boundary condition function;
    ``None`` outputs become NaN (pararealml/boundary_condition.py:274-287)."""

    def vectorized(x: np.ndarray, t: Optional[float]) -> np.ndarray:
        return np.array([bc_function(row, t) for row in x], dtype=float)

    return vectorized
