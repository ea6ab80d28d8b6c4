"""Initial conditions (host-side problem setup).

Semantics follow pararealml/initial_condition.py:14-405.  Vertex-oriented
discretisations always have the problem's static Dirichlet constraints
applied (initial_condition.py:86-89, 226-229); ``discrete_y_0`` returns a
fresh copy.
"""

from abc import ABC, abstractmethod
from copy import deepcopy
from typing import Callable, Optional, Sequence, Tuple

import numpy as np
from scipy.interpolate import interpn
from scipy.stats import beta, multivariate_normal

from pararealml_amd.constrained_problem import ConstrainedProblem
from pararealml_amd.constraint import apply_constraints_along_last_axis
from pararealml_amd.mesh import to_cartesian_coordinates

VectorizedInitialConditionFunction = Callable[[Optional[np.ndarray]], np.ndarray]


class InitialCondition(ABC):
    """pararealml/initial_condition.py:14-46."""

    @abstractmethod
    def y_0(self, x: Optional[np.ndarray]) -> np.ndarray:
        """Initial y at the points x, shape (n, y_dimension)."""

    @abstractmethod
    def discrete_y_0(self, vertex_oriented: Optional[bool] = None) -> np.ndarray:
        """Initial y on the mesh vertices or cell centres."""


class DiscreteInitialCondition(InitialCondition):
    """A fixed array of initial values (pararealml/initial_condition.py:49-121)."""

    def __init__(
        self,
        cp: ConstrainedProblem,
        y_0: np.ndarray,
        vertex_oriented: Optional[bool] = None,
        interpolation_method: str = "linear",
    ):
        if cp.differential_equation.x_dimension and vertex_oriented is None:
            raise ValueError("vertex orientation must be defined for PDEs")
        if y_0.shape != cp.y_shape(vertex_oriented):
            raise ValueError(
                f"discrete initial value shape {y_0.shape} must match "
                "constrained problem solution shape "
                f"{cp.y_shape(vertex_oriented)}"
            )
        self._cp = cp
        self._y_0 = np.array(y_0, copy=True)
        self._vertex_oriented = vertex_oriented
        self._interpolation_method = interpolation_method
        if vertex_oriented:
            apply_constraints_along_last_axis(cp.static_y_vertex_constraints, self._y_0)

    def y_0(self, x: Optional[np.ndarray]) -> np.ndarray:
        if not self._cp.differential_equation.x_dimension:
            return np.copy(self._y_0)
        return interpn(
            self._cp.mesh.axis_coordinates(self._vertex_oriented),
            self._y_0,
            x,
            method=self._interpolation_method,
            bounds_error=False,
            fill_value=None,
        )

    def discrete_y_0(self, vertex_oriented: Optional[bool] = None) -> np.ndarray:
        if vertex_oriented is None:
            vertex_oriented = self._vertex_oriented
        if (
            not self._cp.differential_equation.x_dimension
            or vertex_oriented == self._vertex_oriented
        ):
            return np.copy(self._y_0)
        y_0 = self.y_0(self._cp.mesh.all_index_coordinates(vertex_oriented))
        if vertex_oriented:
            apply_constraints_along_last_axis(
                self._cp.static_y_vertex_constraints, y_0
            )
        return y_0


class ConstantInitialCondition(DiscreteInitialCondition):
    """pararealml/initial_condition.py:124-147."""

    def __init__(self, cp: ConstrainedProblem, constant_y_0s: Sequence[float]):
        y_dim = cp.differential_equation.y_dimension
        if len(constant_y_0s) != y_dim:
            raise ValueError(
                f"length of constant y0 values ({len(constant_y_0s)}) must "
                f"match number of y components ({y_dim})"
            )
        values = np.empty(cp.y_shape(True))
        for i, c in enumerate(constant_y_0s):
            values[..., i] = c
        super().__init__(cp, values, True)


class ContinuousInitialCondition(InitialCondition):
    """An initial condition given by a vectorised function
    (pararealml/initial_condition.py:150-247)."""

    def __init__(
        self,
        cp: ConstrainedProblem,
        y_0_func: VectorizedInitialConditionFunction,
        multipliers: Optional[Sequence[float]] = None,
    ):
        diff_eq = cp.differential_equation
        if multipliers is not None:
            if len(multipliers) != diff_eq.y_dimension:
                raise ValueError(
                    f"length of multipliers ({len(multipliers)}) must match "
                    f"number of y dimensions ({diff_eq.y_dimension})"
                )
            self._multipliers = np.array(multipliers)
        else:
            self._multipliers = np.ones(diff_eq.y_dimension)
        self._cp = cp
        self._y_0_func = y_0_func
        self._discrete_y_0_vertices = self._discretise(True)
        self._discrete_y_0_cells = self._discretise(False)

    def y_0(self, x: Optional[np.ndarray]) -> np.ndarray:
        return np.multiply(self._y_0_func(x), self._multipliers)

    def discrete_y_0(self, vertex_oriented: Optional[bool] = None) -> np.ndarray:
        src = (
            self._discrete_y_0_vertices
            if vertex_oriented
            else self._discrete_y_0_cells
        )
        return np.copy(src)

    def _discretise(self, vertex_oriented: bool) -> np.ndarray:
        diff_eq = self._cp.differential_equation
        if not diff_eq.x_dimension:
            y_0 = np.array(self.y_0(None))
            if y_0.shape != self._cp.y_shape():
                raise ValueError(
                    "expected initial condition function output shape to be "
                    f"{self._cp.y_shape()} but got {y_0.shape}"
                )
            return y_0
        x = self._cp.mesh.all_index_coordinates(vertex_oriented, flatten=True)
        y_0 = self.y_0(x)
        if y_0.shape != (len(x), diff_eq.y_dimension):
            raise ValueError(
                "expected initial condition function output shape to be "
                f"{(len(x), diff_eq.y_dimension)} but got {y_0.shape}"
            )
        y_0 = y_0.reshape(self._cp.y_shape(vertex_oriented))
        if vertex_oriented:
            apply_constraints_along_last_axis(
                self._cp.static_y_vertex_constraints, y_0
            )
        return y_0

    def _convert_coordinates_to_cartesian(self, x: np.ndarray) -> np.ndarray:
        cart = to_cartesian_coordinates(
            [x[:, i] for i in range(x.shape[1])],
            self._cp.mesh.coordinate_system_type,
        )
        return np.stack(cart, axis=-1)


class GaussianInitialCondition(ContinuousInitialCondition):
    """Multivariate normal PDFs per y element
    (pararealml/initial_condition.py:250-308)."""

    def __init__(
        self,
        cp: ConstrainedProblem,
        means_and_covs: Sequence[Tuple[np.ndarray, np.ndarray]],
        multipliers: Optional[Sequence[float]] = None,
    ):
        diff_eq = cp.differential_equation
        if not diff_eq.x_dimension:
            raise ValueError("constrained problem must be a PDE")
        if len(means_and_covs) != diff_eq.y_dimension:
            raise ValueError(
                f"number of means and covariances ({len(means_and_covs)}) "
                f"must match number of y dimensions ({diff_eq.y_dimension})"
            )
        for mean, cov in means_and_covs:
            if mean.shape != (diff_eq.x_dimension,):
                raise ValueError(
                    f"expected mean shape to be {(diff_eq.x_dimension,)} but "
                    f"got {mean.shape}"
                )
            if cov.shape != (diff_eq.x_dimension, diff_eq.x_dimension):
                raise ValueError(
                    "expected covariance shape to be "
                    f"{(diff_eq.x_dimension, diff_eq.x_dimension)} but got "
                    f"{cov.shape}"
                )
        self._means_and_covs = deepcopy(means_and_covs)
        super().__init__(cp, self._pdf_values, multipliers)

    def _pdf_values(self, x: Optional[np.ndarray]) -> np.ndarray:
        cart = self._convert_coordinates_to_cartesian(x)
        out = np.empty((len(x), self._cp.differential_equation.y_dimension))
        for i, (mean, cov) in enumerate(self._means_and_covs):
            out[:, i] = multivariate_normal.pdf(cart, mean=mean, cov=cov)
        return out


class MarginalBetaProductInitialCondition(ContinuousInitialCondition):
    """Products of per-axis Beta PDFs (pararealml/initial_condition.py:311-381)."""

    def __init__(
        self,
        cp: ConstrainedProblem,
        all_alphas_and_betas: Sequence[Sequence[Tuple[float, float]]],
        multipliers: Optional[Sequence[float]] = None,
    ):
        diff_eq = cp.differential_equation
        if len(all_alphas_and_betas) != diff_eq.y_dimension:
            raise ValueError(
                "number of alphas and betas sequences "
                f"({len(all_alphas_and_betas)}) must match the number of y "
                f"dimensions ({diff_eq.y_dimension})"
            )
        if any(len(ab) != diff_eq.x_dimension for ab in all_alphas_and_betas):
            raise ValueError(
                "all sequences of alphas and betas must have same length as "
                f"number of spatial dimensions ({diff_eq.x_dimension})"
            )
        self._all_alphas_and_betas = deepcopy(all_alphas_and_betas)
        super().__init__(cp, self._pdf_values, multipliers)

    def _pdf_values(self, x: Optional[np.ndarray]) -> np.ndarray:
        cart = self._convert_coordinates_to_cartesian(x)
        columns = []
        for alphas_and_betas in self._all_alphas_and_betas:
            prod = np.ones((len(x), 1))
            for axis, (a, b) in enumerate(alphas_and_betas):
                prod = prod * beta.pdf(cart[:, axis : axis + 1], a, b)
            columns.append(prod)
        return np.concatenate(columns, axis=-1)


def vectorize_ic_function(
    ic_function: Callable[[Optional[Sequence[float]]], Sequence[float]]
) -> VectorizedInitialConditionFunction:
    """Row-by-row vectorisation (pararealml/initial_condition.py:384-405)."""

    def vectorized(x: Optional[np.ndarray]) -> np.ndarray:
        if x is None:
            return np.array(ic_function(None))
        return np.array([ic_function(row) for row in x])

    return vectorized
