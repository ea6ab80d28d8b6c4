"""Solutions of initial value problems.

Semantics follow pararealml/solution.py:22-258 and 329-336 (constructor
validation and copying, interpolation, orientation conversion, and the
time-matched differences).  Plotting (solution.py:260-326) is presentation
and is not part of this package.
"""

from __future__ import annotations

from typing import List, NamedTuple, Optional, Sequence

import numpy as np
from scipy.interpolate import interpn

from pararealml_amd.constraint import apply_constraints_along_last_axis
from pararealml_amd.initial_value_problem import InitialValueProblem


class Solution:
    """A discrete solution: time coordinates and ``(n_t, *y_shape)`` values."""

    def __init__(
        self,
        ivp: InitialValueProblem,
        t_coordinates: np.ndarray,
        discrete_y: np.ndarray,
        vertex_oriented: Optional[bool] = None,
        d_t: Optional[float] = None,
    ):
        if t_coordinates.ndim != 1:
            raise ValueError(
                f"number of t coordinate dimensions ({t_coordinates.ndim}) "
                "must be 1"
            )
        if len(t_coordinates) == 0:
            raise ValueError("length of t coordinates must be greater than 0")
        cp = ivp.constrained_problem
        if cp.differential_equation.x_dimension and vertex_oriented is None:
            raise ValueError(
                "vertex orientation must be defined for solutions to PDEs"
            )
        y_shape = cp.y_shape(vertex_oriented)
        if discrete_y.shape != (len(t_coordinates),) + y_shape:
            raise ValueError(
                "expected solution shape to be "
                f"{(len(t_coordinates),) + y_shape} but got {discrete_y.shape}"
            )
        self._ivp = ivp
        self._t_coordinates = np.array(t_coordinates, copy=True)
        self._discrete_y = np.array(discrete_y, copy=True)
        self._vertex_oriented = vertex_oriented
        self._t_coordinates.setflags(write=False)
        if d_t is None:
            d_t = (
                0.0
                if len(t_coordinates) == 1
                else t_coordinates[1] - t_coordinates[0]
            )
        self._d_t = d_t

    @property
    def initial_value_problem(self) -> InitialValueProblem:
        return self._ivp

    @property
    def vertex_oriented(self) -> Optional[bool]:
        return self._vertex_oriented

    @property
    def d_t(self) -> float:
        return self._d_t

    @property
    def t_coordinates(self) -> np.ndarray:
        return self._t_coordinates

    def y(
        self, x: Optional[np.ndarray] = None, interpolation_method: str = "linear"
    ) -> np.ndarray:
        """Values at arbitrary spatial points for every time step
        (pararealml/solution.py:114-145)."""
        cp = self._ivp.constrained_problem
        diff_eq = cp.differential_equation
        if not diff_eq.x_dimension:
            return np.copy(self._discrete_y)
        values = interpn(
            cp.mesh.axis_coordinates(self._vertex_oriented),
            np.moveaxis(self._discrete_y, 0, -2),
            x,
            method=interpolation_method,
            bounds_error=False,
            fill_value=None,
        )
        values = np.moveaxis(values, -2, 0)
        values = values.reshape(
            (len(self._t_coordinates),) + x.shape[:-1] + (diff_eq.y_dimension,)
        )
        return np.ascontiguousarray(values)

    def discrete_y(
        self,
        vertex_oriented: Optional[bool] = None,
        interpolation_method: str = "linear",
    ) -> np.ndarray:
        """The solution on vertices or cell centres
        (pararealml/solution.py:147-180)."""
        if vertex_oriented is None:
            vertex_oriented = self._vertex_oriented
        cp = self._ivp.constrained_problem
        if (
            not cp.differential_equation.x_dimension
            or self._vertex_oriented == vertex_oriented
        ):
            return np.copy(self._discrete_y)
        x = cp.mesh.all_index_coordinates(vertex_oriented)
        values = self.y(x, interpolation_method)
        if vertex_oriented:
            apply_constraints_along_last_axis(cp.static_y_vertex_constraints, values)
        return values

    def diff(self, solutions: Sequence["Solution"], atol: float = 1e-8) -> "Diffs":
        """Differences to other solutions at matching time points
        (pararealml/solution.py:182-258)."""
        if len(solutions) == 0:
            raise ValueError("length of solutions must be greater than 0")
        all_t = [self._t_coordinates] + [s.t_coordinates for s in solutions]
        all_dt = [self._d_t] + [s.d_t for s in solutions]
        others = [s.discrete_y(self._vertex_oriented) for s in solutions]
        shortest = int(np.argmin([len(t) for t in all_t]))

        matching: List[float] = []
        diffs: List[List[np.ndarray]] = [[] for _ in solutions]
        for i, t in enumerate(all_t[shortest]):
            indices = []
            for j, t_j in enumerate(all_t):
                if j == shortest:
                    indices.append(i)
                    continue
                k = int(round((t - t_j[0]) / all_dt[j]))
                if 0 <= k < len(t_j) and np.isclose(t, t_j[k], atol=atol, rtol=0.0):
                    indices.append(k)
                else:
                    indices = None
                    break
            if indices is None:
                continue
            matching.append(t)
            for j, other in enumerate(others):
                diffs[j].append(other[indices[j + 1]] - self._discrete_y[indices[0]])
        return Diffs(np.array(matching), [np.array(d) for d in diffs])


class Diffs(NamedTuple):
    """pararealml/solution.py:329-336."""

    matching_time_points: np.ndarray
    differences: Sequence[np.ndarray]
