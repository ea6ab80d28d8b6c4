"""Differential equations as symbolic right-hand sides (host-side setup).

The symbol vocabulary (``y``, ``x``, ``y-gradient``, ``y-hessian``,
``y-divergence``, ``y-curl``, ``y-laplacian``, ``y-vector-laplacian``), the
left-hand side kinds and the equation catalogue follow
pararealml/differential_equation.py:10-850.  The symbolic systems are what
the CPU oracle evaluates generically; the device path instead recognises the
equation class and runs a hand-fused kernel for it, with the coefficients
exposed through ``device_parameters``.
"""

from abc import ABC, abstractmethod
from copy import copy, deepcopy
from enum import Enum
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np
from sympy import Expr, Symbol, symarray


class Symbols:
    """Symbol arrays for an equation with the given x and y dimensions
    (pararealml/differential_equation.py:10-139)."""

    def __init__(self, x_dimension: int, y_dimension: int):
        self._t = Symbol("t")
        self._y = symarray("y", (y_dimension,))
        self._x = None
        self._y_gradient = None
        self._y_hessian = None
        self._y_divergence = None
        self._y_curl = None
        self._y_laplacian = None
        self._y_vector_laplacian = None
        if x_dimension:
            self._x = symarray("x", (x_dimension,))
            self._y_gradient = symarray("y-gradient", (y_dimension, x_dimension))
            self._y_hessian = symarray(
                "y-hessian", (y_dimension, x_dimension, x_dimension)
            )
            self._y_divergence = symarray(
                "y-divergence", (y_dimension,) * x_dimension
            )
            if 2 <= x_dimension <= 3:
                curl_shape = (
                    (y_dimension,) * x_dimension + (x_dimension,)
                    if x_dimension == 3
                    else ()
                )
                self._y_curl = symarray("y-curl", curl_shape)
            self._y_laplacian = symarray("y-laplacian", (y_dimension,))
            self._y_vector_laplacian = symarray(
                "y-vector-laplacian",
                (y_dimension,) * x_dimension + (x_dimension,),
            )

    t = property(lambda self: self._t)
    y = property(lambda self: copy(self._y))
    x = property(lambda self: copy(self._x))
    y_gradient = property(lambda self: copy(self._y_gradient))
    y_hessian = property(lambda self: copy(self._y_hessian))
    y_divergence = property(lambda self: copy(self._y_divergence))
    y_curl = property(lambda self: copy(self._y_curl))
    y_laplacian = property(lambda self: copy(self._y_laplacian))
    y_vector_laplacian = property(lambda self: copy(self._y_vector_laplacian))


class LHS(Enum):
    """Left-hand side kinds (pararealml/differential_equation.py:142-150)."""

    D_Y_OVER_D_T = 0
    Y = 1
    Y_LAPLACIAN = 2


class SymbolicEquationSystem:
    """pararealml/differential_equation.py:153-211."""

    def __init__(
        self,
        rhs: Union[Sequence[Expr], np.ndarray],
        lhs_types: Optional[Sequence[LHS]] = None,
    ):
        if len(rhs) < 1:
            raise ValueError("number of equations must be greater than 0")
        if lhs_types is None:
            lhs_types = [LHS.D_Y_OVER_D_T] * len(rhs)
        if len(rhs) != len(lhs_types):
            raise ValueError(
                f"length of right-hand side ({len(rhs)}) must match length of "
                f"left-hand side ({len(lhs_types)})"
            )
        self._rhs = copy(rhs)
        self._lhs_types = copy(lhs_types)
        self._by_type: Dict[LHS, List[int]] = {kind: [] for kind in LHS}
        for i, kind in enumerate(lhs_types):
            self._by_type[kind].append(i)

    @property
    def rhs(self) -> Union[Sequence[Expr], np.ndarray]:
        return copy(self._rhs)

    @property
    def lhs_types(self) -> Sequence[LHS]:
        return copy(self._lhs_types)

    def equation_indices_by_type(self, lhs_type: LHS) -> Sequence[int]:
        return copy(self._by_type[lhs_type])


class DifferentialEquation(ABC):
    """pararealml/differential_equation.py:214-352."""

    def __init__(
        self,
        x_dimension: int,
        y_dimension: int,
        all_vector_field_indices: Optional[Sequence[Sequence[int]]] = None,
    ):
        if x_dimension < 0:
            raise ValueError(
                f"number of x dimensions ({x_dimension}) must be non-negative"
            )
        if y_dimension < 1:
            raise ValueError(
                f"number of y dimensions ({y_dimension}) must be at least 1"
            )
        if all_vector_field_indices:
            for indices in all_vector_field_indices:
                if len(indices) != x_dimension:
                    raise ValueError(
                        f"length of vector field indices {indices} must match "
                        f"x dimensions ({x_dimension})"
                    )
                for index in indices:
                    if not (0 <= index < y_dimension):
                        raise ValueError(
                            "all indices must be non-negative and less than "
                            f"the number of y dimensions ({y_dimension})"
                        )
        self._x_dimension = x_dimension
        self._y_dimension = y_dimension
        self._all_vector_field_indices = deepcopy(all_vector_field_indices)
        self._symbols = Symbols(x_dimension, y_dimension)
        self._validate_equations()

    @property
    def x_dimension(self) -> int:
        return self._x_dimension

    @property
    def y_dimension(self) -> int:
        return self._y_dimension

    @property
    def symbols(self) -> Symbols:
        return self._symbols

    @property
    def all_vector_field_indices(self) -> Optional[Sequence[Sequence[int]]]:
        return deepcopy(self._all_vector_field_indices)

    @property
    @abstractmethod
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        """The symbolic right-hand sides and their left-hand side kinds."""

    def device_parameters(self) -> Optional[Tuple[str, Tuple[float, ...]]]:
        """(kernel family, coefficients) when a fused device kernel exists for
        this equation, else None."""
        return None

    def device_y_laplacian(self) -> List[Tuple[int, int, float]]:
        """The LHS.Y_LAPLACIAN equations as (field, source field, c): the
        field's Laplacian is c times the source field (the device solves
        them after each step); empty when there are none."""
        return []

    def _validate_equations(self):
        system = self.symbolic_equation_system
        if len(system.rhs) != self._y_dimension:
            raise ValueError(
                f"number of equations ({len(system.rhs)}) must match "
                f"number of y dimensions ({self._y_dimension})"
            )
        allowed = {self._symbols.t}
        allowed.update(self._symbols.y)
        if self._x_dimension:
            allowed.update(self._symbols.x)
            allowed.update(self._symbols.y_gradient.flatten())
            allowed.update(self._symbols.y_hessian.flatten())
            allowed.update(self._symbols.y_divergence.flatten())
            if 2 <= self._x_dimension <= 3:
                allowed.update(self._symbols.y_curl.flatten())
            allowed.update(self._symbols.y_laplacian)
            allowed.update(self._symbols.y_vector_laplacian.flatten())
        for i, expr in enumerate(system.rhs):
            if not expr.free_symbols.issubset(allowed):
                raise ValueError(
                    "invalid symbol in right-hand side symbols "
                    f"({expr.free_symbols}) of equation {i}"
                )
        n_time = len(system.equation_indices_by_type(LHS.D_Y_OVER_D_T))
        if self._x_dimension:
            if n_time == 0:
                raise ValueError(
                    "at least one equation's left-hand side must be of type "
                    "D_Y_OVER_D_T"
                )
        elif n_time != self._y_dimension:
            raise ValueError(
                "ordinary differential equation systems can only contain "
                "equations with D_Y_OVER_D_T type left-hand sides"
            )


# --------------------------------------------------------------------- ODEs
class PopulationGrowthEquation(DifferentialEquation):
    """pararealml/differential_equation.py:355-370."""

    def __init__(self, r: float = 0.01):
        self._r = r
        super().__init__(0, 1)

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        return SymbolicEquationSystem([self._r * self._symbols.y[0]])


class LotkaVolterraEquation(DifferentialEquation):
    """pararealml/differential_equation.py:373-411."""

    def __init__(
        self,
        alpha: float = 2.0,
        beta: float = 0.04,
        gamma: float = 1.06,
        delta: float = 0.02,
    ):
        if min(alpha, beta, gamma, delta) < 0.0:
            raise ValueError("all coefficients must be non-negative")
        self._alpha, self._beta = alpha, beta
        self._gamma, self._delta = gamma, delta
        super().__init__(0, 2)

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        prey, pred = self._symbols.y
        return SymbolicEquationSystem(
            [
                self._alpha * prey - self._beta * prey * pred,
                self._delta * prey * pred - self._gamma * pred,
            ]
        )


class LorenzEquation(DifferentialEquation):
    """pararealml/differential_equation.py:414-448."""

    def __init__(
        self, sigma: float = 10.0, rho: float = 28.0, beta: float = 8.0 / 3.0
    ):
        if min(sigma, rho, beta) < 0.0:
            raise ValueError("all coefficients must be non-negative")
        self._sigma, self._rho, self._beta = sigma, rho, beta
        super().__init__(0, 3)

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        c, h, v = self._symbols.y
        return SymbolicEquationSystem(
            [
                self._sigma * (h - c),
                c * (self._rho - v) - h,
                c * h - self._beta * v,
            ]
        )


class SIREquation(DifferentialEquation):
    """pararealml/differential_equation.py:451-483."""

    def __init__(self, beta: float = 0.2, gamma: float = 0.1):
        if beta < 0.0 or gamma < 0.0:
            raise ValueError("beta and gamma must be non-negative")
        self._beta, self._gamma = beta, gamma
        super().__init__(0, 3)

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        s, i, r = self._symbols.y
        n = s + i + r
        return SymbolicEquationSystem(
            [
                -self._beta * s * i / n,
                self._beta * s * i / n - self._gamma * i,
                self._gamma * i,
            ]
        )


class VanDerPolEquation(DifferentialEquation):
    """pararealml/differential_equation.py:486-507."""

    def __init__(self, mu: float = 1.0):
        if mu < 0.0:
            raise ValueError("mu must be non-negative")
        self._mu = mu
        super().__init__(0, 2)

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        u, v = self.symbols.y
        return SymbolicEquationSystem([v, self._mu * (1.0 - u**2) * v - u])


class NBodyGravitationalEquation(DifferentialEquation):
    """Positions then velocities of n bodies in 2 or 3 dimensions
    (pararealml/differential_equation.py:510-605)."""

    def __init__(
        self, n_dims: int, masses: Sequence[float], g: float = 6.6743e-11
    ):
        if n_dims < 2 or n_dims > 3:
            raise ValueError(
                f"number of dimensions ({n_dims}) must be either 2 or 3"
            )
        if len(masses) < 2:
            raise ValueError(
                f"number of masses ({len(masses)}) must be at least 2"
            )
        if np.any(np.array(masses) <= 0.0):
            raise ValueError(f"all masses ({masses}) must be greater than 0")
        self._dims = n_dims
        self._masses = tuple(masses)
        self._n_objects = len(masses)
        self._g = g
        super().__init__(0, 2 * len(masses) * n_dims)

    @property
    def spatial_dimension(self) -> int:
        return self._dims

    @property
    def masses(self) -> Tuple[float, ...]:
        return copy(self._masses)

    @property
    def n_objects(self) -> int:
        return self._n_objects

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        y = np.array(self._symbols.y, dtype=object)
        k, dims = self._n_objects, self._dims
        n_pos = k * dims
        rhs = np.empty(self._y_dimension, dtype=object)
        rhs[:n_pos] = y[n_pos:]
        forces = np.zeros((k, k, dims), dtype=object)
        for i in range(k):
            p_i = y[i * dims : (i + 1) * dims]
            for j in range(i + 1, k):
                p_j = y[j * dims : (j + 1) * dims]
                disp = p_j - p_i
                dist = np.power(np.power(disp, 2).sum(axis=-1), 0.5)
                force = (self._g * self._masses[i] * self._masses[j]) * (
                    disp / np.power(dist, 3)
                )
                forces[i, j, :] = force
                forces[j, i, :] = -force
            rhs[n_pos + i * dims : n_pos + (i + 1) * dims] = (
                forces[i, :, :].sum(axis=0) / self._masses[i]
            )
        return SymbolicEquationSystem(rhs)


# --------------------------------------------------------------------- PDEs
class DiffusionEquation(DifferentialEquation):
    """``y_t = d * lap(y)`` (pararealml/differential_equation.py:608-630)."""

    def __init__(self, x_dimension: int, d: float = 1.0):
        if x_dimension <= 0:
            raise ValueError(
                f"number of x dimensions ({x_dimension}) must be at least 1"
            )
        self._d = d
        super().__init__(x_dimension, 1)

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        return SymbolicEquationSystem([self._d * self._symbols.y_laplacian[0]])

    def device_parameters(self):
        return "diffusion", (float(self._d),)


class ConvectionDiffusionEquation(DifferentialEquation):
    """``y_t = d * lap(y) - v . grad(y)``
    (pararealml/differential_equation.py:633-670)."""

    def __init__(
        self, x_dimension: int, velocity: Sequence[float], d: float = 1.0
    ):
        if x_dimension <= 0:
            raise ValueError(
                f"number of x dimensions ({x_dimension}) must be at least 1"
            )
        if len(velocity) != x_dimension:
            raise ValueError(
                f"length of the velocity vector ({len(velocity)}) must match "
                f"number of x dimensions ({x_dimension})"
            )
        self._velocity = copy(velocity)
        self._d = d
        super().__init__(x_dimension, 1)

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        return SymbolicEquationSystem(
            [
                self._d * self._symbols.y_laplacian[0]
                - np.dot(self._velocity, self._symbols.y_gradient[0, :])
            ]
        )

    def device_parameters(self):
        vel = [float(v) for v in self._velocity] + [0.0] * (2 - len(self._velocity))
        return "convection_diffusion", (float(self._d), vel[0], vel[1])


class WaveEquation(DifferentialEquation):
    """``(y0, y1)_t = (y1, c^2 lap(y0))``
    (pararealml/differential_equation.py:673-697)."""

    def __init__(self, x_dimension: int, c: float = 1.0):
        if x_dimension <= 0:
            raise ValueError(
                f"number of x dimensions ({x_dimension}) must be at least 1"
            )
        self._c = c
        super().__init__(x_dimension, 2)

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        return SymbolicEquationSystem(
            [self._symbols.y[1], (self._c**2) * self._symbols.y_laplacian[0]]
        )

    def device_parameters(self):
        return "wave", (float(self._c**2),)


class CahnHilliardEquation(DifferentialEquation):
    """``c_t = d lap(mu)`` with the algebraic ``mu = c^3 - c - gamma lap(c)``
    (pararealml/differential_equation.py:700-731)."""

    def __init__(self, x_dimension: int, d: float = 0.1, gamma: float = 0.01):
        if x_dimension <= 0:
            raise ValueError(
                f"number of x dimensions ({x_dimension}) must be at least 1"
            )
        self._d = d
        self._gamma = gamma
        super().__init__(x_dimension, 2)

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        sym = self._symbols
        return SymbolicEquationSystem(
            [
                self._d * sym.y_laplacian[1],
                sym.y[0] ** 3 - sym.y[0] - self._gamma * sym.y_laplacian[0],
            ],
            [LHS.D_Y_OVER_D_T, LHS.Y],
        )

    def device_parameters(self):
        return "cahn_hilliard", (float(self._d), float(self._gamma))


class BurgersEquation(DifferentialEquation):
    """``u_t = lap(u)/re - (u . grad) u``
    (pararealml/differential_equation.py:734-765)."""

    def __init__(self, x_dimension: int, re: float = 4000.0):
        if x_dimension <= 0:
            raise ValueError(
                f"number of x dimensions ({x_dimension}) must be at least 1"
            )
        self._re = re
        super().__init__(x_dimension, x_dimension, [tuple(range(x_dimension))])

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        sym = self._symbols
        return SymbolicEquationSystem(
            [
                (1.0 / self._re) * sym.y_laplacian[i]
                - np.dot(sym.y, sym.y_gradient[i, :])
                for i in range(self._x_dimension)
            ]
        )

    def device_parameters(self):
        return "burgers", (1.0 / float(self._re),)


class ShallowWaterEquation(DifferentialEquation):
    """Non-conservative shallow water equations in (eta, u, v)
    (pararealml/differential_equation.py:768-819)."""

    def __init__(
        self,
        h: float,
        b: float = 0.01,
        v: float = 0.1,
        f: float = 0.0,
        g: float = 9.80665,
    ):
        self._h, self._b, self._v, self._f, self._g = h, b, v, f, g
        super().__init__(2, 3, [(1, 2)])

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        s = self._symbols
        y, grad = s.y, s.y_gradient
        return SymbolicEquationSystem(
            [
                -self._h * s.y_divergence[1, 2]
                - y[0] * grad[1, 0]
                - y[1] * grad[0, 0]
                - y[0] * grad[2, 1]
                - y[2] * grad[0, 1],
                self._v * s.y_laplacian[1]
                - y[1] * grad[1, 0]
                - y[2] * grad[1, 1]
                - self._g * grad[0, 0]
                - self._b * y[1]
                + self._f * y[2],
                self._v * s.y_laplacian[2]
                - y[1] * grad[2, 0]
                - y[2] * grad[2, 1]
                - self._g * grad[0, 1]
                - self._b * y[2]
                - self._f * y[1],
            ]
        )

    def device_parameters(self):
        return "shallow_water", tuple(
            float(c) for c in (self._h, self._b, self._v, self._f, self._g)
        )


class NavierStokesEquation(DifferentialEquation):
    """Vorticity / stream-function formulation in 2D
    (pararealml/differential_equation.py:822-850)."""

    def __init__(self, re: float = 4000.0):
        self._re = re
        super().__init__(2, 4, [(2, 3)])

    @property
    def symbolic_equation_system(self) -> SymbolicEquationSystem:
        s = self._symbols
        return SymbolicEquationSystem(
            [
                (1.0 / self._re) * s.y_laplacian[0]
                - np.dot(s.y[2:], s.y_gradient[0, :]),
                -s.y[0],
                s.y_gradient[1, 1],
                -s.y_gradient[1, 0],
            ],
            [LHS.D_Y_OVER_D_T, LHS.Y_LAPLACIAN, LHS.Y, LHS.Y],
        )

    def device_parameters(self):
        return "navier_stokes", (1.0 / float(self._re),)

    def device_y_laplacian(self):
        # lap(psi) = -w: (LHS.Y_LAPLACIAN field, source field, coefficient)
        return [(1, 0, -1.0)]
