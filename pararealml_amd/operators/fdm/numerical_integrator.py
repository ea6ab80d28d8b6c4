"""Explicit Runge-Kutta integrators.

Inside ``FDMOperator`` an integrator selects the fused device kernel
(``pml_fdm_solve``) that evaluates all of its stages in LDS; the classes
also keep the reference's ``integral`` method (pararealml/operators/fdm/
numerical_integrator.py:16-132) for callers that drive an integrator with
their own right-hand side.  ``integral`` evaluates the stage formulas with
the operators of whatever array type it is given -- NumPy arrays on the host
or torch tensors in HBM -- and applies the constraints in place.

The implicit methods (BackwardEulerMethod, CrankNicolsonMethod,
numerical_integrator.py:135-270) solve for the next state with the
element-wise secant method of ``scipy.optimize.newton``: ``integral`` calls
SciPy on host arrays as the reference does; inside ``FDMOperator`` the same
iteration runs on the device (``pml_secant_step``).
"""

from abc import ABC, abstractmethod
from typing import Callable, Optional, Sequence, Union

import numpy as np

from pararealml_amd.constraint import Constraint, apply_constraints_along_last_axis

ConstraintFunction = Callable[
    [Optional[float]], Optional[Union[Sequence[Constraint], np.ndarray]]
]


def _constrain(constraints, array):
    if constraints is None:
        return array
    if isinstance(array, np.ndarray):
        return apply_constraints_along_last_axis(constraints, array)
    # torch tensors: same masked assignment, performed on the tensor's device
    import torch

    if len(constraints) != array.shape[-1]:
        raise ValueError(
            f"number of constraints ({len(constraints)}) must match the "
            f"size of the input array's last axis ({array.shape[-1]})"
        )
    for i, c in enumerate(constraints):
        if c is None:
            continue
        mask = torch.as_tensor(c.mask, device=array.device)
        vals = torch.as_tensor(c.values, dtype=array.dtype, device=array.device)
        array[..., i : i + 1][..., mask] = vals
    return array


class NumericalIntegrator(ABC):
    """pararealml/operators/fdm/numerical_integrator.py:10-39."""

    #: number of right-hand-side evaluations per step
    stages: int = 1

    @abstractmethod
    def integral(
        self,
        y,
        t: float,
        d_t: float,
        d_y_over_d_t: Callable,
        y_constraint_function: ConstraintFunction,
    ):
        """Estimates y(t + d_t)."""


class ForwardEulerMethod(NumericalIntegrator):
    """``C(y + dt f(t, y))`` (numerical_integrator.py:42-62)."""

    stages = 1

    def integral(self, y, t, d_t, d_y_over_d_t, y_constraint_function):
        return _constrain(y_constraint_function(t + d_t), y + d_t * d_y_over_d_t(t, y))


class ExplicitMidpointMethod(NumericalIntegrator):
    """Two-stage explicit midpoint method (numerical_integrator.py:65-90)."""

    stages = 2

    def integral(self, y, t, d_t, d_y_over_d_t, y_constraint_function):
        half = d_t / 2.0
        y_hat = _constrain(y_constraint_function(t + half), y + half * d_y_over_d_t(t, y))
        return _constrain(
            y_constraint_function(t + d_t), y + d_t * d_y_over_d_t(t + half, y_hat)
        )


class RK4(NumericalIntegrator):
    """Classical fourth-order Runge-Kutta (numerical_integrator.py:93-132)."""

    stages = 4

    def integral(self, y, t, d_t, d_y_over_d_t, y_constraint_function):
        half = d_t / 2.0
        c_half = y_constraint_function(t + half)
        c_full = y_constraint_function(t + d_t)
        k1 = d_t * d_y_over_d_t(t, y)
        k2 = d_t * d_y_over_d_t(t + half, _constrain(c_half, y + k1 / 2.0))
        k3 = d_t * d_y_over_d_t(t + half, _constrain(c_half, y + k2 / 2.0))
        k4 = d_t * d_y_over_d_t(t + d_t, _constrain(c_full, y + k3))
        return _constrain(c_full, y + (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0)


class ImplicitMethod(NumericalIntegrator, ABC):
    """numerical_integrator.py:135-176: the next state is the root of a
    residual, found by scipy.optimize.newton without a derivative (the
    element-wise secant method) with tolerance ``tol`` and at most
    ``max_iterations`` iterations."""

    def __init__(self, tol: float = 1.48e-8, max_iterations: int = 50):
        if tol < 0.0:
            raise ValueError("tolerance must be non-negative")
        if max_iterations <= 0:
            raise ValueError("number of maximum iterations must be greater than 0")
        self._tol = tol
        self._max_iterations = max_iterations

    @property
    def tol(self) -> float:
        return self._tol

    @property
    def max_iterations(self) -> int:
        return self._max_iterations

    def _solve(self, residual, y_next_init):
        from scipy.optimize import newton

        return newton(residual, y_next_init, tol=self._tol, maxiter=self._max_iterations)


class BackwardEulerMethod(ImplicitMethod):
    """numerical_integrator.py:179-217."""

    #: weight of the implicit right-hand side (the explicit one gets 1 - a)
    a = 1.0

    def __init__(self, tol: float = 1.48e-8, max_iterations: int = 50):
        super().__init__(tol, max_iterations)

    def integral(self, y, t, d_t, d_y_over_d_t, y_constraint_function):
        t_next = t + d_t
        c_next = y_constraint_function(t_next)
        init = _constrain(c_next, y + d_t * d_y_over_d_t(t, y))

        def residual(y_next):
            return y_next - _constrain(c_next, y + d_t * d_y_over_d_t(t_next, y_next))

        return self._solve(residual, init)


class CrankNicolsonMethod(ImplicitMethod):
    """numerical_integrator.py:220-270."""

    def __init__(self, a: float = 0.5, tol: float = 1.48e-8, max_iterations: int = 50):
        if not (0.0 <= a <= 1.0):
            raise ValueError("the value of 'a' must be between 0 and 1")
        self.a = a
        self._b = 1.0 - a
        super().__init__(tol, max_iterations)

    def integral(self, y, t, d_t, d_y_over_d_t, y_constraint_function):
        t_next = t + d_t
        forward = d_t * d_y_over_d_t(t, y)
        c_next = y_constraint_function(t_next)
        init = _constrain(c_next, y + forward)

        def residual(y_next):
            return y_next - _constrain(
                c_next, y + self.a * d_t * d_y_over_d_t(t_next, y_next) + self._b * forward
            )

        return self._solve(residual, init)
