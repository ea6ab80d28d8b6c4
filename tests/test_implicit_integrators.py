"""Host side of the implicit integrators (numerical_integrator.py:135-270):
validation as the reference, and ``integral`` on NumPy arrays equal to the
oracle's restatement (both run SciPy's secant iteration)."""

import numpy as np
import pytest

from oracle import fdm_numpy as ora
from pararealml_amd.operators.fdm import BackwardEulerMethod, CrankNicolsonMethod


def test_validation():
    with pytest.raises(ValueError):
        BackwardEulerMethod(tol=-1.0)
    with pytest.raises(ValueError):
        BackwardEulerMethod(max_iterations=0)
    with pytest.raises(ValueError):
        CrankNicolsonMethod(a=1.5)


def test_integral_matches_oracle_on_a_linear_system():
    rng = np.random.RandomState(3)
    m = -np.eye(4) + 0.1 * rng.standard_normal((4, 4))
    y = rng.standard_normal(4)

    def f(_t, v):
        return m @ v

    def cf(_t):
        return None

    got = BackwardEulerMethod().integral(y, 0.0, 0.1, f, cf)
    np.testing.assert_array_equal(got, ora.integrate_backward_euler(y, 0.0, 0.1, f, cf))
    got = CrankNicolsonMethod().integral(y, 0.0, 0.1, f, cf)
    np.testing.assert_array_equal(got, ora.integrate_crank_nicolson(y, 0.0, 0.1, f, cf))
    # the root of the backward Euler residual: (I - dt m) y_next = y
    want = np.linalg.solve(np.eye(4) - 0.1 * m, y)
    np.testing.assert_allclose(BackwardEulerMethod(tol=1e-13).integral(y, 0.0, 0.1, f, cf),
                               want, rtol=1e-6)
