"""The oracle's time-dependent boundary conditions (fdm_operator.py:55-63,
199-231): with conditions that do not change in time the dynamic path must
reproduce the static one exactly, and the Dirichlet values of every step are
those of the step's end time."""

import numpy as np
import pytest

import pararealml_amd as pml
from oracle import fdm_numpy as ora


def _problem(static, g=lambda t: 0.5, s=lambda t: -0.2):
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.2)], [0.1, 0.1])
    bcs = [
        (pml.DirichletBoundaryCondition(lambda x, t: np.full((len(x), 1), g(t)),
                                        is_static=static),
         pml.NeumannBoundaryCondition(lambda x, t: np.zeros((len(x), 1)), is_static=static)),
        (pml.NeumannBoundaryCondition(lambda x, t: np.zeros((len(x), 1)), is_static=static),
         pml.NeumannBoundaryCondition(lambda x, t: np.full((len(x), 1), s(t)),
                                      is_static=static)),
    ]
    cp = pml.ConstrainedProblem(pml.DiffusionEquation(2, 0.4), mesh, bcs)
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.5, 0.6]), 0.05 * np.eye(2))])
    return cp, ic


@pytest.mark.parametrize("integ", ["fe", "midpoint", "rk4"])
def test_constant_dynamic_conditions_equal_static(integ):
    cp_s, ic = _problem(True)
    cp_d, _ = _problem(False)
    y0 = ic.discrete_y_0(True)
    _, ys = ora.fdm_solve(cp_s, y0, (0.0, 0.02), 0.002, integ)
    _, yd = ora.fdm_solve(cp_d, y0, (0.0, 0.02), 0.002, integ)
    np.testing.assert_array_equal(yd, ys)


def test_dirichlet_values_follow_time():
    cp, ic = _problem(False, g=lambda t: np.sin(5.0 * t))
    t, y = ora.fdm_solve(cp, ic.discrete_y_0(True), (0.0, 0.02), 0.002, "rk4")
    for k, t_k in enumerate(t):
        np.testing.assert_allclose(y[k, 0, :, 0], np.sin(5.0 * t_k), rtol=0, atol=1e-15)
