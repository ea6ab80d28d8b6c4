"""Pins of the oracle's restatement of the supervised-learning data protocol
(oracle/supervised_numpy.py; pararealml/operators/ml/supervised/
supervised_ml_operator.py:362-551) by what its definition fixes."""

import numpy as np
import pytest

import pararealml_amd as pml
from oracle import fdm_numpy as ora
from oracle import supervised_numpy as osup


def _problem():
    mesh = pml.Mesh([(0.0, 1.0)], [0.125])
    bc = pml.DirichletBoundaryCondition(lambda x, t: np.full((len(x), 1), 0.25),
                                        is_static=True)
    cp = pml.ConstrainedProblem(pml.DiffusionEquation(1, 0.1), mesh, [(bc, bc)])
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.4]), np.array([[0.02]]))])
    return cp, ic.discrete_y_0(True)


def test_identity_chain_is_the_oracle_solve_sampled_every_step():
    cp, y0 = _problem()
    inputs, targets = osup.generate_data(cp, y0, (0.0, 0.05), 0.01, 0.0025, "rk4", 2,
                                         lambda _, y: y)
    _, ys = ora.fdm_solve(cp, y0, (0.0, 0.05), 0.0025, "rk4")
    n = 9
    assert inputs.shape == (2 * 5 * n, n + 1) and targets.shape == (2 * 5 * n, 1)
    for it in range(2):
        for i in range(5):
            rows = slice((it * 5 + i) * n, (it * 5 + i + 1) * n)
            np.testing.assert_allclose(targets[rows, 0], ys[4 * (i + 1) - 1][:, 0],
                                       rtol=0, atol=1e-14)
            prev = y0[:, 0] if i == 0 else ys[4 * i - 1][:, 0]
            np.testing.assert_allclose(inputs[rows, :n], np.tile(prev, (n, 1)),
                                       rtol=0, atol=1e-14)
            # the coordinate column
            np.testing.assert_array_equal(inputs[rows, n], np.linspace(0.0, 1.0, n))


def test_isolated_inputs_are_perturbed_unperturbed_states_and_columns():
    cp, y0 = _problem()
    shift = lambda t, y: y + 0.5 + t  # noqa: E731
    inputs, targets = osup.generate_data(cp, y0, (0.0, 0.04), 0.01, 0.005, "fe", 3, shift,
                                         time_variant=True, isolate_perturbations=True)
    _, ys = ora.fdm_solve(cp, y0, (0.0, 0.04), 0.005, "fe")
    n = 9
    assert inputs.shape == (3 * 4 * n, n + 2)
    starts = [y0] + [ys[2 * (i + 1) - 1] for i in range(3)]
    for it in range(3):
        for i in range(4):
            rows = slice((it * 4 + i) * n, (it * 4 + i + 1) * n)
            want = starts[i][:, 0] + 0.5 + 0.01 * i
            want[[0, -1]] = 0.25  # Dirichlet values imposed after the perturbation
            np.testing.assert_allclose(inputs[rows, :n], np.tile(want, (n, 1)),
                                       rtol=0, atol=1e-14)
            np.testing.assert_allclose(inputs[rows, n], 0.01 * (i + 1), rtol=0, atol=1e-15)
            p = want.reshape(-1, 1)
            _, one = ora.fdm_solve(cp, p, (0.01 * i, 0.01 * (i + 1)), 0.005, "fe")
            np.testing.assert_array_equal(targets[rows], one[-1])


def test_whole_solve_mode_and_step_column():
    cp, y0 = _problem()
    scale = lambda t, y: 2.0 * y  # noqa: E731
    inputs, targets = osup.generate_data(cp, y0, (0.0, 0.03), 0.01, 0.005, "midpoint", 2,
                                         scale, auto_regressive=False, time_variant=True)
    p = 2.0 * y0
    p[[0, -1]] = 0.25
    _, ys = ora.fdm_solve(cp, p, (0.0, 0.03), 0.005, "midpoint")
    n = 9
    for it in range(2):
        for i in range(3):
            rows = slice((it * 3 + i) * n, (it * 3 + i + 1) * n)
            np.testing.assert_array_equal(inputs[rows, :n], np.tile(p[:, 0], (n, 1)))
            np.testing.assert_array_equal(targets[rows], ys[2 * (i + 1) - 1])
    inputs, _ = osup.generate_data(cp, y0, (0.0, 0.02), 0.01, 0.005, "fe", 1,
                                   lambda _, y: y, input_d_t=True)
    np.testing.assert_array_equal(inputs[:, n], 0.01)


def test_random_draws_follow_the_seed_in_iteration_order():
    cp, y0 = _problem()
    noise = lambda t, y: y + np.random.normal(0.0, 0.01, y.shape)  # noqa: E731
    a = osup.generate_data(cp, y0, (0.0, 0.02), 0.01, 0.005, "fe", 3, noise,
                           isolate_perturbations=True, seed=7)
    b = osup.generate_data(cp, y0, (0.0, 0.02), 0.01, 0.005, "fe", 3, noise,
                           isolate_perturbations=True, seed=7)
    np.testing.assert_array_equal(a[0], b[0])
    # the draws in the order iteration 0 (steps 0, 1), iteration 1, ...
    np.random.seed(7)
    draws = [np.random.normal(0.0, 0.01, y0.shape) for _ in range(6)]
    n = 9
    for k in range(6):
        got = a[0][k * n, 1:n - 1]
        start = y0 if k % 2 == 0 else ora.fdm_solve(cp, y0, (0.0, 0.01), 0.005, "fe")[1][-1]
        np.testing.assert_allclose(got, (start + draws[k])[1:-1, 0], rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        osup.generate_data(cp, y0, (0.0, 0.02), 0.01, 0.005, "fe", 1,
                           lambda _, y: y[:-1])
