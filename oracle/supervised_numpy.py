"""NumPy restatement of the reference's supervised-learning data protocol.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``generate_data`` restates SupervisedMLOperator.generate_data with one job
(pararealml/operators/ml/supervised/supervised_ml_operator.py:152-245) and
its worker ``_generate_data`` (:381-499) with ``_perturb_and_solve_ivp``
(:501-551) and ``_create_input_placeholder`` (:362-379), for vertex-oriented
PDE problems, the oracle operator being FDM (``oracle.fdm_numpy.fdm_solve``):

* the operator's time grid ``t = discretize_time_domain(ivp.t_interval,
  d_t)``;
* auto-regressive mode: per iteration a chain of ``len(t) - 1`` sub-IVPs
  ``(t_i, t_i + d_t)``; each constrains its initial state, perturbs it with
  ``perturbation_function(t_i, y)``, constrains the result, solves it with
  the FDM oracle and records (perturbed state, final state); the next
  sub-IVP starts from that final state or, with ``isolate_perturbations``,
  from the unperturbed solution at ``t_{i+1}`` (one oracle solve over
  ``(t_0, t_{-2})`` before the iterations);
* otherwise one perturbed solve of the whole IVP per iteration, the targets
  at ``rint((t[1:] - t_0) / oracle_d_t) - 1``;
* inputs: one row per (iteration, time step, spatial point): the recorded
  state flattened channel-last, then the time (``time_variant``) or the
  operator's step (``input_d_t``), then the point's coordinates; targets:
  one row per (iteration, time step, point) with the final state's values
  there.

Pinning: with the identity perturbation the chained targets equal the
oracle's own solve sampled every ``d_t`` and each chain step's input equals
the previous step's target; with ``isolate_perturbations`` the inputs are
the perturbed unperturbed solution; the coordinate, time and step columns,
the row counts and the non-auto-regressive targets are checked against
their definitions -- tests/test_oracle_supervised.py.
"""

from typing import Callable, Optional

import numpy as np

from oracle import fdm_numpy as ora


def _placeholder(cp, time_column: bool) -> np.ndarray:
    x = cp.mesh.all_index_coordinates(True, flatten=True)
    y = np.empty((len(x), cp.differential_equation.y_dimension * len(x)))
    if time_column:
        return np.hstack([y, np.empty((len(x), 1)), x])
    return np.hstack([y, x])


def _perturb_and_solve(cp, y_0, t_interval, perturbation, oracle_d_t, method):
    """(perturbed and constrained initial state, final state)."""
    y_c = ora.apply_constraints_last_axis(cp.static_y_vertex_constraints, np.array(y_0))
    p = perturbation(t_interval[0], y_c)
    if p.shape != y_c.shape:
        raise ValueError(f"perturbed y shape {p.shape} must match input y shape {y_c.shape}")
    p = ora.apply_constraints_last_axis(cp.static_y_vertex_constraints, np.array(p))
    _, ys = ora.fdm_solve(cp, p, t_interval, oracle_d_t, method)
    return p, ys


def generate_data(cp, y_0: np.ndarray, t_interval, d_t: float, oracle_d_t: float,
                  method: str, iterations: int,
                  perturbation: Callable[[float, np.ndarray], np.ndarray],
                  auto_regressive: bool = True, time_variant: bool = False,
                  input_d_t: bool = False, isolate_perturbations: bool = False,
                  seed: Optional[int] = None):
    if seed is not None:
        np.random.seed(seed)
    y_dim = cp.differential_equation.y_dimension
    x_dim = cp.differential_equation.x_dimension
    t = ora.discretize_time_domain(t_interval, d_t)

    unperturbed = None
    if auto_regressive and isolate_perturbations:
        _, ys = _perturb_and_solve(cp, y_0, (t[0], t[-2]), lambda _, y: y, oracle_d_t,
                                   method)
        unperturbed = ys[np.rint((t[1:-1] - t[0]) / oracle_d_t).astype(int) - 1]

    single = _placeholder(cp, time_variant or input_d_t)
    n_pts = single.shape[0]
    epoch = np.tile(single, (len(t) - 1, 1))
    if time_variant:
        epoch[:, -x_dim - 1] = np.repeat(t[1:], n_pts)
    elif input_d_t:
        epoch[:, -x_dim - 1] = d_t
    inputs = np.tile(epoch, (iterations, 1))
    targets = np.empty((inputs.shape[0], y_dim))
    n_y = y_dim * n_pts
    for it in range(iterations):
        offset = it * n_pts * (len(t) - 1)
        if auto_regressive:
            y_i = y_0
            for i, t_i in enumerate(t[:-1]):
                p, ys = _perturb_and_solve(cp, y_i, (t_i, t_i + d_t), perturbation,
                                           oracle_d_t, method)
                rows = slice(offset + i * n_pts, offset + (i + 1) * n_pts)
                inputs[rows, :n_y] = p.reshape((1, -1))
                targets[rows] = ys[-1].reshape((-1, y_dim))
                if isolate_perturbations and i < len(t) - 2:
                    y_i = unperturbed[i]
                else:
                    y_i = ys[-1]
        else:
            p, ys = _perturb_and_solve(cp, y_0, tuple(t_interval), perturbation, oracle_d_t,
                                       method)
            rows = slice(offset, offset + (len(t) - 1) * n_pts)
            inputs[rows, :inputs.shape[1] - x_dim - time_variant] = p.reshape((1, -1))
            targets[rows] = ys[np.rint((t[1:] - t[0]) / oracle_d_t).astype(int) - 1].reshape(
                (-1, y_dim))
    return inputs, targets
