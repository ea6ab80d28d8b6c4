"""Pins the oracle's time loop, grid transfers, Parareal emulation and Adam
step by facts the mathematics fixes, independently of the oracle's own
restatement of the reference:

* ``fdm_solve`` on an ODE: the reference's closed-form test
  (tests/operators/fdm/test_fdm_operator.py:38-61, transcribed as the
  ``fdm_exponential_ode`` known answer) and, tighter, RK4's amplification
  polynomial raised to the step count;
* ``fdm_solve`` on 2-D diffusion with zero Dirichlet edges: a discrete sine
  mode is an eigenvector of the vertex-grid Laplacian, so every step
  multiplies it by the integrator's amplification factor of
  ``z = dt d lambda_km``;
* ``fdm_solve`` under zero-Neumann (reflective-ghost) edges: the
  trapezoid-weighted sum of the state telescopes to zero change per step,
  for diffusion and for Cahn-Hilliard's ``c``;
* the Cahn-Hilliard ``LHS.Y`` lag (fdm_operator.py:127-139): after one RK4
  step ``c' = c + dt d lap(mu)`` and ``mu' = c^3 - c - gamma lap(c)``, both
  from the pre-step state, with the KAT-pinned Laplacian;
* ``restrict_injection`` / ``prolong_bilinear``: exact on bilinear fields,
  restriction after prolongation is the identity;
* ``parareal_solve``: one slice equals the fine solve, N iterations equal the
  serial fine solution, G = F terminates after the first iteration, and the
  iteration count is the first k whose update is below the tolerance;
* ``adam_step``: the first bias-corrected step is ``lr g / (|g| + eps)`` and
  a constant gradient keeps that step size at every step.
"""

import json
from pathlib import Path

import numpy as np
import pytest

import pararealml_amd as pml
from oracle import fdm_numpy as ora
from oracle import surrogate_numpy as so
from tests.helpers import cahn_hilliard, zero_bc

ROOT = Path(__file__).resolve().parent.parent

AMPLIFICATION = {
    "fe": lambda z: 1.0 + z,
    "midpoint": lambda z: 1.0 + z + z * z / 2.0,
    "rk4": lambda z: 1.0 + z + z * z / 2.0 + z**3 / 6.0 + z**4 / 24.0,
}


def test_fdm_solve_exponential_ode_known_answer(kats):
    kat = kats["fdm_exponential_ode"]
    cp = pml.ConstrainedProblem(pml.PopulationGrowthEquation(kat["r"]))
    t, y = ora.fdm_solve(cp, np.array([kat["y0"]]), tuple(kat["t_interval"]), kat["d_t"], "rk4")
    assert y.shape == (kat["n_steps"], 1)
    np.testing.assert_allclose(t[-1], kat["t_interval"][1], rtol=1e-12)
    # the reference's assertion: np.allclose against y0 e^{r t}
    assert np.allclose(kat["y0"] * np.exp(kat["r"] * t), y[:, 0])
    # tighter: n steps of RK4 multiply y0 by P(r dt)^n exactly (to rounding)
    amp = AMPLIFICATION["rk4"](kat["r"] * kat["d_t"])
    n = np.arange(1, kat["n_steps"] + 1)
    np.testing.assert_allclose(y[:, 0], kat["y0"] * amp**n, rtol=1e-11, atol=0)


@pytest.mark.parametrize("integrator", ["fe", "midpoint", "rk4"])
def test_fdm_solve_dirichlet_sine_mode_amplification(integrator):
    n, length, d = 33, 2.0, 0.7
    dx = length / (n - 1)
    eq = pml.DiffusionEquation(2, d)
    mesh = pml.Mesh([(0.0, length), (0.0, length)], [dx, dx])
    bc = zero_bc(1)
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc), (bc, bc)])
    k, m = 3, 5
    x0, x1 = mesh.vertex_coordinate_grids
    mode = np.sin(k * np.pi * x0 / length) * np.sin(m * np.pi * x1 / length)
    y0 = mode[..., np.newaxis]
    lam = -(4.0 / dx**2) * (
        np.sin(k * np.pi * dx / (2 * length)) ** 2 + np.sin(m * np.pi * dx / (2 * length)) ** 2
    )
    d_t = 0.2 * dx**2 / d
    steps = 40
    _, y = ora.fdm_solve(cp, y0, (0.0, steps * d_t), d_t, integrator)
    amp = AMPLIFICATION[integrator](d_t * d * lam)
    inner = (slice(1, -1), slice(1, -1), 0)
    for s in (0, steps // 2, steps - 1):
        expected = amp ** (s + 1) * y0[inner]
        np.testing.assert_allclose(y[s][inner], expected, rtol=0, atol=1e-12)
        # boundary vertices hold the Dirichlet value
        assert np.all(y[s][0] == 0.0) and np.all(y[s][:, -1] == 0.0)


def _trapezoid_weights(shape):
    w = []
    for nv in shape:
        v = np.ones(nv)
        v[0] = v[-1] = 0.5
        w.append(v)
    return np.multiply.outer(*w)


@pytest.mark.parametrize("integrator", ["fe", "rk4"])
def test_fdm_solve_zero_neumann_conserves_trapezoid_mass(integrator):
    eq = pml.DiffusionEquation(2, 1.3)
    mesh = pml.Mesh([(0.0, 1.9), (0.0, 2.3)], [0.1, 0.1])
    bc = zero_bc(1, "neumann")
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc), (bc, bc)])
    y0 = np.random.RandomState(4).uniform(0.0, 1.0, mesh.vertices_shape + (1,))
    w = _trapezoid_weights(mesh.vertices_shape)
    mass0 = np.sum(w * y0[..., 0])
    _, y = ora.fdm_solve(cp, y0, (0.0, 0.04), 1e-3, integrator)
    masses = np.array([np.sum(w * s[..., 0]) for s in y])
    np.testing.assert_allclose(masses, mass0, rtol=1e-13, atol=0)
    # and it does change the state (not a trivially conserved fixed point)
    assert np.max(np.abs(y[-1] - y0)) > 1e-2


def test_fdm_solve_cahn_hilliard_lag_relations_and_mass():
    cp, ic = cahn_hilliard(24, seed=5)
    d, gamma = 0.1, 0.01
    y0 = ic.discrete_y_0(True)
    d_x = cp.mesh.d_x
    dbc = cp.static_boundary_vertex_constraints[1]
    d_t = 1e-4
    _, y = ora.fdm_solve(cp, y0, (0.0, 3 * d_t), d_t, "rk4")
    prev = y0
    w = _trapezoid_weights(cp.mesh.vertices_shape)
    for s in range(3):
        c, mu = prev[..., 0:1], prev[..., 1:2]
        lap_mu = ora.laplacian(mu, d_x, dbc[:, 1:2])
        lap_c = ora.laplacian(c, d_x, dbc[:, 0:1])
        np.testing.assert_allclose(y[s][..., 0:1], c + d_t * d * lap_mu, rtol=0, atol=1e-15)
        np.testing.assert_allclose(y[s][..., 1:2], c**3 - c - gamma * lap_c, rtol=0,
                                   atol=1e-15)
        np.testing.assert_allclose(np.sum(w * y[s][..., 0]), np.sum(w * y0[..., 0]),
                                   rtol=0, atol=1e-15)
        prev = y[s]


def test_grid_transfers_exact_on_bilinear_fields():
    m0, m1 = 9, 7
    rng = np.random.RandomState(0)
    a, b, c, e = rng.standard_normal(4)
    fx0, fx1 = np.meshgrid(np.linspace(0, 1, 2 * m0 - 1), np.linspace(-1, 2, 2 * m1 - 1),
                           indexing="ij")
    fine = (a + b * fx0 + c * fx1 + e * fx0 * fx1)[..., np.newaxis]
    coarse = ora.restrict_injection(fine)
    assert coarse.shape == (m0, m1, 1)
    np.testing.assert_array_equal(coarse, fine[::2, ::2])
    np.testing.assert_allclose(ora.prolong_bilinear(coarse), fine, rtol=0, atol=1e-14)
    # restriction after prolongation is the identity, for any coarse field
    v = rng.standard_normal((m0, m1, 2))
    np.testing.assert_array_equal(ora.restrict_injection(ora.prolong_bilinear(v)), v)
    # midpoints are the mean of their two neighbours (a non-bilinear field)
    p = ora.prolong_bilinear(v)
    np.testing.assert_allclose(p[1, 0], 0.5 * (v[0, 0] + v[1, 0]), rtol=0, atol=1e-15)
    np.testing.assert_allclose(p[1, 1], 0.25 * (v[0, 0] + v[1, 0] + v[0, 1] + v[1, 1]),
                               rtol=0, atol=1e-15)


def test_cubic_prolongation_exact_on_cubics():
    m0, m1 = 7, 9
    rng = np.random.RandomState(1)
    coef = rng.standard_normal((4, 4))
    g0, g1 = np.linspace(-1, 2, 2 * m0 - 1), np.linspace(0, 1, 2 * m1 - 1)
    X, Y = np.meshgrid(g0, g1, indexing="ij")
    fine = sum(coef[a, b] * X**a * Y**b for a in range(4) for b in range(4))[..., None]
    np.testing.assert_allclose(ora.prolong_cubic(fine[::2, ::2]), fine, rtol=0, atol=1e-12)
    # the 4-point midpoint weights, one-sided at the ends (a non-polynomial field)
    v = rng.standard_normal((m0, m1, 1))
    p = ora.prolong_cubic(v)
    np.testing.assert_array_equal(ora.restrict_injection(p), v)
    np.testing.assert_allclose(p[3, 0], (-v[0, 0] + 9 * v[1, 0] + 9 * v[2, 0] - v[3, 0]) / 16,
                               rtol=0, atol=1e-15)
    np.testing.assert_allclose(p[1, 0], (5 * v[0, 0] + 15 * v[1, 0] - 5 * v[2, 0] + v[3, 0]) / 16,
                               rtol=0, atol=1e-15)
    np.testing.assert_allclose(p[-2, 0],
                               (5 * v[-1, 0] + 15 * v[-2, 0] - 5 * v[-3, 0] + v[-4, 0]) / 16,
                               rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        ora.prolong_cubic(np.zeros((3, 5, 1)))


def _lorenz_setup():
    cp = pml.ConstrainedProblem(pml.LorenzEquation())
    y0 = np.array([1.0, 1.0, 1.0])

    def solver(integrator, d_t):
        def run(y, t0, t1):
            return ora.fdm_solve(cp, y, (t0, t1), d_t, integrator)[1]
        return run

    return y0, solver


def test_parareal_emulation_invariants():
    y0, solver = _lorenz_setup()
    f, g = solver("rk4", 1e-3), solver("fe", 1e-2)
    n = 5
    ident = lambda y: y  # noqa: E731
    serial = f(y0, 0.0, 0.5)
    # one slice: the fine solve itself
    _, traj1, it1, _ = ora.parareal_solve(y0, (0.0, 0.5), 1, f, g, 1e-3, 1e-2, ident, 0.0)
    assert it1 == 1
    # (the output shift U[1] - F[-1] = (G + (F - G)) - F is zero up to rounding)
    np.testing.assert_allclose(traj1, serial, rtol=1e-15, atol=1e-15)
    # N iterations: the serial fine solution (slices chain exactly)
    _, traj_n, it_n, u = ora.parareal_solve(y0, (0.0, 0.5), n, f, g, 1e-3, 1e-2, ident, 0.0)
    assert it_n == n
    np.testing.assert_allclose(traj_n, serial, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(u[1:], serial[99::100], rtol=1e-12, atol=1e-12)
    # G = F: the corrections are exactly zero, the update is zero and the
    # first iteration ends the loop with the serial fine solution
    _, traj_ff, it_ff, _ = ora.parareal_solve(y0, (0.0, 0.5), n, f, f, 1e-3, 1e-3, ident, 1e-12)
    assert it_ff == 1
    np.testing.assert_array_equal(traj_ff, serial)


def test_parareal_emulation_stops_at_first_small_update():
    y0, solver = _lorenz_setup()
    f, g = solver("rk4", 1e-3), solver("fe", 1e-2)
    n = 5
    ident = lambda y: y  # noqa: E731
    updates = []

    def record(old, new):
        # a scalar tolerance compares, per field, the largest update over
        # slices (the RMS over a slice's single ODE point is its modulus)
        updates.append(float(np.max(np.abs(new - old))))
        return False

    _, _, it, _ = ora.parareal_solve(y0, (0.0, 0.5), n, f, g, 1e-3, 1e-2, ident, record)
    assert it == n and len(updates) == n
    k = next(k for k in range(1, n) if updates[k] < min(updates[:k]))
    tol = 0.5 * (updates[k] + min(updates[:k]))
    _, _, it_tol, _ = ora.parareal_solve(y0, (0.0, 0.5), n, f, g, 1e-3, 1e-2, ident, tol)
    assert it_tol == k + 1
    # a per-field sequence with the same value everywhere behaves the same
    _, _, it_seq, _ = ora.parareal_solve(y0, (0.0, 0.5), n, f, g, 1e-3, 1e-2, ident,
                                         [tol] * 3)
    assert it_seq == k + 1


def test_parareal_increment_form_equals_the_reference_update_for_linear_g():
    """U_new[j+1] = F(U_old[j]) + G(U_new[j] - U_old[j]) is the reference's
    G(U_new[j]) + F(U_old[j]) - G(U_old[j]) when G is linear and the
    constraints homogeneous: 2-D diffusion, zero Dirichlet edges, G = FE on
    the 2x grid with injection and cubic prolongation."""
    n = 17
    eq = pml.DiffusionEquation(2, 0.5)
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.0)], [1.0 / (n - 1)] * 2)
    bc = zero_bc(1)
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc), (bc, bc)])
    cmesh = pml.Mesh(mesh.x_intervals, [2.0 / (n - 1)] * 2)
    ccp = pml.ConstrainedProblem(eq, cmesh, [(bc, bc), (bc, bc)])
    y0 = pml.GaussianInitialCondition(
        cp, [(np.array([0.4, 0.55]), 0.01 * np.eye(2))]).discrete_y_0(True)

    def cons(y):
        return ora.apply_constraints_last_axis(cp.static_y_vertex_constraints, y)

    def f(y, a, b):
        return ora.fdm_solve(cp, y, (a, b), 2e-4, "rk4")[1]

    def g(y, a, b):
        yc = ora.restrict_injection(y).copy()
        traj = ora.fdm_solve(ccp, yc, (a, b), 1e-3, "fe")[1]
        return np.stack([cons(ora.prolong_cubic(s)) for s in traj])

    # (two of four iterations: after four both are the serial fine solution)
    runs = [ora.parareal_solve(y0, (0.0, 0.04), 4, f, g, 2e-4, 1e-3, cons, 0.0,
                               max_iterations=2, coarse_increments=inc) for inc in (False, True)]
    assert runs[0][2] == runs[1][2] == 2
    serial = f(y0, 0.0, 0.04)[49::50]
    assert np.max(np.abs(runs[0][3][1:] - serial)) > 1e-9  # not yet converged
    np.testing.assert_allclose(runs[1][3], runs[0][3], rtol=0,
                               atol=1e-12 * np.abs(runs[0][3]).max())
    # and it differs from the update when G is not linear (G(0) != 0)
    shifted = ora.parareal_solve(y0, (0.0, 0.04), 4, f, lambda y, a, b: g(y, a, b) + 1e-3,
                                 2e-4, 1e-3, cons, 0.0, max_iterations=1,
                                 coarse_increments=True)
    plain = ora.parareal_solve(y0, (0.0, 0.04), 4, f, lambda y, a, b: g(y, a, b) + 1e-3,
                               2e-4, 1e-3, cons, 0.0, max_iterations=1)
    assert np.max(np.abs(shifted[3] - plain[3])) > 1e-4


def test_adam_step_closed_forms():
    params = so.init_params(1, 1, seed=3)
    rng = np.random.RandomState(7)
    grads = [(rng.standard_normal(w.shape), rng.standard_normal(b.shape)) for w, b in params]
    masters = [(w.copy(), b.copy()) for w, b in params]
    zeros = [(np.zeros_like(w), np.zeros_like(b)) for w, b in params]
    state = (masters, zeros, [(z0.copy(), z1.copy()) for z0, z1 in zeros], 0)
    lr, eps = 1e-3, 1e-8
    new_params, state1 = so.adam_step(params, grads, state, lr, eps=eps)
    for (w0, b0), (gw, gb), (w1, b1) in zip(masters, grads, state1[0]):
        np.testing.assert_allclose(w1, w0 - lr * gw / (np.abs(gw) + eps), rtol=0, atol=1e-15)
        np.testing.assert_allclose(b1, b0 - lr * gb / (np.abs(gb) + eps), rtol=0, atol=1e-15)
    # layer 0 keeps its FP32 master, the others are BF16-rounded
    np.testing.assert_array_equal(new_params[0][0], state1[0][0][0])
    np.testing.assert_array_equal(new_params[1][0], so.bf16(state1[0][1][0]))
    # a constant gradient: bias correction keeps the step at lr g / (|g| + eps)
    st = state1
    for step in range(2, 5):
        before = [w.copy() for w, _ in st[0]]
        _, st = so.adam_step(params, grads, st, lr, eps=eps)
        assert st[3] == step
        for w_prev, (w_new, _), (gw, _) in zip(before, st[0], grads):
            np.testing.assert_allclose(w_new - w_prev, -lr * gw / (np.abs(gw) + eps),
                                       rtol=1e-9, atol=1e-18)


def test_oracle_header_names_every_pin():
    text = (ROOT / "oracle" / "fdm_numpy.py").read_text()
    assert "tests/test_oracle_pins.py" in text
    kats = json.loads((ROOT / "tests" / "golden" / "reference_kats.json").read_text())
    assert kats["fdm_exponential_ode"]["n_steps"] == 100000
