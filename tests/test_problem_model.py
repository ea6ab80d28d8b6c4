"""Host-side problem model: the reference's own behaviour (mesh shapes,
constraint construction and corner precedence, boundary conditions, initial
conditions, solutions, time discretisation), ported as tests of this
package."""

import numpy as np
import pytest

import pararealml_amd as pml
from pararealml_amd.constraint import Constraint
from pararealml_amd.operator import discretize_time_domain


def test_mesh_shapes_and_coordinates():
    mesh = pml.Mesh([(2.0, 6.0), (-3.0, 3.0), (10.0, 12.0)], [0.1, 0.2, 0.5])
    assert mesh.shape(True) == (41, 31, 5)
    assert mesh.shape(False) == (40, 30, 4)
    np.testing.assert_allclose(mesh.vertex_axis_coordinates[1], np.linspace(-3, 3, 31))
    np.testing.assert_allclose(mesh.cell_center_axis_coordinates[2], [10.25, 10.75, 11.25, 11.75])
    assert mesh.volume == pytest.approx(4 * 6 * 2)
    assert mesh.all_index_coordinates(True, flatten=True).shape == (41 * 31 * 5, 3)
    with pytest.raises(ValueError):
        pml.Mesh([(0.0, 1.0)], [0.1, 0.1])
    with pytest.raises(ValueError):
        pml.Mesh([(1.0, 0.0)], [0.1])
    with pytest.raises(ValueError):
        pml.Mesh([(0.0, 1.0)], [-0.1])


def test_polar_mesh_validation():
    m = pml.Mesh([(1.0, 5.0), (0.0, 2 * np.pi)], [2.0, np.pi], pml.CoordinateSystem.POLAR)
    assert m.vertices_shape == (3, 3)
    with pytest.raises(ValueError):
        pml.Mesh([(-1.0, 5.0), (0.0, np.pi)], [1.0, 0.5], pml.CoordinateSystem.POLAR)


def test_constraint_apply_and_multiply_and_add():
    c = Constraint(np.array([1.0, 2.0]), np.array([True, False, True]))
    a = np.zeros((2, 3))
    c.apply(a)
    np.testing.assert_array_equal(a, [[1, 0, 2], [1, 0, 2]])
    res = np.zeros(3)
    c.multiply_and_add(np.array([10.0, 20.0, 30.0]), 2.0, res)
    np.testing.assert_array_equal(res, [12.0, 0.0, 34.0])
    with pytest.raises(ValueError):
        Constraint(np.array([1.0]), np.array([True, True]))
    with pytest.raises(ValueError):
        c.apply(np.zeros(4))


def test_constant_boundary_condition_values():
    bc = pml.ConstantBoundaryCondition([1.0, None], [None, -1.0])
    x = np.zeros((4, 2))
    y = bc.y_condition(x, None)
    assert np.all(y[:, 0] == 1.0)
    assert all(v is None for v in y[:, 1])
    with pytest.raises(ValueError):
        pml.ConstantBoundaryCondition(None, None)
    d = pml.DirichletBoundaryCondition(lambda x, t: np.zeros((len(x), 1)))
    with pytest.raises(RuntimeError):
        d.d_y_condition(x, None)


def test_vectorized_bc_function_nan_for_none():
    f = pml.vectorize_bc_function(lambda x, t: (x[0], -1.0, 2 * x[1], None, t**2))
    out = f(np.array([[0.0, 1.0], [2.0, 3.0]]), 3.0)
    np.testing.assert_array_equal(
        out, [[0.0, -1.0, 2.0, np.nan, 9.0], [2.0, -1.0, 6.0, np.nan, 9.0]]
    )


def test_constrained_problem_corner_precedence_and_gradients():
    """Later axes overwrite earlier ones at corners (the reference's
    tests/test_constrained_problem.py:80-163 case)."""
    diff_eq = pml.WaveEquation(2)
    mesh = pml.Mesh([(2.0, 6.0), (-3.0, 3.0)], [0.1, 0.2])
    bcs = (
        (
            pml.DirichletBoundaryCondition(
                pml.vectorize_bc_function(lambda x, t: (999.0, None)), is_static=True
            ),
            pml.NeumannBoundaryCondition(
                pml.vectorize_bc_function(lambda x, t: (100.0, -100.0)), is_static=True
            ),
        ),
        (
            pml.NeumannBoundaryCondition(
                pml.vectorize_bc_function(lambda x, t: (-x[0], None)), is_static=True
            ),
            pml.DirichletBoundaryCondition(
                pml.vectorize_bc_function(lambda x, t: (x[0], x[1])), is_static=True
            ),
        ),
    )
    cp = pml.ConstrainedProblem(diff_eq, mesh, bcs)
    assert cp.are_all_boundary_conditions_static
    assert cp.are_there_boundary_conditions_on_y
    y = np.full(cp.y_shape(True), 13.0)
    pml.apply_constraints_along_last_axis(cp.static_y_vertex_constraints, y)
    assert np.all(y[0, :-1, 0] == 999.0)
    assert np.all(y[0, :-1, 1] == 13.0)
    assert np.all(y[-1, :-1, :] == 13.0)
    assert np.all(y[1:, 0, :] == 13.0)
    np.testing.assert_allclose(y[:, -1, 0], np.linspace(2.0, 6.0, y.shape[0]))
    assert np.all(y[:, -1, 1] == 3.0)
    cells = cp.static_boundary_cell_constraints[0]
    np.testing.assert_allclose(cells[1, 0][1].values, np.linspace(2.05, 5.95, cp.y_cells_shape[0]))
    assert cells[1, 0][0] is None


def test_constrained_problem_dynamic_and_errors():
    diff_eq = pml.DiffusionEquation(2)
    mesh = pml.Mesh([(0.0, 5.0), (-5.0, 5.0)], [0.1, 0.2])
    dyn = [(pml.DirichletBoundaryCondition(lambda x, t: np.full((len(x), 1), t)),) * 2] * 2
    cp = pml.ConstrainedProblem(diff_eq, mesh, dyn)
    assert not cp.are_all_boundary_conditions_static
    y_bcs, _ = cp.create_boundary_constraints(True, 2.0)
    assert np.all(y_bcs[0, 0][0].values == 2.0)
    bad = [(pml.DirichletBoundaryCondition(lambda x, t: np.zeros((13, 1)), is_static=True),) * 2] * 2
    with pytest.raises(ValueError):
        pml.ConstrainedProblem(diff_eq, mesh, bad)
    with pytest.raises(ValueError):
        pml.ConstrainedProblem(diff_eq, None, dyn)


def test_ode_constrained_problem():
    cp = pml.ConstrainedProblem(pml.LotkaVolterraEquation())
    assert cp.mesh is None and cp.static_y_vertex_constraints is None
    assert cp.y_shape(True) == (2,)
    assert not cp.are_all_boundary_conditions_static


def test_gaussian_initial_condition_kat(kats):
    k = kats["gaussian_ic_2d"]
    diff_eq = pml.WaveEquation(2)
    mesh = pml.Mesh([(0.0, 2.0), (0.0, 2.0)], [1.0, 1.0])
    mk = lambda a, b: pml.DirichletBoundaryCondition(  # noqa: E731
        pml.vectorize_bc_function(lambda x, t: (a, b)), is_static=True
    )
    cp = pml.ConstrainedProblem(diff_eq, mesh, [(mk(0.0, 2.0), mk(1.0, 2.0)), (mk(3.0, 2.0), mk(4.0, 2.0))])
    ic = pml.GaussianInitialCondition(
        cp, [(np.array(m), np.array(c)) for m, c in k["means_covs"]], k["multipliers"]
    )
    np.testing.assert_allclose(ic.y_0(np.array(k["x"])), k["expected_y0"], rtol=1e-7)
    np.testing.assert_allclose(ic.discrete_y_0(True), k["expected_vertex"], rtol=1e-7)
    np.testing.assert_allclose(ic.discrete_y_0(False), k["expected_cell"], rtol=1e-7)


def test_discrete_and_constant_initial_conditions():
    diff_eq = pml.DiffusionEquation(1)
    mesh = pml.Mesh([(0.0, 1.0)], [0.25])
    bc = pml.DirichletBoundaryCondition(lambda x, t: np.full((len(x), 1), 7.0), is_static=True)
    cp = pml.ConstrainedProblem(diff_eq, mesh, [(bc, bc)])
    ic = pml.DiscreteInitialCondition(cp, np.ones((5, 1)), True)
    np.testing.assert_array_equal(ic.discrete_y_0(True)[:, 0], [7, 1, 1, 1, 7])
    c = pml.ConstantInitialCondition(cp, [3.0])
    np.testing.assert_array_equal(c.discrete_y_0(True)[:, 0], [7, 3, 3, 3, 7])
    with pytest.raises(ValueError):
        pml.DiscreteInitialCondition(cp, np.ones((4, 1)), True)


def test_discretize_time_domain(kats):
    k = kats["discretize_time_domain"]
    for interval, dt in k["cases"]:
        np.testing.assert_allclose(discretize_time_domain(interval, dt), k["expected"])


def test_solution_shapes_interpolation_and_diff():
    diff_eq = pml.DiffusionEquation(1)
    mesh = pml.Mesh([(0.0, 1.0)], [0.5])
    bc = pml.NeumannBoundaryCondition(lambda x, t: np.zeros((len(x), 1)), is_static=True)
    cp = pml.ConstrainedProblem(diff_eq, mesh, [(bc, bc)])
    ivp = pml.InitialValueProblem(cp, (0.0, 1.0), pml.ConstantInitialCondition(cp, [0.0]))
    y = np.arange(6, dtype=float).reshape(2, 3, 1)
    sol = pml.Solution(ivp, np.array([0.5, 1.0]), y, vertex_oriented=True)
    assert sol.d_t == 0.5
    np.testing.assert_allclose(sol.discrete_y(False)[:, :, 0], [[0.5, 1.5], [3.5, 4.5]])
    other = pml.Solution(ivp, np.array([0.25, 0.5, 0.75, 1.0]), np.ones((4, 3, 1)), vertex_oriented=True)
    d = sol.diff([other])
    np.testing.assert_allclose(d.matching_time_points, [0.5, 1.0])
    np.testing.assert_allclose(d.differences[0][0][:, 0], 1 - y[0, :, 0])
    with pytest.raises(ValueError):
        pml.Solution(ivp, np.array([]), np.zeros((0, 3, 1)), vertex_oriented=True)
    with pytest.raises(ValueError):
        pml.Solution(ivp, np.array([1.0]), np.zeros((1, 4, 1)), vertex_oriented=True)


def test_operator_and_ivp_validation():
    from pararealml_amd.operators.fdm import ForwardEulerMethod, FDMOperator, ThreePointCentralDifferenceMethod

    with pytest.raises(ValueError):
        FDMOperator(ForwardEulerMethod(), ThreePointCentralDifferenceMethod(), 0.0)
    with pytest.raises(ValueError):
        FDMOperator(ForwardEulerMethod(), ThreePointCentralDifferenceMethod(), 0.1, precision="half")
    with pytest.raises(ValueError):
        ThreePointCentralDifferenceMethod(-0.1)
    cp = pml.ConstrainedProblem(pml.LotkaVolterraEquation())
    with pytest.raises(ValueError):
        pml.InitialValueProblem(cp, (1.0, 0.0), pml.ContinuousInitialCondition(cp, lambda _: np.ones(2)))


def test_integrator_integral_api_on_host_arrays(kats):
    """The integrators' integral() keeps the reference's semantics for
    caller-supplied right-hand sides (tests/operators/fdm/
    test_numerical_integrator.py:13-138)."""
    from pararealml_amd.operators.fdm import RK4, ExplicitMidpointMethod, ForwardEulerMethod

    k = kats["integrators"]
    fe = ForwardEulerMethod().integral(np.array([1.0]), 1.0, 0.5, lambda t, y: 5 * y + t**2, lambda _: None)
    np.testing.assert_allclose(fe, k["forward_euler"]["expected"])
    mp = ExplicitMidpointMethod().integral(np.array([2.0]), 0.0, 0.5, lambda t, y: 2 * y - 4 * t, lambda _: None)
    np.testing.assert_allclose(mp, k["explicit_midpoint"]["expected"])
    value = np.full((8, 1), np.nan)
    value[0] = value[-1] = 0.0
    mask = ~np.isnan(value)
    c = Constraint(value[mask], mask)
    out = RK4().integral(np.zeros((8, 1)), 1.0, 1.0, lambda _, y: 2 * y + 1, lambda _: [c])
    np.testing.assert_allclose(out[1:-1, 0], 3.0)
    assert out[0, 0] == 0.0 and out[-1, 0] == 0.0


def test_equation_validation():
    with pytest.raises(ValueError):
        pml.DiffusionEquation(0)
    with pytest.raises(ValueError):
        pml.ConvectionDiffusionEquation(2, [1.0])
    with pytest.raises(ValueError):
        pml.NBodyGravitationalEquation(4, [1.0, 2.0])
    eq = pml.CahnHilliardEquation(2)
    system = eq.symbolic_equation_system
    assert system.equation_indices_by_type(pml.LHS.Y) == [1]
    assert eq.device_parameters() == ("cahn_hilliard", (0.1, 0.01))
    assert pml.NBodyGravitationalEquation(2, [1.0, 2.0]).y_dimension == 8
