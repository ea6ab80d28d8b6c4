"""Device evaluation of symbolic right-hand sides (the compiled programs of
``pararealml_amd.operators.device_program`` run by ``pml_expr_eval`` and
``pml_ode_solve``) against the oracle, which lambdifies the same expressions
as the reference (operators/symbol_mapper.py:222-253):

* ODE systems through FDMOperator.solve: the reference's closed-form test
  (tests/operators/fdm/test_fdm_operator.py:38-61) and shape test (96-110),
  every reference ODE equation under FE, midpoint and RK4 (FP64 to 1e-11
  relative, FP32 to a stated tolerance), the examples' settings;
* PDEs without a fused kernel: the reference's t- and x-dependent
  right-hand side with time-dependent Neumann edges (test_fdm_operator.py:
  322-355), a user system using every derivative symbol (gradient, Hessian,
  Laplacian, divergence, curl, vector Laplacian) in 2-D and a 3-D one;
* the equations with fused kernels forced through the symbolic path,
  against the fused kernels (diffusion, wave, convection-diffusion,
  Burgers, shallow water, Cahn-Hilliard with LHS.Y, Navier-Stokes with
  LHS.Y_LAPLACIAN);
* Parareal over ODE systems (the reference's van der Pol, Lorenz and
  n-body examples, scaled down) against the oracle's emulation.
"""

import numpy as np
import pytest
import sympy as sp

import pararealml_amd as pml
from oracle import fdm_numpy as ora
from pararealml_amd.differential_equation import DifferentialEquation, SymbolicEquationSystem
from pararealml_amd.operators.fdm import (
    RK4,
    ExplicitMidpointMethod,
    FDMOperator,
    ForwardEulerMethod,
    ThreePointCentralDifferenceMethod,
)
from pararealml_amd.operators.parareal import PararealOperator

pytestmark = pytest.mark.gpu

DIFF = ThreePointCentralDifferenceMethod()
INTEGRATORS = {"fe": ForwardEulerMethod, "midpoint": ExplicitMidpointMethod, "rk4": RK4}


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def _ode_ivp(eq, y0, t_interval):
    cp = pml.ConstrainedProblem(eq)
    ic = pml.ContinuousInitialCondition(cp, lambda _: np.asarray(y0, dtype=float))
    return cp, pml.InitialValueProblem(cp, t_interval, ic)


def test_ode_exponential_known_answer(kats):
    kat = kats["fdm_exponential_ode"]
    print("case: ode exponential kat", flush=True)
    cp, ivp = _ode_ivp(pml.PopulationGrowthEquation(kat["r"]), [kat["y0"]],
                       tuple(kat["t_interval"]))
    sol = FDMOperator(RK4(), DIFF, kat["d_t"]).solve(ivp)
    y = sol.discrete_y()
    assert sol.d_t == kat["d_t"] and y.shape == (kat["n_steps"], 1)
    analytic = kat["y0"] * np.exp(kat["r"] * sol.t_coordinates)
    assert np.allclose(analytic, y[:, 0])
    # RK4's amplification polynomial, step by step
    z = kat["r"] * kat["d_t"]
    amp = 1 + z + z * z / 2 + z**3 / 6 + z**4 / 24
    np.testing.assert_allclose(y[:, 0], kat["y0"] * amp ** np.arange(1, kat["n_steps"] + 1),
                               rtol=1e-11)


def test_ode_lorenz_shape_as_the_reference_test():
    print("case: ode lorenz shape", flush=True)
    _, ivp = _ode_ivp(pml.LorenzEquation(), np.ones(3), (0.0, 10.0))
    sol = FDMOperator(ForwardEulerMethod(), DIFF, 0.01).solve(ivp)
    assert sol.vertex_oriented and sol.d_t == 0.01
    assert sol.discrete_y().shape == (1000, 3)
    assert np.all(np.isfinite(sol.discrete_y()))


ODE_CASES = {
    "population": (lambda: pml.PopulationGrowthEquation(0.3), [2.0], 3.0, 1e-3),
    "lotka_volterra": (lambda: pml.LotkaVolterraEquation(), [100.0, 15.0], 2.0, 1e-3),
    "lorenz": (lambda: pml.LorenzEquation(), [1.0, 1.0, 1.0], 2.0, 1e-3),
    "sir": (lambda: pml.SIREquation(), [999.0, 1.0, 0.0], 20.0, 1e-2),
    "van_der_pol": (lambda: pml.VanDerPolEquation(1.5), [1.0, 0.0], 3.0, 1e-3),
    "n_body_2d": (lambda: pml.NBodyGravitationalEquation(2, [5.0, 3.0, 1.0], g=1.0),
                  [0.0, 0.0, 1.0, 0.0, 0.0, 1.5, 0.0, 0.0, 0.0, 1.2, -0.8, 0.0], 1.0, 1e-3),
}


def _n_body_3d():
    rng = np.random.RandomState(3)
    masses = list(rng.uniform(5e4, 5e8, 5))
    y0 = np.append(40 * rng.rand(15) - 20.0, 5 * rng.rand(15))
    return pml.NBodyGravitationalEquation(3, masses), y0


@pytest.mark.parametrize("integ", sorted(INTEGRATORS))
@pytest.mark.parametrize("name", sorted(ODE_CASES) + ["n_body_3d"])
def test_ode_systems_match_oracle(name, integ):
    print(f"case: ode {name} {integ}", flush=True)
    if name == "n_body_3d":
        eq, y0 = _n_body_3d()
        t1, d_t = 0.5, 1e-3
    else:
        make, y0, t1, d_t = ODE_CASES[name]
        eq = make()
    cp, ivp = _ode_ivp(eq, y0, (0.0, t1))
    got = FDMOperator(INTEGRATORS[integ](), DIFF, d_t).solve(ivp).discrete_y()
    _, want = ora.fdm_solve(cp, np.asarray(y0, dtype=float), (0.0, t1), d_t, integ)
    assert got.shape == want.shape
    assert rel_err(got, want) <= 1e-11, rel_err(got, want)


def test_ode_fp32_snapshots_and_offset_start():
    print("case: ode fp32 + snapshot_every + t0", flush=True)
    eq = pml.VanDerPolEquation(1.0)
    cp, ivp = _ode_ivp(eq, [1.0, 0.0], (0.5, 2.5))
    f64 = FDMOperator(RK4(), DIFF, 1e-3, snapshot_every=10).solve(ivp)
    f32 = FDMOperator(RK4(), DIFF, 1e-3, precision="float32", snapshot_every=10).solve(ivp)
    assert f64.discrete_y().shape == (200, 2) and f64.d_t == pytest.approx(1e-2)
    np.testing.assert_allclose(f64.t_coordinates, np.linspace(0.5, 2.5, 201)[1:], rtol=1e-12)
    _, want = ora.fdm_solve(cp, np.array([1.0, 0.0]), (0.5, 2.5), 1e-3, "rk4")
    assert rel_err(f64.discrete_y(), want[9::10]) <= 1e-11
    # FP32 rounding over 2000 RK4 steps of an O(1) oscillator
    assert rel_err(f32.discrete_y(), f64.discrete_y()) <= 1e-4


class _TXEquation(DifferentialEquation):
    """The reference's test equation with a t- and x-dependent right-hand
    side (tests/operators/fdm/test_fdm_operator.py:322-335)."""

    def __init__(self):
        super().__init__(2, 1)

    @property
    def symbolic_equation_system(self):
        return SymbolicEquationSystem(
            [self.symbols.t / 100.0 * (self.symbols.x[0] + self.symbols.x[1]) ** 2])


def test_pde_with_t_and_x_dependent_rhs_as_the_reference_test():
    print("case: pde t,x-dependent rhs", flush=True)
    eq = _TXEquation()
    mesh = pml.Mesh([(-5.0, 5.0), (0.0, 3.0)], [2.0, 1.0])
    bcs = [(pml.NeumannBoundaryCondition(lambda x, t: np.zeros((len(x), 1))),
            pml.NeumannBoundaryCondition(lambda x, t: np.zeros((len(x), 1))))] * 2
    cp = pml.ConstrainedProblem(eq, mesh, bcs)
    ic = pml.ContinuousInitialCondition(cp, lambda x: np.zeros((len(x), 1)))
    ivp = pml.InitialValueProblem(cp, (0.0, 5.0), ic)
    sol = FDMOperator(RK4(), DIFF, 0.25).solve(ivp)
    y = sol.discrete_y()
    assert sol.vertex_oriented and sol.d_t == 0.25 and y.shape == (20, 6, 4, 1)
    _, want = ora.fdm_solve(cp, ic.discrete_y_0(True), (0.0, 5.0), 0.25, "rk4")
    assert rel_err(y, want) <= 1e-12
    # the closed form: y(t) = t^2 / 200 (x0 + x1)^2, exact for RK4 (a
    # quadratic in t)
    x0, x1 = mesh.vertex_coordinate_grids
    np.testing.assert_allclose(y[-1, ..., 0], 25.0 / 200.0 * (x0 + x1) ** 2, rtol=1e-12)


class _AllSymbols2D(DifferentialEquation):
    """Every derivative symbol of a 2-D system with three fields except the
    2-D curl: the reference's 2-D ``y-curl`` symbol carries no component
    indices (differential_equation.py:46-52), which its symbol mapper cannot
    map (symbol_mapper.py:198-210); the 3-D test covers the curl."""

    def __init__(self):
        super().__init__(2, 3, [(1, 2)])

    @property
    def symbolic_equation_system(self):
        s = self.symbols
        y, g, h = s.y, s.y_gradient, s.y_hessian
        return SymbolicEquationSystem([
            0.05 * s.y_laplacian[0] - y[1] * g[0, 0] - y[2] * g[0, 1]
            + 0.01 * h[0, 0, 1] + 0.02 * sp.sin(s.t) * s.x[0],
            0.03 * s.y_vector_laplacian[1, 2, 0] - 0.1 * s.y_divergence[1, 2] * y[1]
            + 0.01 * h[1, 1, 1],
            0.03 * s.y_vector_laplacian[1, 2, 1] + 0.05 * g[1, 0] * y[0]
            - 0.2 * sp.tanh(y[2]) + 0.01 * sp.sqrt(1.0 + y[0] ** 2) * h[2, 1, 0],
        ])


def test_pde_every_derivative_symbol_2d():
    print("case: pde all symbols 2d", flush=True)
    eq = _AllSymbols2D()
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.4)], [0.1, 0.1])
    bcs = [
        (pml.DirichletBoundaryCondition(lambda x, t: np.zeros((len(x), 3)), is_static=True),
         pml.NeumannBoundaryCondition(lambda x, t: np.full((len(x), 3), 0.1), is_static=True)),
        (pml.NeumannBoundaryCondition(lambda x, t: np.zeros((len(x), 3)), is_static=True),
         pml.DirichletBoundaryCondition(lambda x, t: np.full((len(x), 3), 0.2),
                                        is_static=True)),
    ]
    cp = pml.ConstrainedProblem(eq, mesh, bcs)
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.5, 0.7]), 0.05 * np.eye(2))] * 3,
                                      [1.0, 0.3, -0.2])
    ivp = pml.InitialValueProblem(cp, (0.0, 0.02), ic)
    for integ in ("fe", "rk4"):
        got = FDMOperator(INTEGRATORS[integ](), DIFF, 0.002).solve(ivp).discrete_y()
        _, want = ora.fdm_solve(cp, ic.discrete_y_0(True), (0.0, 0.02), 0.002, integ)
        assert rel_err(got, want) <= 1e-12, (integ, rel_err(got, want))
        f32 = FDMOperator(INTEGRATORS[integ](), DIFF, 0.002, precision="float32").solve(ivp)
        assert rel_err(f32.discrete_y(), want) <= 1e-5


class _Curl3D(DifferentialEquation):
    def __init__(self):
        super().__init__(3, 3, [(0, 1, 2)])

    @property
    def symbolic_equation_system(self):
        s = self.symbols
        return SymbolicEquationSystem([
            0.02 * s.y_laplacian[0] + 0.1 * s.y_curl[0, 1, 2, 0],
            0.02 * s.y_laplacian[1] + 0.1 * s.y_curl[0, 1, 2, 1] - 0.05 * s.y_divergence[0, 1, 2],
            0.02 * s.y_vector_laplacian[0, 1, 2, 2] + 0.1 * s.y_curl[0, 1, 2, 2] * s.x[2],
        ])


def test_pde_3d_user_system():
    print("case: pde 3d", flush=True)
    eq = _Curl3D()
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 0.8), (0.0, 0.6)], [0.1, 0.1, 0.1])
    bc = pml.DirichletBoundaryCondition(lambda x, t: np.zeros((len(x), 3)), is_static=True)
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc)] * 3)
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.5, 0.4, 0.3]), 0.03 * np.eye(3))] * 3,
                                      [1.0, 0.5, 0.25])
    ivp = pml.InitialValueProblem(cp, (0.0, 0.01), ic)
    got = FDMOperator(RK4(), DIFF, 0.001).solve(ivp).discrete_y()
    _, want = ora.fdm_solve(cp, ic.discrete_y_0(True), (0.0, 0.01), 0.001, "rk4")
    assert rel_err(got, want) <= 1e-12


def _fused_cases():
    from tests.helpers import cahn_hilliard, navier_stokes_cavity, shallow_water, zero_bc

    mesh = pml.Mesh([(0.0, 2.0), (0.0, 1.5)], [0.1, 0.1])
    out = {}
    for name, eq, n_f in (("diffusion", pml.DiffusionEquation(2, 0.3), 1),
                          ("wave", pml.WaveEquation(2, 1.2), 2),
                          ("convection_diffusion",
                           pml.ConvectionDiffusionEquation(2, [0.3, -0.2], 0.05), 1),
                          ("burgers", pml.BurgersEquation(2, 50.0), 2)):
        bcs = [(pml.DirichletBoundaryCondition(lambda x, t, n=n_f: np.full((len(x), n), 0.5),
                                               is_static=True),
                pml.NeumannBoundaryCondition(lambda x, t, n=n_f: np.full((len(x), n), -0.2),
                                             is_static=True)),
               (zero_bc(n_f, "neumann"), zero_bc(n_f, "neumann"))]
        cp = pml.ConstrainedProblem(eq, mesh, bcs)
        ic = pml.GaussianInitialCondition(cp, [(np.array([1.0, 0.7]), 0.1 * np.eye(2))] * n_f)
        out[name] = (cp, ic, 0.0025)
    cp, ic = cahn_hilliard(20, seed=2)
    out["cahn_hilliard"] = (cp, ic, 1e-4)
    cp, ic = shallow_water(33, seed=1)
    out["shallow_water"] = (cp, ic, 1e-3)
    cp, y0 = navier_stokes_cavity(16, re=100.0)
    out["navier_stokes"] = (cp, pml.DiscreteInitialCondition(cp, y0, True), 1e-3)
    return out


@pytest.mark.parametrize("name", ["burgers", "cahn_hilliard", "convection_diffusion",
                                  "diffusion", "navier_stokes", "shallow_water", "wave"])
def test_symbolic_path_matches_fused_kernels(name):
    """The generic path (compiled symbolic right-hand side + device
    operators) reproduces the hand-fused kernels, LHS.Y (Cahn-Hilliard) and
    LHS.Y_LAPLACIAN (Navier-Stokes, Jacobi tolerance 1e-12) included."""
    print(f"case: symbolic vs fused {name}", flush=True)
    cp, ic, d_t = _fused_cases()[name]
    diff = ThreePointCentralDifferenceMethod(tol=1e-12)
    ivp = pml.InitialValueProblem(cp, (0.0, 10 * d_t), ic)
    for integ in ("fe", "rk4"):
        fused = FDMOperator(INTEGRATORS[integ](), diff, d_t).solve(ivp).discrete_y()
        sym = FDMOperator(INTEGRATORS[integ](), diff, d_t, symbolic_rhs=True).solve(ivp)
        tol = 1e-10 if name in ("navier_stokes", "shallow_water") else 1e-12
        assert rel_err(sym.discrete_y(), fused) <= tol, (integ, rel_err(sym.discrete_y(), fused))


def _ode_parareal_oracle(cp, y0, t_interval, n, f_dt, g_dt, f_int, g_int, tol):
    def f_solve(y, a, b):
        return ora.fdm_solve(cp, y, (a, b), f_dt, f_int)[1]

    def g_solve(y, a, b):
        return ora.fdm_solve(cp, y, (a, b), g_dt, g_int)[1]

    return ora.parareal_solve(np.asarray(y0, dtype=float), t_interval, n, f_solve, g_solve,
                              f_dt, g_dt, lambda y: y, tol)


@pytest.mark.parametrize("case", ["van_der_pol", "lorenz", "n_body_3d"])
def test_ode_parareal_matches_oracle_emulation(case):
    """The reference's ODE Parareal examples (van_der_pol_parareal.py,
    lorenz_parareal.py, n_body_3d_parareal.py), scaled down: slices batched
    into one ODE kernel launch per fine phase, iteration counts equal to the
    oracle's emulation of the SPMD algorithm, trajectories within 1e-9."""
    print(f"case: ode parareal {case}", flush=True)
    if case == "van_der_pol":
        eq, y0, t1, n, fd, gd, fi, gi, tol = (pml.VanDerPolEquation(), [1.0, 0.0], 4.0, 8,
                                             1e-4, 1e-3, "fe", "fe", 1e-3)
    elif case == "lorenz":
        eq, y0, t1, n, fd, gd, fi, gi, tol = (pml.LorenzEquation(), [1.0, 1.0, 1.0], 2.0, 8,
                                             1e-4, 1e-2, "rk4", "rk4", 1e-4)
    else:
        eq, y0 = _n_body_3d()
        t1, n, fd, gd, fi, gi, tol = 0.5, 5, 1e-3, 1e-2, "rk4", "rk4", 0.5
    cp, ivp = _ode_ivp(eq, y0, (0.0, t1))
    f = FDMOperator(INTEGRATORS[fi](), DIFF, fd)
    g = FDMOperator(INTEGRATORS[gi](), DIFF, gd)
    p = PararealOperator(f, g, tol, n_slices=n)
    sol = p.solve(ivp)
    _, traj, iters, _ = _ode_parareal_oracle(cp, y0, (0.0, t1), n, fd, gd, fi, gi, tol)
    assert p.last_iterations == iters
    y = sol.discrete_y()
    assert y.shape == traj.shape
    assert rel_err(y, traj) <= 1e-9, rel_err(y, traj)
    if case == "van_der_pol":
        serial = f.solve(ivp).discrete_y()
        assert rel_err(y, serial) <= 1e-2  # converged to the fine solution's tolerance
