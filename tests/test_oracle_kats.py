"""Pins the CPU oracle against the reference's own known answers and
against what the mathematics fixes (exactness on polynomials, discrete
eigenvalues, integrator amplification factors)."""

import numpy as np
import pytest

from oracle import fdm_numpy as ora
from tests.helpers import dbc_from_spec


def _arr(v):
    return np.array(v, dtype=float)


@pytest.mark.parametrize(
    "name", ["gradient_2d", "gradient_1d_constrained", "gradient_2d_constrained"]
)
def test_oracle_gradient_kats(kats, name):
    k = kats[name]
    y = _arr(k["y"])
    dbc = dbc_from_spec(k.get("dbc"), len(k["d_x"]), y.shape[-1])
    out = ora.gradient(y, k["d_x"], k["x_axis"], dbc)
    np.testing.assert_allclose(out, _arr(k["expected"]), rtol=0, atol=1e-12)


@pytest.mark.parametrize(
    "name",
    ["hessian_2d", "hessian_1d_constrained", "hessian_2d_constrained", "hessian_mixed"],
)
def test_oracle_hessian_kats(kats, name):
    k = kats[name]
    y = _arr(k["y"])
    dbc = dbc_from_spec(k.get("dbc"), len(k["d_x"]), y.shape[-1])
    out = ora.hessian(y, k["d_x"], k["x_axis1"], k["x_axis2"], dbc)
    np.testing.assert_allclose(out, _arr(k["expected"]), rtol=0, atol=1e-12)


@pytest.mark.parametrize("name", ["divergence_2d", "divergence_3d"])
def test_oracle_divergence_kats(kats, name):
    k = kats[name]
    out = ora.divergence(_arr(k["y"]), k["d_x"])
    np.testing.assert_allclose(out, _arr(k["expected"]), rtol=0, atol=1e-12)


def test_oracle_curl_kats(kats):
    k = kats["curl_2d"]
    np.testing.assert_allclose(
        ora.curl(_arr(k["y"]), k["d_x"], 0), _arr(k["expected"]), atol=1e-12
    )
    k = kats["curl_3d"]
    exp = _arr(k["expected"])
    for i in range(3):
        np.testing.assert_allclose(
            ora.curl(_arr(k["y"]), k["d_x"], i), exp[..., i : i + 1], atol=1e-12
        )


def test_oracle_laplacian_kat(kats):
    k = kats["laplacian_2d"]
    np.testing.assert_allclose(
        ora.laplacian(_arr(k["y"]), k["d_x"]), _arr(k["expected"]), atol=1e-12
    )
    # vector Laplacian component == scalar Laplacian of that component
    y = _arr(k["y"])
    for i in range(2):
        np.testing.assert_array_equal(
            ora.vector_laplacian(y, k["d_x"], i), ora.laplacian(y[..., i : i + 1], k["d_x"])
        )


def test_oracle_hessian_symmetry_and_trace(kats):
    k = kats["hessian_symmetry_3d_y"]
    y, dx = _arr(k["y"]), k["d_x"]
    for a in range(3):
        for b in range(a + 1, 3):
            np.testing.assert_allclose(ora.hessian(y, dx, a, b), ora.hessian(y, dx, b, a))
    trace = sum(ora.hessian(y, dx, a, a) for a in range(3))
    np.testing.assert_allclose(ora.laplacian(y, dx), trace)


def test_oracle_derivatives_exact_on_polynomials():
    # interior first derivative exact on linear fields, second on quadratics
    x0 = np.linspace(0, 2, 9)
    x1 = np.linspace(-1, 1, 7)
    X0, X1 = np.meshgrid(x0, x1, indexing="ij")
    dx = [x0[1] - x0[0], x1[1] - x1[0]]
    lin = (3 * X0 - 2 * X1 + 1)[..., None]
    g0 = ora.gradient(lin, dx, 0)
    g1 = ora.gradient(lin, dx, 1)
    np.testing.assert_allclose(g0[1:-1], 3.0, atol=1e-12)
    np.testing.assert_allclose(g1[:, 1:-1], -2.0, atol=1e-12)
    quad = (X0**2 + 2 * X1**2 - X0 * X1)[..., None]
    lap = ora.laplacian(quad, dx)
    np.testing.assert_allclose(lap[1:-1, 1:-1], 6.0, atol=1e-10)
    mixed = ora.hessian(quad, dx, 0, 1)
    np.testing.assert_allclose(mixed[1:-1, 1:-1], -1.0, atol=1e-10)


def test_oracle_laplacian_eigenmodes():
    """Zero halos = Dirichlet-zero just outside; the sine modes of the
    n interior-like vertices are eigenvectors with eigenvalue
    -(4/dx^2) sin^2(k pi / (2 (n + 1))) per axis."""
    n0, n1 = 13, 9
    dx = [0.3, 0.2]
    i0 = np.arange(1, n0 + 1)[:, None]
    i1 = np.arange(1, n1 + 1)[None, :]
    k0, k1 = 3, 2
    u = (np.sin(k0 * np.pi * i0 / (n0 + 1)) * np.sin(k1 * np.pi * i1 / (n1 + 1)))[..., None]
    lam = -(4 / dx[0] ** 2) * np.sin(k0 * np.pi / (2 * (n0 + 1))) ** 2 - (
        4 / dx[1] ** 2
    ) * np.sin(k1 * np.pi / (2 * (n1 + 1))) ** 2
    np.testing.assert_allclose(ora.laplacian(u, dx), lam * u, atol=1e-10)


def test_oracle_neumann_ghost_cosine_modes():
    """Zero-Neumann ghosts reflect (y[-1] = y[1]); cosine modes of the
    vertex grid are eigenvectors with eigenvalue -(4/dx^2) sin^2(k pi/(2(n-1)))."""
    import pararealml_amd as pml
    from tests.helpers import zero_bc

    n = 11
    mesh = pml.Mesh([(0.0, 1.0)], [0.1])
    bc = zero_bc(1, "neumann")
    cp = pml.ConstrainedProblem(pml.DiffusionEquation(1), mesh, [(bc, bc)])
    dbc = cp.static_boundary_vertex_constraints[1]
    k = 2
    i = np.arange(n)
    u = np.cos(k * np.pi * i / (n - 1))[:, None]
    lam = -(4 / 0.1**2) * np.sin(k * np.pi / (2 * (n - 1))) ** 2
    np.testing.assert_allclose(ora.laplacian(u, mesh.d_x, dbc), lam * u, atol=1e-9)


def test_oracle_integrator_kats(kats):
    k = kats["integrators"]
    fe = k["forward_euler"]
    out = ora.integrate_fe(_arr(fe["y0"]), fe["t0"], fe["d_t"], lambda t, y: 5 * y + t**2, lambda _: None)
    np.testing.assert_allclose(out, fe["expected"])
    mp = k["explicit_midpoint"]
    out = ora.integrate_midpoint(
        _arr(mp["y0"]), mp["t0"], mp["d_t"], lambda t, y: 2 * y - 4 * t, lambda _: None
    )
    np.testing.assert_allclose(out, mp["expected"])
    rk = k["rk4"]
    out = ora.integrate_rk4(_arr(rk["y0"]), rk["t0"], rk["d_t"], lambda t, y: 2 * y + 1, lambda _: None)
    np.testing.assert_allclose(out, rk["expected"])


@pytest.mark.parametrize(
    "name,poly",
    [
        ("fe", lambda z: 1 + z),
        ("midpoint", lambda z: 1 + z + z**2 / 2),
        ("rk4", lambda z: 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24),
    ],
)
def test_oracle_integrator_amplification(name, poly):
    lam, dt = -1.7, 0.3
    y = np.array([2.0])
    out = ora.INTEGRATORS[name](y, 0.0, dt, lambda t, v: lam * v, lambda _: None)
    np.testing.assert_allclose(out, poly(lam * dt) * y, rtol=1e-14)


def test_oracle_discretize_time_domain(kats):
    k = kats["discretize_time_domain"]
    for (interval, dt) in k["cases"]:
        np.testing.assert_allclose(ora.discretize_time_domain(interval, dt), k["expected"])


def test_oracle_anti_laplacian_round_trip():
    from pararealml_amd.constraint import Constraint

    y = np.random.random((12, 10, 1))
    dx = [0.05, 0.025]
    value = np.full((12, 10, 1), np.nan)
    value[0], value[-1] = 1.0, 2.0
    value[:, 0], value[:, -1] = 3.0, 4.0
    mask = ~np.isnan(value)
    c = Constraint(value[mask], mask)
    c.apply(y)
    lap = ora.laplacian(y, dx)
    sol = ora.anti_laplacian(lap, dx, [c], tol=1e-12)
    np.testing.assert_allclose(sol, y, atol=1e-8)
