"""Pins of the oracle's Solution post-processing (oracle/solution_numpy.py):
interpolation against scipy.interpolate.interpn -- the routine the
reference's Solution.y calls (pararealml/solution.py:132-139) -- and against
what the mathematics fixes; time matching against grids whose matches are
known (solution.py:182-258)."""

import numpy as np
import pytest
from scipy.interpolate import interpn

from oracle import solution_numpy as osol


def _axes(shape, rng):
    return [np.linspace(rng.uniform(-1, 0), rng.uniform(1, 2), n) for n in shape]


def _points(axes, n, rng, pad=0.3):
    lo = np.array([a[0] for a in axes])
    hi = np.array([a[-1] for a in axes])
    span = hi - lo
    return lo - pad * span + (1 + 2 * pad) * span * rng.random((n, len(axes)))


def _scipy(axes, values, xi, method):
    out = interpn(tuple(axes), np.moveaxis(values, 0, -2), xi, method=method,
                  bounds_error=False, fill_value=None)
    out = np.moveaxis(out, -2, 0)
    return out.reshape((values.shape[0],) + xi.shape[:-1] + (values.shape[-1],))


@pytest.mark.parametrize("shape", [(17,), (9, 13), (5, 6, 7), (1, 8), (6, 1, 4)])
@pytest.mark.parametrize("method", ["linear", "nearest"])
def test_interp_equals_scipy_bitwise(shape, method):
    rng = np.random.default_rng(len(shape) * 7 + shape[0])
    axes = _axes(shape, rng)
    values = rng.standard_normal((4,) + shape + (3,))
    xi = _points(axes, 500, rng)
    # exact vertices, the upper boundary and a NaN coordinate
    xi[:len(axes[0])] = np.stack(
        [a[np.arange(len(axes[0])) % len(a)] for a in axes], axis=-1)
    xi[-2] = [a[-1] for a in axes]
    xi[-1, 0] = np.nan
    xi = xi.reshape(25, 20, len(shape))
    ours = osol.interp_regular(axes, values, xi, method)
    ref = _scipy(axes, values, xi, method)
    assert ours.shape == ref.shape == (4, 25, 20, 3)
    np.testing.assert_array_equal(ours, ref)


def test_interp_exact_on_multilinear_fields_inside_and_outside():
    rng = np.random.default_rng(3)
    axes = _axes((11, 9), rng)
    g0, g1 = np.meshgrid(*axes, indexing="ij")
    c = rng.standard_normal(4)

    def f(x0, x1):
        return c[0] + c[1] * x0 + c[2] * x1 + c[3] * x0 * x1

    values = f(g0, g1)[None, ..., None]
    xi = _points(axes, 2000, rng, pad=0.5)
    got = osol.interp_regular(axes, values, xi)[0, :, 0]
    np.testing.assert_allclose(got, f(xi[:, 0], xi[:, 1]), rtol=0, atol=1e-12)


def test_nearest_returns_vertex_values_at_jittered_vertices():
    rng = np.random.default_rng(5)
    axes = _axes((12, 10), rng)
    values = rng.standard_normal((2, 12, 10, 1))
    i0 = rng.integers(0, 12, 300)
    i1 = rng.integers(0, 10, 300)
    h0 = axes[0][1] - axes[0][0]
    h1 = axes[1][1] - axes[1][0]
    xi = np.stack([axes[0][i0] + h0 * rng.uniform(-0.4, 0.4, 300),
                   axes[1][i1] + h1 * rng.uniform(-0.4, 0.4, 300)], axis=-1)
    got = osol.interp_regular(axes, values, xi, "nearest")
    np.testing.assert_array_equal(got, values[:, i0, i1, :])


def _stack(n, seed):
    return np.random.default_rng(seed).standard_normal((n, 4, 3))


def test_diff_every_other_point_of_a_finer_grid():
    t_a = np.linspace(0.1, 1.0, 10)  # d_t 0.1
    t_b = np.linspace(0.05, 1.0, 20)  # d_t 0.05
    ya, yb = _stack(10, 1), _stack(20, 2)
    times, d = osol.diff(ya, [yb], [t_a, t_b], [0.1, 0.05])
    np.testing.assert_allclose(times, t_a)
    np.testing.assert_array_equal(d[0], yb[1::2] - ya)
    # the finer solution as self: the same matches, opposite sign
    times2, d2 = osol.diff(yb, [ya], [t_b, t_a], [0.05, 0.1])
    np.testing.assert_allclose(times2, t_a)
    np.testing.assert_array_equal(d2[0], ya - yb[1::2])


def test_diff_tolerance_and_no_match():
    t_a = np.linspace(0.1, 1.0, 10)
    ya, yb = _stack(10, 3), _stack(10, 4)
    # shifted by less than atol: every point matches
    times, d = osol.diff(ya, [yb], [t_a, t_a + 5e-9], [0.1, 0.1], atol=1e-8)
    assert len(times) == 10
    np.testing.assert_array_equal(d[0], yb - ya)
    # shifted by more than atol: nothing matches, the reference's empty array
    times, d = osol.diff(ya, [yb], [t_a, t_a + 0.05], [0.1, 0.1], atol=1e-8)
    assert times.shape == (0,) and d[0].shape == (0,)
    # a longer other grid: only the common stretch matches
    t_c = np.linspace(0.1, 2.0, 20)
    yc = _stack(20, 5)
    times, d = osol.diff(ya, [yc], [t_a, t_c], [0.1, 0.1])
    np.testing.assert_allclose(times, t_a)
    np.testing.assert_array_equal(d[0], yc[:10] - ya)


def test_oracle_diff_equals_host_solution_diff():
    """The package's host Solution.diff (which keeps the reference's host
    semantics for host-held solutions) agrees with the oracle."""
    from pararealml_amd import (
        ConstrainedProblem, DiffusionEquation, GaussianInitialCondition,
        InitialValueProblem, Mesh, NeumannBoundaryCondition,
    )
    from pararealml_amd.solution import Solution

    mesh = Mesh([(0.0, 1.0)], [0.25])
    bcs = [(NeumannBoundaryCondition(lambda x, t: np.zeros((len(x), 1)), is_static=True),) * 2]
    cp = ConstrainedProblem(DiffusionEquation(1), mesh, bcs)
    ivp = InitialValueProblem(cp, (0.0, 1.0), GaussianInitialCondition(
        cp, [(np.array([0.5]), np.array([[0.1]]))]))
    t_a = np.linspace(0.1, 1.0, 10)
    t_b = np.linspace(0.05, 1.0, 20)
    ya = np.random.default_rng(6).standard_normal((10, 5, 1))
    yb = np.random.default_rng(7).standard_normal((20, 5, 1))
    sa = Solution(ivp, t_a, ya, vertex_oriented=True, d_t=0.1)
    sb = Solution(ivp, t_b, yb, vertex_oriented=True, d_t=0.05)
    host = sa.diff([sb])
    times, d = osol.diff(ya, [yb], [t_a, t_b], [0.1, 0.05])
    np.testing.assert_array_equal(host.matching_time_points, times)
    np.testing.assert_array_equal(host.differences[0], d[0])
