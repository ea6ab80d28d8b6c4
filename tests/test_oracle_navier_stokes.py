"""The oracle's LHS.Y_LAPLACIAN step (fdm_operator.py:141-161) on
Navier-Stokes, pinned by relations the step must satisfy rather than by its
own right-hand side: after a step psi solves lap(psi) = -w of the pre-step
state, the next step's u and v are the derivatives of that psi, and with
u = v = 0 the vorticity only diffuses."""

import numpy as np

from oracle import fdm_numpy as ora
from tests.helpers import navier_stokes_cavity


def test_navier_stokes_step_relations():
    cp, y0 = navier_stokes_cavity(8, re=50.0)
    d_t = 1e-3
    _, y = ora.fdm_solve(cp, y0, (0.0, 2 * d_t), d_t, "fe", poisson_tol=1e-13,
                         poisson_rng=np.random.RandomState(3))
    d_x = cp.mesh.d_x
    dbc = cp.static_boundary_vertex_constraints[1]
    inner = (slice(1, -1), slice(1, -1))
    # lap(psi_1) = -w_0 and lap(psi_2) = -w_1 at the interior vertices
    for k, w_pre in ((0, y0), (1, y[0])):
        lap = ora.laplacian(y[k][..., 1:2], d_x, dbc[:, 1:2])[..., 0]
        np.testing.assert_allclose(lap[inner], -w_pre[..., 0][inner], rtol=0, atol=1e-9)
    # u_2 = d psi_1 / dx1, v_2 = -d psi_1 / dx0 (from the pre-step state)
    u2 = ora.derivative(y[0][..., 1:2], d_x[1], 1, dbc[1, 1])[..., 0]
    v2 = -ora.derivative(y[0][..., 1:2], d_x[0], 0, dbc[0, 1])[..., 0]
    np.testing.assert_allclose(y[1][..., 2][inner], u2[inner], rtol=0, atol=1e-12)
    np.testing.assert_allclose(y[1][..., 3][inner], v2[inner], rtol=0, atol=1e-12)
    # step 1: u_0 = v_0 = 0 inside, so w_1 = w_0 + dt lap(w_0) / re there
    lap_w0 = ora.laplacian(y0[..., 0:1], d_x, dbc[:, 0:1])[..., 0]
    np.testing.assert_allclose(y[0][..., 0][inner], (y0[..., 0] + d_t * lap_w0 / 50.0)[inner],
                               rtol=1e-13, atol=1e-15)
    # the lid and the walls hold their Dirichlet values
    assert np.all(y[1][:, -1, 2] == 1.0) and np.all(y[1][0, :, 0] == 0.0)


def test_poisson_result_does_not_depend_on_the_start_at_tight_tolerance():
    cp, y0 = navier_stokes_cavity(8)
    a = ora.fdm_solve(cp, y0, (0.0, 1e-3), 1e-3, "fe", poisson_tol=1e-13,
                      poisson_rng=np.random.RandomState(1))[1]
    b = ora.fdm_solve(cp, y0, (0.0, 1e-3), 1e-3, "fe", poisson_tol=1e-13,
                      poisson_rng=np.random.RandomState(2))[1]
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-10)
