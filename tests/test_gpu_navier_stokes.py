"""Navier-Stokes (vorticity / stream function) on the device: the fused step
for w, u, v plus the per-step Jacobi solve of lap(psi) = -w, against the
oracle's restatement of fdm_operator.py:141-161."""

import numpy as np
import pytest

from oracle import fdm_numpy as ora
from tests.helpers import navier_stokes_cavity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("integ", ["fe", "midpoint", "rk4"])
def test_navier_stokes_matches_oracle(integ):
    import pararealml_amd as pml
    from pararealml_amd.operators.fdm import (
        RK4,
        ExplicitMidpointMethod,
        FDMOperator,
        ForwardEulerMethod,
        ThreePointCentralDifferenceMethod,
    )

    print(f"case: navier-stokes {integ}", flush=True)
    cp, y0 = navier_stokes_cavity(16, re=100.0)
    d_t, steps = 1e-3, 6
    ic = pml.DiscreteInitialCondition(cp, y0, True)
    ivp = pml.InitialValueProblem(cp, (0.0, steps * d_t), ic)
    integrator = {"fe": ForwardEulerMethod(), "midpoint": ExplicitMidpointMethod(),
                  "rk4": RK4()}[integ]
    # A tight Jacobi tolerance makes psi independent of the iteration's
    # starting value (random in the reference, the previous psi here).
    op = FDMOperator(integrator, ThreePointCentralDifferenceMethod(tol=1e-12), d_t)
    y = op.solve(ivp).discrete_y()
    # the initial condition carries the Dirichlet values (initial_condition.py:86-89)
    _, y_ref = ora.fdm_solve(cp, ic.discrete_y_0(True), (0.0, steps * d_t), d_t, integ,
                             poisson_tol=1e-12)
    assert y.shape == y_ref.shape == (steps, 17, 17, 4)
    scale = np.abs(y_ref).max(axis=(0, 1, 2))
    err = np.abs(y - y_ref).max(axis=(0, 1, 2)) / scale
    assert np.all(err <= 1e-8), err
    # psi solves the Poisson equation of the previous step's vorticity
    lap = ora.laplacian(y[-1][..., 1:2], cp.mesh.d_x, cp.static_boundary_vertex_constraints[1][:, 1:2])
    inner = (slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(lap[inner][..., 0], -y[-2][..., 0][inner], rtol=0,
                               atol=1e-6 * np.abs(y[-2][..., 0]).max())


def test_navier_stokes_rejects_unsupported_uses():
    import torch

    import pararealml_amd as pml
    from pararealml_amd import _device as dev
    from pararealml_amd.operators.fdm import (
        FDMOperator,
        ForwardEulerMethod,
        ThreePointCentralDifferenceMethod,
    )

    print("case: navier-stokes guards", flush=True)
    cp, y0 = navier_stokes_cavity(8)
    ivp = pml.InitialValueProblem(cp, (0.0, 2e-3), pml.DiscreteInitialCondition(cp, y0, True))
    op = FDMOperator(ForwardEulerMethod(), ThreePointCentralDifferenceMethod(), 1e-3,
                     precision="float32c")
    with pytest.raises(NotImplementedError):
        op.solve(ivp)
    problem = dev.device_problem(cp)
    with pytest.raises(NotImplementedError):  # batched stepping has no Poisson phase
        dev.FDMBatch([problem], [0], "float64", dev.integrator_id(ForwardEulerMethod()),
                     torch.cuda.current_stream())
