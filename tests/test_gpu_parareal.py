"""Parity of the one-GPU Parareal driver with the oracle's serial emulation
of the reference's SPMD algorithm, and the textbook invariants."""

import numpy as np
import pytest

import pararealml_amd as pml
from oracle import fdm_numpy as ora
from pararealml_amd.operators.fdm import (
    RK4,
    FDMOperator,
    ForwardEulerMethod,
    ThreePointCentralDifferenceMethod,
)
from pararealml_amd.operators.parareal import CoarseGridFDMOperator, PararealOperator
from tests.helpers import cahn_hilliard, shallow_water, zero_bc

pytestmark = pytest.mark.gpu

DIFF = ThreePointCentralDifferenceMethod()


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def diffusion_problem(n=33, d=0.5):
    eq = pml.DiffusionEquation(2, d)
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.0)], [1.0 / (n - 1)] * 2)
    bc = zero_bc(1)
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc), (bc, bc)])
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.4, 0.55]), 0.01 * np.eye(2))])
    return cp, ic


def oracle_parareal(cp, ic, t_interval, n_slices, f_dt, g_dt, f_int, g_int, tol,
                    coarse_grid=False, max_iterations=10**9):
    y0 = ic.discrete_y_0(True)
    y_cons = cp.static_y_vertex_constraints

    def constrain(y):
        return ora.apply_constraints_last_axis(y_cons, y)

    def f_solve(y, a, b):
        return ora.fdm_solve(cp, y, (a, b), f_dt, f_int)[1]

    if coarse_grid:
        mesh = cp.mesh
        cmesh = pml.Mesh(mesh.x_intervals, [2 * d for d in mesh.d_x])
        ccp = pml.ConstrainedProblem(cp.differential_equation, cmesh, cp.boundary_conditions)

        def g_solve(y, a, b):
            # injection, the coarse solve, cubic prolongation; Dirichlet
            # values on both grids (CoarseGridFDMOperator's transfers)
            yc = ora.apply_constraints_last_axis(ccp.static_y_vertex_constraints,
                                                 ora.restrict_injection(y).copy())
            traj = ora.fdm_solve(ccp, yc, (a, b), g_dt, g_int)[1]
            return np.stack([constrain(ora.prolong_cubic(s)) for s in traj])
    else:
        def g_solve(y, a, b):
            return ora.fdm_solve(cp, y, (a, b), g_dt, g_int)[1]

    return ora.parareal_solve(y0, t_interval, n_slices, f_solve, g_solve, f_dt, g_dt,
                              constrain, tol, max_iterations)


def test_parareal_matches_oracle_emulation_fp64():
    print("case: parareal 4 slices fp64", flush=True)
    cp, ic = diffusion_problem()
    t = (0.0, 0.04)
    f = FDMOperator(RK4(), DIFF, 1e-4)
    g = FDMOperator(ForwardEulerMethod(), DIFF, 5e-4)
    p = PararealOperator(f, g, 1e-6, n_slices=4)
    sol = p.solve(pml.InitialValueProblem(cp, t, ic))
    t_ref, y_ref, it_ref, _ = oracle_parareal(cp, ic, t, 4, 1e-4, 5e-4, "rk4", "fe", 1e-6)
    assert p.last_iterations == it_ref
    np.testing.assert_allclose(sol.t_coordinates, t_ref)
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12


def test_parareal_n_iterations_equals_serial_fine():
    """With a tolerance that is never met, N iterations reproduce the serial
    fine solution (Parareal's finite-termination property)."""
    print("case: parareal finite termination", flush=True)
    cp, ic = diffusion_problem(25)
    t = (0.0, 0.02)
    f = FDMOperator(RK4(), DIFF, 1e-4)
    g = FDMOperator(ForwardEulerMethod(), DIFF, 1e-3)
    p = PararealOperator(f, g, 0.0, n_slices=5)
    ivp = pml.InitialValueProblem(cp, t, ic)
    sol = p.solve(ivp)
    assert p.last_iterations == 5
    fine = f.solve(ivp).discrete_y()
    assert rel_err(sol.discrete_y(), fine) <= 1e-12


def test_parareal_coarse_equal_to_fine_converges_in_one_iteration():
    print("case: parareal g == f", flush=True)
    cp, ic = diffusion_problem(25)
    t = (0.0, 0.02)
    f = FDMOperator(RK4(), DIFF, 1e-4)
    p = PararealOperator(f, f, 1e-12, n_slices=4)
    ivp = pml.InitialValueProblem(cp, t, ic)
    sol = p.solve(ivp)
    assert p.last_iterations == 1
    assert rel_err(sol.discrete_y(), f.solve(ivp).discrete_y()) <= 1e-12


def test_parareal_single_slice_and_serial_mode():
    cp, ic = diffusion_problem(17)
    ivp = pml.InitialValueProblem(cp, (0.0, 0.01), ic)
    f = FDMOperator(RK4(), DIFF, 1e-4)
    g = FDMOperator(ForwardEulerMethod(), DIFF, 1e-3)
    fine = f.solve(ivp).discrete_y()
    p = PararealOperator(f, g, 1e-6)
    np.testing.assert_allclose(p.solve(ivp).discrete_y(), fine, rtol=0, atol=1e-14)
    np.testing.assert_array_equal(p.solve(ivp, parallel_enabled=False).discrete_y(), fine)


def test_parareal_coarse_grid_fp64_and_compensated_fine():
    """2x-coarsened coarse operator; compensated-FP32 fine solves compared
    against the FP64 oracle emulation at a stated 1e-6 relative tolerance."""
    print("case: parareal coarse grid", flush=True)
    cp, ic = diffusion_problem(33)
    t = (0.0, 0.04)
    g = CoarseGridFDMOperator(FDMOperator(ForwardEulerMethod(), DIFF, 5e-4))
    t_ref, y_ref, it_ref, _ = oracle_parareal(cp, ic, t, 4, 1e-4, 5e-4, "rk4", "fe", 1e-7,
                                              coarse_grid=True)
    f64 = PararealOperator(FDMOperator(RK4(), DIFF, 1e-4), g, 1e-7, n_slices=4)
    sol = f64.solve(pml.InitialValueProblem(cp, t, ic))
    assert f64.last_iterations == it_ref
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12
    f32c = PararealOperator(FDMOperator(RK4(), DIFF, 1e-4, precision="float32c"), g, 1e-7,
                            n_slices=4, output="endpoints")
    sol = f32c.solve(pml.InitialValueProblem(cp, t, ic))
    assert abs(f32c.last_iterations - it_ref) <= 1
    ends = y_ref.reshape(4, -1, *y_ref.shape[1:])[:, -1]
    assert rel_err(sol.discrete_y(), ends) <= 1e-6


def test_parareal_cahn_hilliard_multi_field_tolerances():
    print("case: parareal cahn-hilliard", flush=True)
    cp, ic = cahn_hilliard(20, seed=4)
    t = (0.0, 0.004)
    f = FDMOperator(RK4(), DIFF, 1e-4)
    g = FDMOperator(ForwardEulerMethod(), DIFF, 5e-4)
    p = PararealOperator(f, g, [1e-8, 1e-6], n_slices=4)
    sol = p.solve(pml.InitialValueProblem(cp, t, ic))
    t_ref, y_ref, it_ref, _ = oracle_parareal(cp, ic, t, 4, 1e-4, 5e-4, "rk4", "fe",
                                              [1e-8, 1e-6])
    assert p.last_iterations == it_ref
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12


def test_parareal_shallow_water_coarse_grid_fp64():
    """Three fields through the grid transfers: shallow water with a
    2x-coarsened RK4 coarse operator matches the oracle's emulation
    (iteration count and trajectory, FP64 to 1e-12), and N iterations
    reproduce the serial fine solution."""
    print("case: parareal shallow water coarse grid", flush=True)
    cp, ic = shallow_water(33, seed=3)
    t = (0.0, 0.08)
    g = CoarseGridFDMOperator(FDMOperator(RK4(), DIFF, 5e-3))
    t_ref, y_ref, it_ref, _ = oracle_parareal(cp, ic, t, 4, 1e-3, 5e-3, "rk4", "rk4", 1e-6,
                                              coarse_grid=True)
    p = PararealOperator(FDMOperator(RK4(), DIFF, 1e-3), g, 1e-6, n_slices=4)
    sol = p.solve(pml.InitialValueProblem(cp, t, ic))
    assert p.last_iterations == it_ref
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12
    full = PararealOperator(FDMOperator(RK4(), DIFF, 1e-3), g, 0.0, 4, n_slices=4)
    sol = full.solve(pml.InitialValueProblem(cp, t, ic))
    assert full.last_iterations == 4
    serial = FDMOperator(RK4(), DIFF, 1e-3).solve(pml.InitialValueProblem(cp, t, ic))
    assert rel_err(sol.discrete_y(), serial.discrete_y()) <= 1e-12


def test_parareal_validation_errors():
    cp, ic = diffusion_problem(17)
    ivp = pml.InitialValueProblem(cp, (0.0, 0.01), ic)
    f = FDMOperator(RK4(), DIFF, 3e-4)
    g = FDMOperator(ForwardEulerMethod(), DIFF, 1e-3)
    with pytest.raises(ValueError):
        PararealOperator(f, g, 1e-6, n_slices=4).solve(ivp)
    f = FDMOperator(RK4(), DIFF, 1e-4)
    with pytest.raises(ValueError):
        PararealOperator(f, g, [1e-6, 1e-6], n_slices=2).solve(ivp)


def test_same_inputs_replay_bit_identically():
    """Deterministic replay: the Parareal update norms and the Jacobi
    iteration's norms are summed in a fixed order, so the same inputs give
    bit-identical norms, iteration counts, sweep counts and results."""
    print("case: deterministic replay", flush=True)
    cp, ic = diffusion_problem(65)
    ivp = pml.InitialValueProblem(cp, (0.0, 0.04), ic)
    runs = []
    for _ in range(2):
        f = FDMOperator(RK4(), DIFF, 1e-4, precision="float32")
        g = CoarseGridFDMOperator(FDMOperator(ForwardEulerMethod(), DIFF, 5e-4))
        p = PararealOperator(f, g, 1e-7, n_slices=8)
        sol = p.solve(ivp)
        runs.append((p.last_iterations, p.last_stats["update_norms"], sol.discrete_y()))
    assert runs[0][0] == runs[1][0]
    assert runs[0][1] == runs[1][1]  # lists of floats, compared exactly
    np.testing.assert_array_equal(runs[0][2], runs[1][2])

    lap = np.random.RandomState(5).standard_normal(cp.y_vertices_shape)
    y_cons = cp.static_y_vertex_constraints
    outs = []
    for _ in range(2):
        np.random.seed(11)  # the reference's random start (numerical_differentiator.py:908)
        diff = ThreePointCentralDifferenceMethod(tol=1e-9)
        res = diff.anti_laplacian(lap, cp.mesh, y_cons)
        outs.append((diff.last_jacobi_sweeps, res))
    assert outs[0][0] == outs[1][0]
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_increment_form_equals_the_reference_update_for_fdm_coarse():
    """coarse_increments=True (U_new[j+1] = F(U_old[j]) + G(U_new[j] -
    U_old[j])) with a linear FDM coarse operator and zero Dirichlet edges
    gives the reference update's iterates (to rounding) and iteration count;
    it is refused for a non-linear equation, non-zero boundary values and a
    frozen learned correction."""
    from pararealml_amd.operators.ml import ConvSurrogate, CorrectedCoarseOperator

    print("case: parareal increment form", flush=True)
    cp, ic = diffusion_problem(33)
    ivp = pml.InitialValueProblem(cp, (0.0, 0.04), ic)
    f = FDMOperator(RK4(), DIFF, 1e-4)
    g = CoarseGridFDMOperator(FDMOperator(ForwardEulerMethod(), DIFF, 5e-4))
    its = []
    for inc in (False, True):
        seen = []
        p = PararealOperator(f, g, 1e-8, n_slices=4, output="endpoints", coarse_increments=inc,
                             monitor=lambda i, u, seen=seen: seen.append(u.cpu().numpy()) and False)
        sol = p.solve(ivp)
        # the end points are one slice apart (a plain float, not a device buffer)
        assert isinstance(sol.d_t, float) and sol.d_t == pytest.approx(0.01)
        its.append((p.last_iterations, seen))
    assert its[0][0] == its[1][0] >= 2
    for a, b in zip(its[0][1], its[1][1]):
        assert rel_err(b, a) <= 1e-12
    cp_ch, ic_ch = cahn_hilliard(17)
    with pytest.raises(ValueError):
        PararealOperator(f, g, 1e-6, n_slices=2, coarse_increments=True).solve(
            pml.InitialValueProblem(cp_ch, (0.0, 0.02), ic_ch))
    eq = pml.DiffusionEquation(2, 0.5)
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.0)], [1.0 / 16] * 2)
    one = pml.DirichletBoundaryCondition(lambda x, t: np.ones((len(x), 1)), is_static=True)
    cp1 = pml.ConstrainedProblem(eq, mesh, [(one, one), (one, one)])
    ic1 = pml.DiscreteInitialCondition(cp1, np.ones(cp1.y_vertices_shape), True)
    with pytest.raises(ValueError):
        PararealOperator(f, g, 1e-6, n_slices=2, coarse_increments=True).solve(
            pml.InitialValueProblem(cp1, (0.0, 0.02), ic1))
    frozen = CorrectedCoarseOperator(FDMOperator(ForwardEulerMethod(), DIFF, 5e-4),
                                     ConvSurrogate(1), base_levels=1, freeze=True)
    with pytest.raises(ValueError):
        PararealOperator(f, frozen, 1e-6, n_slices=2, coarse_increments=True).solve(ivp)


@pytest.mark.parametrize("prec", ["float64", "float32"])
def test_grid_transfers_match_oracle(prec):
    """pml_restrict2x / pml_prolong2x / pml_prolong2x_cubic against the
    oracle's injection and bilinear / cubic prolongation (themselves pinned
    by exactness on bilinear / cubic fields) on a batch of random fields:
    bitwise in FP64 (same operation order), to FP32 rounding in FP32; the
    cubic refuses grids with fewer than 4 points per axis."""
    import ctypes

    import torch

    from pararealml_amd import _native as nat

    print(f"case: grid transfers {prec}", flush=True)
    lib = nat.device_lib()
    dtype = torch.float64 if prec == "float64" else torch.float32
    pid = nat.PML_F64 if prec == "float64" else nat.PML_F32
    rng = np.random.RandomState(3)
    m0, m1, batch = 9, 14, 3
    coarse = rng.standard_normal((batch, m0, m1))
    fine = rng.standard_normal((batch, 2 * m0 - 1, 2 * m1 - 1))
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    c_dev = torch.tensor(coarse, dtype=dtype, device="cuda")
    f_dev = torch.tensor(fine, dtype=dtype, device="cuda")
    out_f = torch.empty_like(f_dev)
    out_c = torch.empty_like(c_dev)
    nat.check(lib.pml_restrict2x(pid, f_dev.data_ptr(), out_c.data_ptr(), batch, m0, m1, sp))
    got_r = out_c.cpu().numpy()
    tol = 0.0 if prec == "float64" else 1e-6
    for name, fn, ref in (("bilinear", lib.pml_prolong2x, ora.prolong_bilinear),
                          ("cubic", lib.pml_prolong2x_cubic, ora.prolong_cubic)):
        nat.check(fn(pid, c_dev.data_ptr(), out_f.data_ptr(), batch, m0, m1, sp))
        got = out_f.cpu().numpy()
        src = c_dev.cpu().numpy().astype(np.float64)
        for b in range(batch):
            want = ref(src[b][..., None])[..., 0]
            np.testing.assert_allclose(got[b], want, rtol=0, atol=tol * 8, err_msg=name)
    for b in range(batch):
        np.testing.assert_array_equal(got_r[b], f_dev.cpu().numpy()[b][::2, ::2])
    with pytest.raises(ValueError):
        nat.check(lib.pml_prolong2x_cubic(pid, c_dev.data_ptr(), out_f.data_ptr(), batch, 3,
                                          m1, sp))
