"""Parity of the fused gfx950 FDM kernels (through FDMOperator and the C ABI)
with the CPU oracle.

Tolerances: FP64 device results must match the FP64 oracle to a relative
max-norm difference of 1e-12 (different but equally rounded op order and
reciprocal multiplications); FP32 results are compared with the FP64 oracle
at a tolerance derived from the step count (stated per test)."""

import numpy as np
import pytest

import pararealml_amd as pml
from oracle import fdm_numpy as ora
from pararealml_amd.operators.fdm import (
    RK4,
    ExplicitMidpointMethod,
    FDMOperator,
    ForwardEulerMethod,
    ThreePointCentralDifferenceMethod,
)
from tests.helpers import cahn_hilliard, diffusion_c1, shallow_water, zero_bc

pytestmark = pytest.mark.gpu

INTEGRATORS = {"fe": ForwardEulerMethod, "midpoint": ExplicitMidpointMethod, "rk4": RK4}


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run_both(cp, ic, t_interval, d_t, integ="fe", precision="float64", **kw):
    ivp = pml.InitialValueProblem(cp, t_interval, ic)
    op = FDMOperator(INTEGRATORS[integ](), ThreePointCentralDifferenceMethod(), d_t,
                     precision=precision, **kw)
    sol = op.solve(ivp)
    t_ref, y_ref = ora.fdm_solve(cp, ic.discrete_y_0(True), t_interval, d_t, integ)
    np.testing.assert_allclose(sol.t_coordinates, t_ref)
    return sol, y_ref


def test_c1_diffusion_fe_fp64_parity():
    """The reference-runnable configuration: 101x101, FE, d_t=0.002, 1000
    steps, zero Dirichlet, Gaussian IC; full trajectory."""
    print("case: c1 diffusion fe fp64 101x101x1000", flush=True)
    cp, ic = diffusion_c1()
    sol, y_ref = run_both(cp, ic, (0.0, 2.0), 0.002, "fe")
    y = sol.discrete_y()
    assert y.shape == (1000, 101, 101, 1)
    assert sol.d_t == 0.002
    assert rel_err(y, y_ref) <= 1e-12
    # Dirichlet edges are exactly zero at every step
    assert np.all(y[:, 0] == 0) and np.all(y[:, :, -1] == 0)


def test_c1_diffusion_fe_fp32_tolerance():
    """FP32 state vs the FP64 oracle after 1000 steps.  The oracle's own
    float32 mode drifts by 1.0e-7 (relative max norm) on this case; the
    stated tolerance is 1e-6."""
    print("case: c1 diffusion fe fp32", flush=True)
    cp, ic = diffusion_c1()
    sol, y_ref = run_both(cp, ic, (0.0, 2.0), 0.002, "fe", precision="float32")
    err = rel_err(sol.discrete_y(), y_ref)
    assert err <= 1e-6, err


def test_compensated_fp32_beats_plain_fp32():
    print("case: c1 diffusion fe fp32c", flush=True)
    cp, ic = diffusion_c1(n=41, length=4.0)
    t = (0.0, 1.0)
    ivp = pml.InitialValueProblem(cp, t, ic)
    _, y_ref = ora.fdm_solve(cp, ic.discrete_y_0(True), t, 0.001, "rk4", keep="last")
    errs = {}
    for prec in ("float32", "float32c"):
        op = FDMOperator(RK4(), ThreePointCentralDifferenceMethod(), 0.001, precision=prec)
        errs[prec] = rel_err(op.solve(ivp).discrete_y()[-1], y_ref[0])
    assert errs["float32c"] <= 1e-6, errs
    assert errs["float32c"] < errs["float32"], errs


@pytest.mark.parametrize("integ", ["fe", "midpoint", "rk4"])
def test_diffusion_mixed_bcs_fp64(integ):
    """Dirichlet on axis 0 (different constants, so the axis-1 Neumann edge
    corners show the later-axis rule), non-zero Neumann on axis 1."""
    print(f"case: diffusion mixed bcs {integ}", flush=True)
    eq = pml.DiffusionEquation(2, 0.7)
    mesh = pml.Mesh([(0.0, 2.0), (0.0, 3.0)], [0.1, 0.125])
    bcs = [
        (
            pml.DirichletBoundaryCondition(lambda x, t: np.full((len(x), 1), 0.5), is_static=True),
            pml.DirichletBoundaryCondition(lambda x, t: x[:, 1:2] * 0.1, is_static=True),
        ),
        (
            pml.NeumannBoundaryCondition(lambda x, t: np.full((len(x), 1), -0.3), is_static=True),
            pml.CauchyBoundaryCondition(
                lambda x, t: np.full((len(x), 1), 1.0),
                lambda x, t: x[:, :1] * 0.2,
                is_static=True,
            ),
        ),
    ]
    cp = pml.ConstrainedProblem(eq, mesh, bcs)
    ic = pml.GaussianInitialCondition(cp, [(np.array([1.0, 1.5]), 0.2 * np.eye(2))], [2.0])
    sol, y_ref = run_both(cp, ic, (0.0, 0.2), 0.002, integ)
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12


def test_partial_nan_boundary_conditions_fp64():
    """Boundary functions returning None (NaN) for some points and fields."""
    print("case: partial nan bcs", flush=True)
    eq = pml.WaveEquation(2, 0.8)

    def lower(x, t):
        return (0.0 if x[1] < 1.0 else None, None)

    bcs = [
        (
            pml.DirichletBoundaryCondition(pml.vectorize_bc_function(lower), is_static=True),
            pml.NeumannBoundaryCondition(
                pml.vectorize_bc_function(lambda x, t: (None, 0.5)), is_static=True
            ),
        ),
        (
            pml.CauchyBoundaryCondition(
                pml.vectorize_bc_function(lambda x, t: (None, 0.1)),
                pml.vectorize_bc_function(lambda x, t: (0.2, None)),
                is_static=True,
            ),
            zero_bc(2, "neumann"),
        ),
    ]
    mesh = pml.Mesh([(0.0, 2.0), (0.0, 2.0)], [0.1, 0.1])
    cp = pml.ConstrainedProblem(eq, mesh, bcs)
    ic = pml.GaussianInitialCondition(
        cp, [(np.array([1.0, 1.0]), 0.1 * np.eye(2))] * 2, [1.0, 0.0]
    )
    sol, y_ref = run_both(cp, ic, (0.0, 0.1), 0.005, "rk4")
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12


@pytest.mark.parametrize("integ", ["fe", "rk4"])
def test_cahn_hilliard_fp64(integ):
    print(f"case: cahn-hilliard {integ}", flush=True)
    cp, ic = cahn_hilliard(24, seed=3)
    sol, y_ref = run_both(cp, ic, (0.0, 0.005), 1e-4, integ)
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12


def test_cahn_hilliard_rk4_fp32_tolerance():
    """FP32 CH (the batched benchmark equation), 200 steps; stated tolerance
    1e-5 relative."""
    print("case: cahn-hilliard fp32", flush=True)
    cp, ic = cahn_hilliard(40, seed=5)
    sol, y_ref = run_both(cp, ic, (0.0, 0.02), 1e-4, "rk4", precision="float32")
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-5


@pytest.mark.parametrize("integ", ["fe", "midpoint", "rk4"])
def test_shallow_water_fp64(integ):
    print(f"case: shallow water {integ}", flush=True)
    cp, ic = shallow_water(21, seed=1)
    sol, y_ref = run_both(cp, ic, (0.0, 0.03), 0.001, integ)
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12


def test_shallow_water_reference_example_bcs_fp64():
    """The reference example's boundary set: Neumann on eta only, u and v
    free (pararealml examples/shallow_water_fdm.py)."""
    print("case: shallow water example bcs", flush=True)
    eq = pml.ShallowWaterEquation(0.5)
    mesh = pml.Mesh([(-5.0, 5.0), (0.0, 5.0)], [0.5, 0.25])
    bc = pml.NeumannBoundaryCondition(
        pml.vectorize_bc_function(lambda x, t: (0.0, None, None)), is_static=True
    )
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc)] * 2)
    ic = pml.GaussianInitialCondition(
        cp, [(np.array([2.5, 1.25]), 0.25 * np.eye(2))] * 3, [1.0, 0.0, 0.0]
    )
    sol, y_ref = run_both(cp, ic, (0.0, 0.05), 0.0025, "rk4")
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12


def test_convection_diffusion_and_burgers_fp64():
    print("case: convection-diffusion / burgers 2d", flush=True)
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.0)], [0.05, 0.05])
    bc1 = zero_bc(1)
    cp = pml.ConstrainedProblem(
        pml.ConvectionDiffusionEquation(2, [0.3, -0.2], 0.05), mesh, [(bc1, bc1)] * 2
    )
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.5, 0.5]), 0.02 * np.eye(2))])
    sol, y_ref = run_both(cp, ic, (0.0, 0.02), 0.001, "rk4")
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12

    bc2 = zero_bc(2)
    cp = pml.ConstrainedProblem(pml.BurgersEquation(2, 100.0), mesh, [(bc2, bc2)] * 2)
    ic = pml.GaussianInitialCondition(
        cp, [(np.array([0.5, 0.5]), 0.02 * np.eye(2))] * 2, [0.5, 0.3]
    )
    sol, y_ref = run_both(cp, ic, (0.0, 0.02), 0.001, "midpoint")
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12


@pytest.mark.parametrize("eq_name", ["diffusion", "burgers", "wave"])
def test_one_dimensional_fp64(eq_name):
    """1-D problems run on the line-along-axis-1 layout; Cauchy lower end,
    Neumann upper end."""
    print(f"case: 1-d {eq_name}", flush=True)
    eq = {
        "diffusion": pml.DiffusionEquation(1, 0.5),
        "burgers": pml.BurgersEquation(1, 50.0),
        "wave": pml.WaveEquation(1, 1.2),
    }[eq_name]
    n_f = eq.y_dimension
    mesh = pml.Mesh([(0.0, 4.0)], [0.1])
    bcs = [
        (
            pml.CauchyBoundaryCondition(
                lambda x, t: np.full((len(x), n_f), 0.25),
                lambda x, t: np.full((len(x), n_f), 0.1),
                is_static=True,
            ),
            pml.NeumannBoundaryCondition(
                lambda x, t: np.full((len(x), n_f), -0.2), is_static=True
            ),
        )
    ]
    cp = pml.ConstrainedProblem(eq, mesh, bcs)
    ic = pml.GaussianInitialCondition(cp, [(np.array([2.0]), np.array([[0.3]]))] * n_f)
    sol, y_ref = run_both(cp, ic, (0.0, 0.1), 0.002, "rk4")
    assert sol.discrete_y().shape == (50, 41, n_f)
    assert rel_err(sol.discrete_y(), y_ref) <= 1e-12


def test_all_tile_configs_bitwise_identical():
    """Every kernel variant (tile height, fused steps per launch) and the
    graph / eager launch paths compute the same operations per vertex."""
    print("case: tile configs", flush=True)
    from pararealml_amd import _device as dev
    from pararealml_amd import _native as nat

    cp, ic = cahn_hilliard(70, seed=2)
    problem = dev.device_problem(cp)
    lib = nat.device_lib()
    results = []
    for prec in ("float32", "float32c"):
        n_cfg = lib.pml_fdm_configs(nat.PML_EQ["cahn_hilliard"], nat.PML_INT_RK4, dev.precision_id(prec))
        assert n_cfg >= 1
        outs = []
        for cfg in range(n_cfg):
            for graph in (False, True):
                b = dev.FDMBatch([problem], [0], prec, nat.PML_INT_RK4)
                b.load_host([ic.discrete_y_0(True)])
                b.advance(1e-4, 130, tile_config=cfg, use_graph=graph)
                outs.append(b.state_f64().cpu().numpy())
        for o in outs[1:]:
            np.testing.assert_array_equal(o, outs[0])
        results.append(outs[0])


def _mixed_bc_diffusion(n0, n1):
    """Diffusion with Dirichlet rows (axis 0) and Neumann / Cauchy columns
    (axis 1) on an n0 x n1 vertex grid."""
    eq = pml.DiffusionEquation(2, 0.7)
    mesh = pml.Mesh([(0.0, 0.1 * (n0 - 1)), (0.0, 0.125 * (n1 - 1))], [0.1, 0.125])
    bcs = [
        (
            pml.DirichletBoundaryCondition(lambda x, t: np.full((len(x), 1), 0.5), is_static=True),
            pml.DirichletBoundaryCondition(lambda x, t: x[:, 1:2] * 0.1, is_static=True),
        ),
        (
            pml.NeumannBoundaryCondition(lambda x, t: np.full((len(x), 1), -0.3), is_static=True),
            pml.CauchyBoundaryCondition(
                lambda x, t: np.full((len(x), 1), 1.0), lambda x, t: x[:, :1] * 0.2, is_static=True
            ),
        ),
    ]
    cp = pml.ConstrainedProblem(eq, mesh, bcs)
    centre = np.array([0.05 * (n0 - 1), 0.06 * (n1 - 1)])
    ic = pml.GaussianInitialCondition(cp, [(centre, 0.3 * np.eye(2))], [2.0])
    return cp, ic


def _mixed_bc_shallow_water(n0, n1):
    """Shallow water with Dirichlet / Neumann rows (axis 0) and Neumann /
    Cauchy columns (axis 1) on an n0 x n1 vertex grid."""
    eq = pml.ShallowWaterEquation(h=1.0, v=0.01, g=9.81)
    mesh = pml.Mesh([(0.0, 0.1 * (n0 - 1)), (0.0, 0.125 * (n1 - 1))], [0.1, 0.125])
    bcs = [
        (
            pml.DirichletBoundaryCondition(lambda x, t: np.full((len(x), 3), 0.05), is_static=True),
            pml.NeumannBoundaryCondition(lambda x, t: np.zeros((len(x), 3)), is_static=True),
        ),
        (
            pml.NeumannBoundaryCondition(lambda x, t: np.full((len(x), 3), -0.02), is_static=True),
            pml.CauchyBoundaryCondition(
                lambda x, t: np.full((len(x), 3), 1.0),
                lambda x, t: np.repeat(x[:, :1] * 0.01, 3, axis=1),
                is_static=True,
            ),
        ),
    ]
    cp = pml.ConstrainedProblem(eq, mesh, bcs)
    centre = np.array([0.05 * (n0 - 1), 0.06 * (n1 - 1)])
    ic = pml.GaussianInitialCondition(cp, [(centre, 0.3 * np.eye(2))] * 3, [0.2, 0.0, 0.0])
    return cp, ic


def _check_stream_variants(eq_name, cases, integ, prec, d_t):
    """Runs every kernel variant, eager and graph, on the ragged batch of
    `cases` with snapshots; asserts bitwise-identical states and snapshots
    and, in FP64, the oracle per grid."""
    from pararealml_amd import _device as dev
    from pararealml_amd import _native as nat

    problems = [dev.device_problem(cp) for cp, _ in cases]
    y0s = [ic.discrete_y_0(True) for _, ic in cases]
    lib = nat.device_lib()
    int_id = {"fe": nat.PML_INT_FE, "midpoint": nat.PML_INT_MIDPOINT, "rk4": nat.PML_INT_RK4}[integ]
    n_cfg = lib.pml_fdm_configs(nat.PML_EQ[eq_name], int_id, dev.precision_id(prec))
    names = [lib.pml_fdm_config_name(nat.PML_EQ[eq_name], int_id,
                                     dev.precision_id(prec), c).decode() for c in range(n_cfg)]
    assert any(("-c2-" if integ == "fe" else "streamrk-") in n for n in names)
    steps, every = 37, 4
    outs = []
    for cfg in range(n_cfg):
        for graph in (False, True):
            b = dev.FDMBatch(problems, list(range(len(cases))), prec, int_id, dev.solver_stream())
            b.load_host(y0s)
            snaps = b.new_snapshots(steps // every)
            b.advance(d_t, steps, snapshots=snaps, snapshot_every=every, tile_config=cfg,
                      use_graph=graph, constrain_input=True)
            outs.append((names[cfg], graph, b.state_f64().cpu().numpy(),
                         b.snapshots_f64(snaps, steps // every).cpu().numpy()))
    for name, graph, st, sn in outs[1:]:
        np.testing.assert_array_equal(st, outs[0][2], err_msg=f"{name} graph={graph}")
        np.testing.assert_array_equal(sn, outs[0][3], err_msg=f"{name} graph={graph} snapshots")
    if prec == "float64":
        flat = outs[0][2]
        for k, (cp, ic) in enumerate(cases):
            _, y_ref = ora.fdm_solve(cp, ic.discrete_y_0(True), (0.0, steps * d_t), d_t, integ,
                                     keep="last")
            got = problems[k].from_device_layout(b.grid_view(flat, k), cp.mesh.vertices_shape)
            assert rel_err(got, y_ref[0]) <= 1e-12


@pytest.mark.parametrize("integ", ["fe", "midpoint", "rk4"])
@pytest.mark.parametrize("prec", ["float64", "float32", "float32c"])
@pytest.mark.parametrize("shapes", [((21, 130), (34, 66)), ((21, 130), (17, 33))])
def test_stream_variants_identical_with_boundaries_and_snapshots(integ, prec, shapes):
    """Every variant (one and two columns per lane, ring depths, two steps
    per launch), eager and graph launches, on a ragged batch with Dirichlet,
    Neumann and Cauchy edges, strips that end inside the grid and snapshots
    (an odd step count, so fused variants also run a single-step launch):
    bitwise-identical states and snapshots.  The second batch has an odd row length, where the
    two-column variants hand over to the default one.  FP64 matches the
    oracle per grid."""
    print(f"case: stream variants {integ} {prec} {shapes}", flush=True)
    _check_stream_variants("diffusion", [_mixed_bc_diffusion(*s) for s in shapes], integ, prec,
                           2e-4)


@pytest.mark.parametrize("integ", ["midpoint", "rk4"])
@pytest.mark.parametrize("prec", ["float64", "float32", "float32c"])
def test_shallow_water_stream_variants_identical(integ, prec):
    """Three fields take the streaming multi-stage kernel's halo-free layout
    (strips of 64 - 2S columns): every variant, eager and graph, on a ragged
    batch with mixed edges and snapshots gives bitwise-identical results,
    and FP64 matches the oracle per grid."""
    print(f"case: shallow water stream variants {integ} {prec}", flush=True)
    cases = [_mixed_bc_shallow_water(21, 130), _mixed_bc_shallow_water(34, 66),
             _mixed_bc_shallow_water(17, 57)]
    _check_stream_variants("shallow_water", cases, integ, prec, 2e-3)


def test_ragged_batch_matches_per_grid():
    """One launch over grids of different sizes equals per-grid solves."""
    print("case: ragged batch", flush=True)
    from pararealml_amd import _device as dev
    from pararealml_amd import _native as nat

    sizes = [16, 24, 40, 33]
    problems, y0s = [], []
    for i, n in enumerate(sizes):
        cp, ic = cahn_hilliard(n, seed=10 + i)
        problems.append(dev.device_problem(cp))
        y0s.append(ic.discrete_y_0(True))
    b = dev.FDMBatch(problems, list(range(len(sizes))), "float64", nat.PML_INT_RK4)
    b.load_host(y0s)
    b.advance(1e-4, 25)
    flat = b.state_f64().cpu().numpy()
    for k, n in enumerate(sizes):
        cp, ic = cahn_hilliard(n, seed=10 + k)
        _, y_ref = ora.fdm_solve(cp, ic.discrete_y_0(True), (0.0, 25e-4), 1e-4, "rk4", keep="last")
        got = problems[k].from_device_layout(b.grid_view(flat, k), (n, n))
        assert rel_err(got, y_ref[0]) <= 1e-12


def test_unsupported_cases_raise():
    from pararealml_amd.constrained_problem import ConstrainedProblem
    from pararealml_amd.operators.fdm import BackwardEulerMethod

    # what the device path does not run is refused, never run on the host:
    # ODEs with an implicit integrator or in compensated FP32
    cp = ConstrainedProblem(pml.LorenzEquation())
    ic = pml.ContinuousInitialCondition(cp, lambda _: np.ones(3))
    ivp = pml.InitialValueProblem(cp, (0.0, 1.0), ic)
    with pytest.raises(NotImplementedError):
        FDMOperator(BackwardEulerMethod(), ThreePointCentralDifferenceMethod(), 0.1).solve(ivp)
    with pytest.raises(NotImplementedError):
        FDMOperator(RK4(), ThreePointCentralDifferenceMethod(), 0.1,
                    precision="float32c").solve(ivp)


def test_c2_full_ragged_batch_fp32_against_fp64():
    """The batched Cahn-Hilliard workload at full size: 256 grids drawn from
    {64, 96, 128, 192, 256}^2, RK4, 2000 steps of 1e-4, the tuned FP32
    streaming kernel (two steps per launch, graph replays) against the FP64
    kernel (itself pinned to the oracle per grid): per grid, max relative
    error <= 2e-5, and the trapezoid mass of c (conserved exactly by the
    zero-Neumann scheme) drifts by less than 1e-5 of the field's scale."""
    import bench
    from pararealml_amd import _device as dev
    from pararealml_amd import _native as nat

    print("case: c2 full ragged batch fp32 vs fp64, 2000 steps", flush=True)
    sizes, cps, c0 = bench.ch_workload()
    y0 = bench.mu0_on_device(cps, sizes, c0)
    size_list = sorted(cps)
    problems = [dev.device_problem(cps[n]) for n in size_list]
    grid_problem = [size_list.index(n) for n in sizes]
    out = {}
    for prec in ("float32", "float64"):
        b = dev.FDMBatch(problems, grid_problem, prec, nat.PML_INT_RK4, dev.solver_stream())
        b.load_host(y0)
        names = [b.lib.pml_fdm_config_name(b.eq_id, b.integrator, b.prec, i).decode()
                 for i in range(b.lib.pml_fdm_configs(b.eq_id, b.integrator, b.prec))]
        cfg = names.index("stream2-rw8-pd2-s2") if prec == "float32" else 0
        b.advance(1e-4, 2000, tile_config=cfg, use_graph=True)
        out[prec] = (b, b.state_f64().cpu().numpy())
    worst = 0.0
    for k, n in enumerate(sizes):
        a = out["float32"][0].grid_view(out["float32"][1], k)
        r = out["float64"][0].grid_view(out["float64"][1], k)
        worst = max(worst, float(np.max(np.abs(a - r)) / np.max(np.abs(r))))
        w = np.ones(n)
        w[0] = w[-1] = 0.5
        w2 = np.outer(w, w)
        m0 = float(np.sum(w2 * c0[k][..., 0]))
        m1 = float(np.sum(w2 * a[0]))
        assert abs(m1 - m0) <= 1e-5 * 0.05 * n * n, (k, m0, m1)
    assert worst <= 2e-5, worst


def test_shallow_water_fp32_against_fp64():
    """Shallow water (h=1, v=0.01, g=9.81, reflective edges), four 257^2
    grids with seeded bumps, 200 RK4 steps of 1e-3 in FP32 against FP64 on
    all three fields: per field, max error <= 2e-5 of the field's largest
    magnitude (the velocity fields start at zero, so the scale is the
    field's own)."""
    from pararealml_amd import _device as dev
    from pararealml_amd import _native as nat

    print("case: shallow water fp32 vs fp64", flush=True)
    eq = pml.ShallowWaterEquation(h=1.0, v=0.01, g=9.81)
    mesh = pml.Mesh([(-5.0, 5.0), (-5.0, 5.0)], [10.0 / 256] * 2)
    bc = pml.NeumannBoundaryCondition(lambda x, t: np.zeros((len(x), 3)), is_static=True)
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc), (bc, bc)])
    rng = np.random.RandomState(1462316)
    y0s = []
    for _ in range(4):
        mean = rng.uniform(-2.5, 2.5, 2)
        ic = pml.GaussianInitialCondition(cp, [(mean, 0.25 * np.eye(2))] * 3, [0.1, 0.0, 0.0])
        y0s.append(ic.discrete_y_0(True))
    problem = dev.device_problem(cp)
    res = {}
    for prec in ("float32", "float64"):
        b = dev.FDMBatch([problem], [0] * 4, prec, nat.PML_INT_RK4, dev.solver_stream())
        b.load_host(y0s)
        b.advance(1e-3, 200, use_graph=True)
        res[prec] = b.state_f64().cpu().numpy().reshape(4, 3, 257, 257)
    for q in range(3):
        scale = float(np.max(np.abs(res["float64"][:, q])))
        assert scale > 0
        err = float(np.max(np.abs(res["float32"][:, q] - res["float64"][:, q]))) / scale
        assert err <= 2e-5, (q, err)
