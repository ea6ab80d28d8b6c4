"""The BF16 MFMA surrogate kernels against the NumPy oracle (same BF16
rounding points, FP64 accumulation there vs FP32 on the device), training,
and the learned coarse propagator with its FDM fallback."""

import ctypes

import numpy as np
import pytest

from oracle import surrogate_numpy as so

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def random_params(c, seed):
    rng = np.random.RandomState(seed)
    params = so.init_params(c, c, seed=seed)
    w5 = so.bf16(rng.normal(0, 0.05, params[-1][0].shape))
    params[-1] = (w5, rng.normal(0, 0.01, c))
    return [(so.bf16(w), rng.normal(0, 0.05, b.shape)) for w, b in params]


@pytest.mark.parametrize("shape", [(2, 13, 37), (1, 32, 32), (3, 9, 70)])
def test_conv64_forward_matches_oracle(shape):
    import torch

    from pararealml_amd import _native as nat

    print(f"case: conv64 forward {shape}", flush=True)
    b, h, w = shape
    rng = np.random.RandomState(0)
    x = so.bf16(rng.randn(b, h, w, 64))
    wt = so.bf16(rng.normal(0, 0.1, (3, 3, 64, 64)))
    bias = rng.normal(0, 0.1, 64)
    z_ref = so.conv3x3(x, wt, bias)
    lib = nat.device_lib()
    xt = torch.tensor(x, dtype=torch.bfloat16, device="cuda").contiguous()
    wf = torch.tensor(wt.reshape(9, 64, 64), dtype=torch.bfloat16, device="cuda")
    wf = wf.permute(0, 2, 1).contiguous()
    bt = torch.tensor(bias, dtype=torch.float32, device="cuda")
    z = torch.empty_like(xt)
    a = torch.empty_like(xt)
    nat.check(lib.pml_conv64(0, xt.data_ptr(), wf.data_ptr(), bt.data_ptr(), None,
                             z.data_ptr(), a.data_ptr(), b, h, w, None))
    torch.cuda.synchronize()
    zg = z.float().cpu().numpy()
    ag = a.float().cpu().numpy()
    # z is stored in BF16: one BF16 ulp (2^-8 relative) of tolerance
    assert np.max(np.abs(zg - z_ref)) <= 2**-7 * np.max(np.abs(z_ref))
    assert rel(zg, z_ref) <= 4e-3
    assert rel(ag, so.gelu(z_ref)) <= 4e-3


def test_conv64_input_gradient_matches_oracle():
    import torch

    from pararealml_amd import _native as nat

    print("case: conv64 input gradient", flush=True)
    b, h, w = 2, 11, 40
    rng = np.random.RandomState(1)
    dz = so.bf16(rng.randn(b, h, w, 64))
    zprev = so.bf16(rng.randn(b, h, w, 64))
    wt = so.bf16(rng.normal(0, 0.1, (3, 3, 64, 64)))
    ref = so.conv3x3_input_grad(dz, wt) * so.gelu_grad(zprev)
    lib = nat.device_lib()
    wb = torch.tensor(wt.reshape(9, 64, 64), dtype=torch.bfloat16, device="cuda").flip(0).contiguous()
    dzt = torch.tensor(dz, dtype=torch.bfloat16, device="cuda")
    zpt = torch.tensor(zprev, dtype=torch.bfloat16, device="cuda")
    out = torch.empty_like(dzt)
    nat.check(lib.pml_conv64(1, dzt.data_ptr(), wb.data_ptr(), None, zpt.data_ptr(), None,
                             out.data_ptr(), b, h, w, None))
    torch.cuda.synchronize()
    assert rel(out.float().cpu().numpy(), ref) <= 4e-3


@pytest.mark.parametrize("shape", [(3, 10, 45), (2, 37, 70)])
@pytest.mark.parametrize("with_bias", [True, False])
def test_wgrad64_matches_oracle(shape, with_bias):
    import torch

    from pararealml_amd import _native as nat

    print(f"case: wgrad64 {shape} bias={with_bias}", flush=True)
    b, h, w = shape
    rng = np.random.RandomState(2)
    x = so.bf16(rng.randn(b, h, w, 64))
    dz = so.bf16(rng.randn(b, h, w, 64))
    ref = so.conv3x3_weight_grad(x, dz).reshape(9, 64, 64)
    lib = nat.device_lib()
    dw = torch.zeros(9 * 64 * 64, dtype=torch.float32, device="cuda")
    db = torch.zeros(64, dtype=torch.float32, device="cuda")
    xt = torch.tensor(x, dtype=torch.bfloat16, device="cuda")
    dzt = torch.tensor(dz, dtype=torch.bfloat16, device="cuda")
    n = ctypes.c_int64(0)
    nat.check(lib.pml_grad_partials_floats(64, 64, b, h, w, ctypes.byref(n)))
    part = torch.empty(n.value, dtype=torch.float32, device="cuda")
    runs = []
    # atomic sums, then twice the per-workgroup partials added in a fixed order
    for buf in (None, part, part):
        dw.zero_()
        db.zero_()
        nat.check(lib.pml_wgrad64(xt.data_ptr(), dzt.data_ptr(), dw.data_ptr(),
                                  db.data_ptr() if with_bias else None, b, h, w,
                                  None if buf is None else buf.data_ptr(),
                                  0 if buf is None else buf.numel(), None))
        torch.cuda.synchronize()
        runs.append((dw.cpu().numpy().copy(), db.cpu().numpy().copy()))
        assert rel(runs[-1][0].reshape(9, 64, 64), ref) <= 1e-5
        if with_bias:
            # the bias gradient is the dz column sum over all pixels
            db_ref = dz.reshape(-1, 64).sum(axis=0)
            assert rel(runs[-1][1], db_ref) <= 1e-5
        else:
            assert not runs[-1][1].any()
    np.testing.assert_array_equal(runs[1][0], runs[2][0])
    np.testing.assert_array_equal(runs[1][1], runs[2][1])
    with pytest.raises(ValueError):  # a partial buffer that is too small is refused
        nat.check(lib.pml_wgrad64(xt.data_ptr(), dzt.data_ptr(), dw.data_ptr(), None, b, h, w,
                                  part.data_ptr(), 16, None))


@pytest.mark.parametrize("c", [1, 3])
def test_network_forward_and_gradients_match_oracle(c):
    import torch

    from pararealml_amd.operators.ml import ConvSurrogate

    print(f"case: network gradients c={c}", flush=True)
    params = random_params(c, seed=4 + c)
    rng = np.random.RandomState(7)
    u = rng.rand(2, c, 20, 35) + 0.1
    target = u + 0.01 * rng.randn(*u.shape)
    net = ConvSurrogate(c)
    net.set_params(params)
    ut = torch.tensor(u, dtype=torch.float32, device="cuda")
    tt = torch.tensor(target, dtype=torch.float32, device="cuda")
    loss = net.loss_and_grads(ut, tt)
    torch.cuda.synchronize()
    # oracle on the same normalisation (per-sample max |u| in FP32)
    u32 = u.astype(np.float32).astype(np.float64)
    s = np.abs(u32).reshape(2, -1).max(axis=1)
    x = np.moveaxis(u32 / s[:, None, None, None], 1, -1)
    t32 = target.astype(np.float32).astype(np.float64)
    inc = np.moveaxis((t32 - u32) / s[:, None, None, None], 1, -1)
    loss_ref, grads = so.loss_and_grads(params, x, inc)
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    for layer, (gw_ref, gb_ref) in enumerate(grads):
        gw = net.gw[layer].cpu().numpy().reshape(gw_ref.shape)
        gb = net.gb[layer].cpu().numpy()
        assert rel(gw, gw_ref) <= 1e-2, (layer, rel(gw, gw_ref))
        assert rel(gb, gb_ref) <= 1e-2, (layer, rel(gb, gb_ref))
    # forward prediction
    pred = net.predict(ut.double()).cpu().numpy()
    ref = np.moveaxis(so.predict(params, np.moveaxis(u32, 1, -1)), -1, 1)
    assert rel(pred, ref) <= 5e-3


def test_adam_step_matches_oracle():
    import torch

    from pararealml_amd.operators.ml import ConvSurrogate

    print("case: adam", flush=True)
    params = random_params(1, seed=9)
    net = ConvSurrogate(1)
    net.set_params(params)
    rng = np.random.RandomState(3)
    grads = []
    for layer, (w, b) in enumerate(params):
        gw = rng.randn(*w.shape).astype(np.float32).astype(np.float64)
        gb = rng.randn(*b.shape).astype(np.float32).astype(np.float64)
        net.gw[layer].copy_(torch.tensor(gw.ravel(), dtype=torch.float32))
        net.gb[layer].copy_(torch.tensor(gb, dtype=torch.float32))
        grads.append((gw, gb))
    for _ in range(3):
        net.adam_step(1e-3)
    masters = [(w.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64))
               for w, b in params]
    zeros = lambda: [(np.zeros_like(w), np.zeros_like(b)) for w, b in params]  # noqa: E731
    state = (masters, zeros(), zeros(), 0)
    for _ in range(3):
        _, state = so.adam_step(params, grads, state, 1e-3)
    for (gw, gb), (rw, rb) in zip(net.params(), state[0]):
        np.testing.assert_allclose(gw, rw, rtol=1e-5, atol=2e-6)
        np.testing.assert_allclose(gb, rb, rtol=1e-5, atol=2e-6)


def test_training_learns_a_diffusion_step():
    """Training on (u, heat-step(u)) pairs reduces the loss well below the
    identity map's and the trained network beats the identity on held-out
    states."""
    import torch

    from pararealml_amd.operators.ml import ConvSurrogate

    print("case: training", flush=True)
    rng = np.random.RandomState(11)
    n, h = 96, 33

    def heat(u, k=3):
        for _ in range(k):
            p = np.pad(u, ((0, 0), (0, 0), (1, 1), (1, 1)))
            u = u + 0.2 * (p[:, :, 2:, 1:-1] + p[:, :, :-2, 1:-1] + p[:, :, 1:-1, 2:] + p[:, :, 1:-1, :-2] - 4 * u)
        return u

    g = np.linspace(0, 1, h)
    X, Y = np.meshgrid(g, g, indexing="ij")
    u = np.zeros((n, 1, h, h))
    for k in range(n):
        for _ in range(3):
            m = rng.uniform(0.2, 0.8, 2)
            v = rng.uniform(0.005, 0.02)
            u[k, 0] += np.exp(-((X - m[0]) ** 2 + (Y - m[1]) ** 2) / (2 * v))
    target = heat(u)
    ut = torch.tensor(u[:80], dtype=torch.float32, device="cuda")
    tt = torch.tensor(target[:80], dtype=torch.float32, device="cuda")
    net = ConvSurrogate(1, seed=0)
    losses = net.fit(ut, tt, steps=300, batch=16, lr=2e-3)
    assert np.mean(losses[-20:]) < 0.2 * np.mean(losses[:5])
    held = torch.tensor(u[80:], dtype=torch.float64, device="cuda")
    pred = net.predict(held).cpu().numpy()
    err_net = np.linalg.norm(pred - target[80:])
    err_id = np.linalg.norm(u[80:] - target[80:])
    assert err_net < 0.5 * err_id, (err_net, err_id)


def test_learned_coarse_fallback_on_non_finite_output():
    """A surrogate producing NaN falls back to the FDM coarse operator, so
    the Parareal result equals the FDM-coarse run."""
    import torch

    import pararealml_amd as pml
    from pararealml_amd.operators.fdm import (
        RK4,
        FDMOperator,
        ForwardEulerMethod,
        ThreePointCentralDifferenceMethod,
    )
    from pararealml_amd.operators.ml import ConvSurrogate, LearnedCoarseOperator
    from pararealml_amd.operators.parareal import CoarseGridFDMOperator, PararealOperator
    from tests.helpers import zero_bc

    print("case: learned coarse fallback", flush=True)
    diff = ThreePointCentralDifferenceMethod()
    eq = pml.DiffusionEquation(2, 0.5)
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.0)], [1 / 32, 1 / 32])
    bc = zero_bc(1)
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc), (bc, bc)])
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.5, 0.45]), 0.01 * np.eye(2))])
    ivp = pml.InitialValueProblem(cp, (0.0, 0.02), ic)
    f = FDMOperator(RK4(), diff, 1e-4)
    coarse_fdm = CoarseGridFDMOperator(FDMOperator(ForwardEulerMethod(), diff, 5e-4))
    net = ConvSurrogate(1)
    net.b32[0].fill_(float("nan"))
    net._refresh_bf16()
    g = LearnedCoarseOperator(net, 1.25e-3, coarsen=True, fallback=coarse_fdm)
    p_learned = PararealOperator(f, g, 1e-7, n_slices=4, output="endpoints")
    y_l = p_learned.solve(ivp).discrete_y()
    assert g.fallbacks > 0
    p_fdm = PararealOperator(f, coarse_fdm, 1e-7, n_slices=4, output="endpoints")
    y_f = p_fdm.solve(ivp).discrete_y()
    assert p_learned.last_iterations == p_fdm.last_iterations
    np.testing.assert_allclose(y_l, y_f, rtol=0, atol=1e-12)
    torch.cuda.synchronize()


def _diffusion_problem(n=32):
    import pararealml_amd as pml
    from tests.helpers import zero_bc

    eq = pml.DiffusionEquation(2, 0.5)
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.0)], [1 / n, 1 / n])
    bc = zero_bc(1)
    return pml.ConstrainedProblem(eq, mesh, [(bc, bc), (bc, bc)])


def test_corrected_coarse_with_identity_network_is_the_coarse_grid_fdm():
    """With an untrained surrogate (exactly the identity) and the base on the
    2x grid, the corrected operator is the coarse-grid FDM operator with the
    same base: identical Parareal iterates."""
    import pararealml_amd as pml
    from pararealml_amd.operators.fdm import (
        RK4,
        FDMOperator,
        ForwardEulerMethod,
        ThreePointCentralDifferenceMethod,
    )
    from pararealml_amd.operators.ml import ConvSurrogate, CorrectedCoarseOperator
    from pararealml_amd.operators.parareal import CoarseGridFDMOperator, PararealOperator

    print("case: corrected coarse identity", flush=True)
    diff = ThreePointCentralDifferenceMethod()
    cp = _diffusion_problem()
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.5, 0.45]), 0.01 * np.eye(2))])
    ivp = pml.InitialValueProblem(cp, (0.0, 0.02), ic)
    f = FDMOperator(RK4(), diff, 1e-4)
    base = FDMOperator(ForwardEulerMethod(), diff, 5e-4)
    g = CorrectedCoarseOperator(base, ConvSurrogate(1), base_levels=1)
    p_c = PararealOperator(f, g, 1e-7, n_slices=4, output="endpoints")
    y_c = p_c.solve(ivp).discrete_y()
    p_f = PararealOperator(f, CoarseGridFDMOperator(base), 1e-7, n_slices=4, output="endpoints")
    y_f = p_f.solve(ivp).discrete_y()
    assert p_c.last_iterations == p_f.last_iterations
    np.testing.assert_array_equal(y_c, y_f)
    with pytest.raises(ValueError):
        CorrectedCoarseOperator(base, ConvSurrogate(1), base_levels=5)
    with pytest.raises(ValueError):
        CorrectedCoarseOperator(base, ConvSurrogate(1), base_levels=1, correct_on=2)


@pytest.mark.parametrize("correct_on", ["coarse", "fine"])
def test_corrected_propagator_applies_the_network_in_stream_order(correct_on):
    """A network whose output is the constant beta (zero weights, last bias
    beta) turns the corrected prediction into the base prediction plus
    s * beta, s = max |base prediction|, at every point whose prolongation
    stencil is interior.  The surrogate is made on the default stream and
    the propagator runs on the solver stream: the correction has to land
    between the base solve and the prolongation."""
    import torch

    import pararealml_amd as pml
    from pararealml_amd import _device as dev
    from pararealml_amd.operators.fdm import (
        FDMOperator,
        ForwardEulerMethod,
        ThreePointCentralDifferenceMethod,
    )
    from pararealml_amd.operators.ml import ConvSurrogate, CorrectedCoarseOperator

    print(f"case: corrected propagator stream order ({correct_on})", flush=True)
    diff = ThreePointCentralDifferenceMethod()
    cp = _diffusion_problem(64)
    base = FDMOperator(ForwardEulerMethod(), diff, 2e-4)
    beta = float(np.float32(0.01))
    net = ConvSurrogate(1)
    params = [(np.zeros((3, 3, ci, co)), np.zeros(co))
              for ci, co in zip(net.chans[:-1], net.chans[1:])]
    params[-1] = (params[-1][0], np.full(1, beta))
    net.set_params(params)
    stream = dev.solver_stream()
    p_net = CorrectedCoarseOperator(base, net, base_levels=1, correct_on=correct_on
                                    ).slice_propagator(cp, 1, 50, stream)
    p_id = CorrectedCoarseOperator(base, ConvSurrogate(1), base_levels=1, correct_on=correct_on
                                   ).slice_propagator(cp, 1, 50, stream)
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.5, 0.45]), 0.01 * np.eye(2))])
    y0 = ic.discrete_y_0(True)
    u = torch.tensor(np.moveaxis(y0, -1, 0).reshape(1, -1), dtype=torch.float64, device="cuda")
    out_net = torch.zeros_like(u)
    out_id = torch.zeros_like(u)
    torch.cuda.synchronize()
    for _ in range(3):  # repeated calls reuse the workspaces
        p_net.run(u, 0, 1, out_net)
        p_id.run(u, 0, 1, out_id)
    torch.cuda.synchronize()
    y_net = out_net.cpu().numpy().reshape(65, 65)
    y_id = out_id.cpu().numpy().reshape(65, 65)
    # the network's scale: the maximum of the state it corrects (the 2x
    # grid's points are the even points of the prolonged output)
    s = float(np.float32(np.abs(y_id[::2, ::2] if correct_on == "coarse" else y_id).max()))
    # (a 2x-grid shift reaches the fine grid through the cubic prolongation,
    # whose 4-point stencils touch the zero Dirichlet edge up to fine index 3)
    r = slice(4, -4) if correct_on == "coarse" else slice(2, -2)
    np.testing.assert_allclose(y_net[r, r] - y_id[r, r], s * beta, rtol=0, atol=1e-12)
    # The correction was frozen at the first call: a perturbed input moves
    # the output exactly as it moves the base solve's.
    du = 1e-3 * torch.roll(u, 7, dims=1)
    out_net2 = torch.zeros_like(u)
    out_id2 = torch.zeros_like(u)
    p_net.run(u + du, 0, 1, out_net2)
    p_id.run(u + du, 0, 1, out_id2)
    torch.cuda.synchronize()
    np.testing.assert_allclose((out_net2 - out_net).cpu().numpy(),
                               (out_id2 - out_id).cpu().numpy(), rtol=0, atol=1e-13)


@pytest.mark.parametrize("correct_on", ["coarse", "fine"])
def test_trained_correction_reduces_the_coarse_error_on_held_out_states(correct_on):
    """What the network gains: trained on (cheap base prediction, fine
    solution) pairs, the corrected prediction over one slice is closer to
    the fine solution than the base prediction alone (at least 2x), on
    initial states not seen in training."""
    import torch

    from pararealml_amd.operators.fdm import (
        RK4,
        FDMOperator,
        ForwardEulerMethod,
        ThreePointCentralDifferenceMethod,
    )
    from pararealml_amd.operators.ml import (
        ConvSurrogate,
        corrected_pairs,
        gaussian_mixture_states,
        generate_pairs,
    )

    print(f"case: corrected coarse gain ({correct_on})", flush=True)
    n, slices = 64, 16
    diff = ThreePointCentralDifferenceMethod()
    cp = _diffusion_problem(n)
    delta = 1.0 / slices
    base = FDMOperator(ForwardEulerMethod(), diff, delta / 40)
    # RK4 at 90% of its stability limit for the 2-D diffusion stencil
    data_steps = int(np.ceil(delta / (0.9 * 2.78 * (1.0 / n) ** 2 / (8 * 0.5))))
    data_op = FDMOperator(RK4(), diff, delta / data_steps, precision="float32")
    coarsen = correct_on == "coarse"
    train = generate_pairs(data_op, cp, gaussian_mixture_states(cp, 64, seed=1), 1.0, delta,
                           coarsen=coarsen)
    inp, tgt = corrected_pairs(base, cp, train, delta, correct_on=correct_on)
    net = ConvSurrogate(1, seed=0)
    losses = net.fit(inp, tgt, steps=1500, batch=32, lr=1e-3)
    assert np.mean(losses[-20:]) < 0.05 * np.mean(losses[:5])
    held = generate_pairs(data_op, cp, gaussian_mixture_states(cp, 4, seed=2), 0.5, delta,
                          coarsen=coarsen)
    h_in, h_tg = corrected_pairs(base, cp, held, delta, correct_on=correct_on)
    x, t = h_in.double(), h_tg.double()
    pred = net.predict(x.clone())
    err_base = float(torch.linalg.norm(x - t) / torch.linalg.norm(t))
    err_net = float(torch.linalg.norm(pred - t) / torch.linalg.norm(t))
    print(f"held-out relative error: base {err_base:.3e}, corrected {err_net:.3e}", flush=True)
    assert err_net < 0.5 * err_base


def test_slice_guard_matches_its_definition():
    """pml_slice_guard sets the flag exactly when the prediction is not
    finite or, the input holding no NaN, max|dst| > guard max(max|src|,
    1e-300) -- the NumPy expression below -- and never clears it."""
    import torch

    from pararealml_amd import _native as nat

    print("case: slice guard", flush=True)
    lib = nat.device_lib()
    rng = np.random.default_rng(5)
    work = torch.empty(4 * nat.PML_REDUCE_SLOTS, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream

    def want(src, dst, guard):
        if not np.all(np.isfinite(dst)):
            return 1
        if np.any(np.isnan(src)):
            return 0
        return int(np.abs(dst).max() > guard * max(np.abs(src).max(), 1e-300))

    n = 300_000
    base = rng.standard_normal(n)
    cases = []
    for scale in (1.0, 9.9, 10.1, 1e3):
        cases.append((base, base * scale))
    bad = base.copy()
    bad[12345] = np.nan
    cases.append((base, bad))
    inf = base.copy()
    inf[-1] = -np.inf
    cases.append((base, inf))
    cases.append((bad, base * 1e6))     # NaN input: no ratio test
    cases.append((np.zeros(n), np.zeros(n)))
    cases.append((np.zeros(n), np.full(n, 1e-290)))
    for src, dst in cases:
        ts = torch.from_numpy(src).cuda()
        td = torch.from_numpy(dst).cuda()
        flag = torch.zeros(1, dtype=torch.int32, device="cuda")
        nat.check(lib.pml_slice_guard(ts.data_ptr(), td.data_ptr(), n, 10.0, work.data_ptr(),
                                      flag.data_ptr(), s))
        assert int(flag.item()) == want(src, dst, 10.0)
        # a set flag stays set
        flag.fill_(1)
        nat.check(lib.pml_slice_guard(ts.data_ptr(), ts.data_ptr(), n, 10.0, work.data_ptr(),
                                      flag.data_ptr(), s))
        assert int(flag.item()) == 1
    with pytest.raises(ValueError):
        nat.check(lib.pml_slice_guard(ts.data_ptr(), td.data_ptr(), 0, 10.0, work.data_ptr(),
                                      flag.data_ptr(), s))


def test_deferred_fallback_tests_equal_per_slice_tests(monkeypatch):
    """The Parareal driver defers the learned propagator's fallback tests to
    one host synchronisation per sweep and re-runs a sweep from the first
    flagged slice; with slices 1 and 2 of every sweep flagged, the iterates,
    the iteration count and the fallback count equal those of testing each
    slice as it is predicted."""
    import torch

    import pararealml_amd as pml
    from pararealml_amd.operators.fdm import (
        RK4,
        FDMOperator,
        ForwardEulerMethod,
        ThreePointCentralDifferenceMethod,
    )
    from pararealml_amd.operators.ml import ConvSurrogate, CorrectedCoarseOperator
    from pararealml_amd.operators.ml import learned_coarse as lc
    from pararealml_amd.operators.parareal import CoarseGridFDMOperator, PararealOperator
    from oracle import surrogate_numpy as so
    from tests.helpers import zero_bc

    print("case: deferred fallback tests", flush=True)
    diff = ThreePointCentralDifferenceMethod()
    eq = pml.DiffusionEquation(2, 0.5)
    mesh = pml.Mesh([(0.0, 1.0), (0.0, 1.0)], [1 / 32, 1 / 32])
    bc = zero_bc(1)
    cp = pml.ConstrainedProblem(eq, mesh, [(bc, bc), (bc, bc)])
    ic = pml.GaussianInitialCondition(cp, [(np.array([0.5, 0.45]), 0.01 * np.eye(2))])
    ivp = pml.InitialValueProblem(cp, (0.0, 0.04), ic)
    f = FDMOperator(RK4(), diff, 1e-4)
    fb = CoarseGridFDMOperator(FDMOperator(ForwardEulerMethod(), diff, 5e-4))
    rng = np.random.RandomState(2)
    params = so.init_params(1, 1, seed=2)
    params = params[:-1] + [(so.bf16(rng.normal(0, 0.005, params[-1][0].shape)),
                             np.zeros(1))]
    net = ConvSurrogate(1)
    net.set_params(params)

    def flag_slices_1_and_2(self, u_f64, out_f64, k, first, flag):
        if first + k in (1, 2):
            flag.fill_(1)

    monkeypatch.setattr(lc._FallbackGuard, "_test", flag_slices_1_and_2)
    runs = []
    for deferral in (True, False):
        monkeypatch.setattr(lc._FallbackGuard, "supports_deferral", deferral)
        g = CorrectedCoarseOperator(FDMOperator(ForwardEulerMethod(), diff, 1e-3), net,
                                    fallback=fb, base_levels=2, correct_on="base",
                                    freeze=False)
        seen = []
        p = PararealOperator(f, g, 1e-9, n_slices=4, output="endpoints",
                             monitor=lambda i, u, seen=seen: seen.append(u.cpu().numpy())
                             and False)
        y = p.solve(ivp).discrete_y()
        runs.append((p.last_iterations, g.fallbacks, seen, y))
    (it_d, fb_d, seen_d, y_d), (it_s, fb_s, seen_s, y_s) = runs
    assert it_d == it_s and fb_d == fb_s > 0
    for a, b in zip(seen_d, seen_s):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(y_d, y_s)
    torch.cuda.synchronize()


def test_training_is_bitwise_reproducible():
    """The weight gradients are summed per workgroup and then in a fixed
    order, so two trainings from the same seed on the same pairs end with
    bitwise-identical parameters (the small layers, the 64 -> 64 layers and
    the biases alike)."""
    import torch

    from pararealml_amd.operators.ml import ConvSurrogate

    print("case: reproducible training", flush=True)
    rng = np.random.RandomState(9)
    u = torch.tensor(rng.standard_normal((24, 1, 33, 33)), dtype=torch.float32, device="cuda")
    t = u + 0.01 * torch.tensor(rng.standard_normal((24, 1, 33, 33)), dtype=torch.float32,
                                device="cuda")
    params = []
    for _ in range(2):
        net = ConvSurrogate(1, seed=4)
        net.fit(u, t, steps=30, batch=8, lr=1e-3, seed=1, mix=0.5)
        torch.cuda.synchronize()
        params.append([(w.cpu().numpy().copy(), b.cpu().numpy().copy())
                       for w, b in zip(net.w32, net.b32)])
    for (w0, b0), (w1, b1) in zip(*params):
        np.testing.assert_array_equal(w0, w1)
        np.testing.assert_array_equal(b0, b1)
    assert any(np.any(w != 0) for w, _ in params[0])
