"""The convolutional surrogate as a Parareal coarse propagator, its training
data generation from the device FDM solver, and the FDM fallback.

Two forms: ``LearnedCoarseOperator`` replaces the coarse solve by repeated
surrogate steps; ``CorrectedCoarseOperator`` runs a cheap FDM solve on a
further coarsened grid and applies the surrogate once per slice as a
learned correction (stable under the Parareal sweep's repeated
application).

``LearnedCoarseOperator(surrogate, d_t, coarsen=True, fallback=None)`` steps
the state by ``d_t`` per network application (the reference's learned
operators step autoregressively by their ``d_t``,
pararealml/operators/ml/supervised/supervised_ml_operator.py:91-128).  With
``coarsen`` the network works on the 2x-coarsened vertex grid.  After every
application the problem's Dirichlet values are imposed.  When the output of
a slice is not finite or its magnitude exceeds ``guard`` times the input's,
that slice is recomputed by the ``fallback`` operator (normally the plain FDM
coarse operator) and counted in ``fallbacks``.
"""

import ctypes
from typing import Callable, Optional

import numpy as np

from pararealml_amd import _device as dev
from pararealml_amd import _native as nat
from pararealml_amd.operators.ml.conv_surrogate import ConvSurrogate


def coarse_problem(cp):
    """The constrained problem on the 2x-coarsened vertex grid."""
    from pararealml_amd.constrained_problem import ConstrainedProblem
    from pararealml_amd.mesh import Mesh

    mesh = cp.mesh
    if mesh.dimensions != 2 or any((n - 1) % 2 for n in mesh.vertices_shape):
        raise ValueError("coarsening needs a 2-D grid with 2m - 1 vertices per axis")
    cmesh = Mesh(mesh.x_intervals, [2.0 * d for d in mesh.d_x], mesh.coordinate_system_type)
    return ConstrainedProblem(cp.differential_equation, cmesh, cp.boundary_conditions)


class LearnedCoarseOperator:
    """Parareal coarse operator backed by a ``ConvSurrogate``."""

    def __init__(self, surrogate: ConvSurrogate, d_t: float, coarsen: bool = True,
                 fallback=None, guard: float = 10.0):
        if d_t <= 0.0:
            raise ValueError("time step size must be greater than 0")
        self.surrogate = surrogate
        self._d_t = d_t
        self.coarsen = coarsen
        self.fallback = fallback
        self.guard = guard
        self.fallbacks = 0

    @property
    def d_t(self) -> float:
        return self._d_t

    @property
    def vertex_oriented(self) -> Optional[bool]:
        return True

    def slice_propagator(self, cp, n_slices, slice_steps, stream, record_trajectory=False,
                         borders=None):
        fb = None
        if self.fallback is not None:
            fb = self.fallback.slice_propagator(
                cp, n_slices, int(round(slice_steps * self._d_t / self.fallback.d_t)), stream
            )
        return LearnedPropagator(self, cp, slice_steps, stream, fb)


class LearnedPropagator:
    def __init__(self, op: LearnedCoarseOperator, cp, steps: int, stream, fallback_prop):
        import torch

        self.op = op
        self.steps = steps
        self.stream = stream
        self.fallback = fallback_prop
        self.torch = torch
        self.fine = dev.device_problem(cp)
        self.work_cp = coarse_problem(cp) if op.coarsen else cp
        self.work = dev.device_problem(self.work_cp)
        self.bc_fine = torch.from_numpy(self.fine.bc).to("cuda")
        self.bc_work = torch.from_numpy(self.work.bc).to("cuda")
        self.buf = torch.empty(
            (1, self.work.n_fields, self.work.n0, self.work.n1), dtype=torch.float64, device="cuda"
        )
        self.lib = nat.device_lib()

    def _dirichlet(self, t, p, bc):
        nat.check(self.lib.pml_apply_dirichlet(
            t.data_ptr(), bc.data_ptr(), p.n_fields, p.n0, p.n1, p.ndim, 1,
            ctypes.c_void_p(self.stream.cuda_stream)))

    def run(self, u_f64, first: int, last: int, out_f64):
        torch = self.torch
        sp = ctypes.c_void_p(self.stream.cuda_stream)
        net = self.op.surrogate
        # every torch operation below (copies, the fallback test) runs on
        # the propagator's stream, in order with its kernels
        with torch.cuda.stream(self.stream):
            for k in range(last - first):
                src = u_f64[k]
                if self.op.coarsen:
                    nat.check(self.lib.pml_restrict2x(
                        nat.PML_F64, src.data_ptr(), self.buf.data_ptr(), self.work.n_fields,
                        self.work.n0, self.work.n1, sp))
                else:
                    self.buf.view(-1).copy_(src)
                self._dirichlet(self.buf, self.work, self.bc_work)
                for _ in range(self.steps):
                    net.predict(self.buf, out=self.buf, stream=self.stream)
                    self._dirichlet(self.buf, self.work, self.bc_work)
                dst = out_f64[k]
                if self.op.coarsen:
                    nat.check(self.lib.pml_prolong2x(
                        nat.PML_F64, self.buf.data_ptr(), dst.data_ptr(), self.work.n_fields,
                        self.work.n0, self.work.n1, sp))
                    self._dirichlet(dst, self.fine, self.bc_fine)
                else:
                    dst.copy_(self.buf.view(-1))
                if self.fallback is not None:
                    bad = (~torch.isfinite(dst)).any() | (
                        dst.abs().max() > self.op.guard * src.abs().max().clamp_min(1e-300)
                    )
                    if bool(bad.item()):
                        self.op.fallbacks += 1
                        self.fallback.run(u_f64[k : k + 1], first + k, first + k + 1,
                                          out_f64[k : k + 1])


class CorrectedCoarseOperator:
    """Parareal coarse operator: a cheap FDM solve on the 2x-coarsened grid
    corrected by one surrogate application per slice, on the 2x grid
    (``correct_on = "coarse"``) or on the problem's own grid after the
    prolongation (``correct_on = "fine"``, where the network also supplies
    the detail that the 2x grid cannot carry).

    Per slice: restrict the state to the 2x grid (and, with
    ``base_levels = 2``, again to the 4x grid), advance it over the slice with
    ``base`` (an ``FDMOperator`` with a step that is stable on its grid but
    large, so the solve is cheap and inaccurate), prolong the result back to
    the 2x grid, apply the
    surrogate once (``u + s net(u / s)``, trained by ``corrected_pairs`` to
    map the base prediction to the fine solution restricted to the 2x grid)
    and prolong to the fine grid; the Dirichlet values are imposed on every
    grid.  The base solve is contractive, so the network only has to learn a
    small, local correction and repeated application stays bounded.  With
    ``base_levels = 1`` the base sees the whole 2x-grid state, so the
    operator's response to perturbations (what sets Parareal's convergence)
    keeps their 2x-grid content; a 4x base discards it.  A
    non-finite slice, or one whose magnitude exceeds ``guard`` times its
    input's, is recomputed by ``fallback`` and counted in ``fallbacks``.

    With ``freeze`` (the default) each slice's correction (the corrected
    minus the base prediction on the 2x grid) is evaluated once, at the
    slice's first coarse evaluation (Parareal's initial sweep), and reused:
    later evaluations return the base prediction plus that fixed field.  The
    operator is then affine per slice and the correction cancels in
    Parareal's G(U_{k+1}) - G(U_k), so the iteration contracts like the base
    solve from the network-corrected initial guess.  A live BF16 network in
    every evaluation does not cancel below its rounding level (about 1e-5
    relative here) and stalls the iteration there."""

    def __init__(self, base, surrogate: ConvSurrogate, fallback=None, guard: float = 10.0,
                 base_levels: int = 2, freeze: bool = True, correct_on: str = "coarse"):
        if base_levels not in (1, 2):
            raise ValueError("the base solve runs on the 2x (1) or 4x (2) coarsened grid")
        if correct_on not in ("coarse", "fine"):
            raise ValueError("the correction acts on the 2x grid ('coarse') or the problem's "
                             "grid ('fine')")
        self.correct_on = correct_on
        self.base = base
        self.surrogate = surrogate
        self.fallback = fallback
        self.guard = guard
        self.base_levels = base_levels
        self.freeze = freeze
        self.fallbacks = 0

    @property
    def d_t(self) -> float:
        return self.base.d_t

    @property
    def vertex_oriented(self) -> Optional[bool]:
        return True

    def slice_propagator(self, cp, n_slices, slice_steps, stream, record_trajectory=False,
                         borders=None):
        fb = None
        if self.fallback is not None:
            fb = self.fallback.slice_propagator(
                cp, n_slices, int(round(slice_steps * self.d_t / self.fallback.d_t)), stream
            )
        return CorrectedPropagator(self, cp, n_slices, slice_steps, stream, fb)


class CorrectedPropagator:
    def __init__(self, op: CorrectedCoarseOperator, cp, n_slices: int, steps: int, stream,
                 fallback_prop):
        import torch

        from pararealml_amd.operators.parareal.propagators import FDMSlicePropagator

        self.op = op
        self.stream = stream
        self.fallback = fallback_prop
        self.torch = torch
        self.fine = dev.device_problem(cp)
        work_cp = coarse_problem(cp)
        self.work = dev.device_problem(work_cp)
        self.base_cp = coarse_problem(work_cp) if op.base_levels == 2 else work_cp
        self.basep = dev.device_problem(self.base_cp)
        self.inner = FDMSlicePropagator(op.base, self.base_cp, 1, steps, stream)
        self.bc_fine = torch.from_numpy(self.fine.bc).to("cuda")
        self.bc_work = torch.from_numpy(self.work.bc).to("cuda")
        f64 = dict(dtype=torch.float64, device="cuda")
        self.buf = torch.empty((1, self.work.n_fields, self.work.n0, self.work.n1), **f64)
        self.b_in = torch.empty((1, self.basep.elems), **f64)
        self.b_out = torch.empty((1, self.basep.elems), **f64)
        self.lib = nat.device_lib()
        # frozen per-slice corrections on the corrected grid (see
        # CorrectedCoarseOperator)
        self.on_fine = op.correct_on == "fine"
        elems = self.fine.elems if self.on_fine else self.work.elems
        self.corr = self.base_pred = None
        if op.freeze:
            self.corr = torch.zeros((n_slices, elems), **f64)
            self.base_pred = torch.empty(elems, **f64)
            self.have = np.zeros(n_slices, dtype=bool)

    def _dirichlet(self, t, p, bc):
        nat.check(self.lib.pml_apply_dirichlet(
            t.data_ptr(), bc.data_ptr(), p.n_fields, p.n0, p.n1, p.ndim, 1,
            ctypes.c_void_p(self.stream.cuda_stream)))

    def _correct(self, x, n: int, p, bc):
        """Applies slice n's correction to x ((1, C, H, W), FP64, Dirichlet
        values imposed): the network, or its frozen result."""
        if self.corr is None:
            self.op.surrogate.predict(x, out=x, stream=self.stream)
            self._dirichlet(x, p, bc)
        elif not self.have[n]:
            self.base_pred.copy_(x.view(-1))
            self.op.surrogate.predict(x, out=x, stream=self.stream)
            self._dirichlet(x, p, bc)
            self.torch.sub(x.view(-1), self.base_pred, out=self.corr[n])
            self.have[n] = True
        else:
            x.view(-1).add_(self.corr[n])

    def run(self, u_f64, first: int, last: int, out_f64):
        torch = self.torch
        sp = ctypes.c_void_p(self.stream.cuda_stream)
        w, b = self.work, self.basep
        # every torch operation below (the fallback test) runs on the
        # propagator's stream, in order with its kernels
        with torch.cuda.stream(self.stream):
            for k in range(last - first):
                src = u_f64[k]
                nat.check(self.lib.pml_restrict2x(nat.PML_F64, src.data_ptr(), self.buf.data_ptr(),
                                                  w.n_fields, w.n0, w.n1, sp))
                self._dirichlet(self.buf, w, self.bc_work)
                if self.op.base_levels == 2:
                    nat.check(self.lib.pml_restrict2x(nat.PML_F64, self.buf.data_ptr(),
                                                      self.b_in.data_ptr(), b.n_fields, b.n0, b.n1,
                                                      sp))
                    self.inner.run(self.b_in, 0, 1, self.b_out)
                    nat.check(self.lib.pml_prolong2x(nat.PML_F64, self.b_out.data_ptr(),
                                                     self.buf.data_ptr(), b.n_fields, b.n0, b.n1, sp))
                else:
                    self.inner.run(self.buf.view(1, -1), 0, 1, self.buf.view(1, -1))
                self._dirichlet(self.buf, w, self.bc_work)
                dst = out_f64[k]
                if not self.on_fine:
                    self._correct(self.buf, first + k, w, self.bc_work)
                nat.check(self.lib.pml_prolong2x(nat.PML_F64, self.buf.data_ptr(), dst.data_ptr(),
                                                 w.n_fields, w.n0, w.n1, sp))
                self._dirichlet(dst, self.fine, self.bc_fine)
                if self.on_fine:
                    fp = self.fine
                    self._correct(dst.view(1, fp.n_fields, fp.n0, fp.n1), first + k, fp,
                                  self.bc_fine)
                if self.fallback is not None:
                    bad = (~torch.isfinite(dst)).any() | (
                        dst.abs().max() > self.op.guard * src.abs().max().clamp_min(1e-300)
                    )
                    if bool(bad.item()):
                        self.op.fallbacks += 1
                        self.fallback.run(u_f64[k : k + 1], first + k, first + k + 1,
                                          out_f64[k : k + 1])


def corrected_pairs(base_op, cp, snaps, slice_dt: float, stream=None, base_levels: int = 2,
                    correct_on: str = "coarse"):
    """Training pairs for ``CorrectedCoarseOperator`` from trajectories
    (``snaps``: FP32 device tensor (T, N, C, H, W), states one slice apart,
    as ``generate_pairs`` returns them: on the 2x-coarsened grid for
    ``correct_on = "coarse"``, on the problem's grid for ``"fine"``).  The
    input is the base operator's prediction over one slice from state t (run
    on the 2x grid, or restricted to the 4x grid and prolonged back for
    ``base_levels = 2``; for ``"fine"`` the state is first restricted to the
    2x grid and the prediction prolonged to the problem's grid; Dirichlet
    values imposed on every grid), the target is state t + 1.  Returns
    (inputs, targets), each ((T-1) N, C, H, W) FP32."""
    import torch

    stream = stream if stream is not None else dev.solver_stream()
    n_t, n, c, h, w = snaps.shape
    on_fine = correct_on == "fine"
    work_cp = coarse_problem(cp)
    wp = dev.device_problem(work_cp)
    fp = dev.device_problem(cp)
    base_cp = coarse_problem(work_cp) if base_levels == 2 else work_cp
    bp = base_op.device_problem(base_cp)
    steps = int(round(slice_dt / base_op.d_t))
    lib = nat.device_lib()
    sp = ctypes.c_void_p(stream.cuda_stream)
    inputs = torch.empty((n_t - 1, n, c, h, w), dtype=torch.float32, device="cuda")
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        batch = dev.FDMBatch([bp], [0] * n, base_op.precision,
                             dev.integrator_id(base_op.integrator), stream)
        bc_work = torch.from_numpy(wp.bc).to("cuda")
        bc_fine = torch.from_numpy(fp.bc).to("cuda")
        x4 = torch.empty(n * bp.elems, dtype=torch.float64, device="cuda")
        y2 = torch.empty((n * c, wp.n0, wp.n1), dtype=torch.float64, device="cuda")
        yf = torch.empty((n * c, fp.n0, fp.n1), dtype=torch.float64, device="cuda")
        for t in range(n_t - 1):
            x2 = snaps[t].to(torch.float64).contiguous()
            if on_fine:
                nat.check(lib.pml_restrict2x(nat.PML_F64, x2.data_ptr(), y2.data_ptr(), n * c,
                                             wp.n0, wp.n1, sp))
                x2 = y2.clone()
            if base_levels == 2:
                nat.check(lib.pml_restrict2x(nat.PML_F64, x2.data_ptr(), x4.data_ptr(), n * c,
                                             bp.n0, bp.n1, sp))
                batch.load_f64(x4)
            else:
                batch.load_f64(x2.view(-1))
            batch.advance(base_op.d_t, steps, tile_config=base_op.tile_config_for(bp, n),
                          use_graph=True, constrain_input=True)
            y4 = batch.state_f64()
            if base_levels == 2:
                nat.check(lib.pml_prolong2x(nat.PML_F64, y4.data_ptr(), y2.data_ptr(), n * c,
                                            bp.n0, bp.n1, sp))
            else:
                y2.view(-1).copy_(y4)
            nat.check(lib.pml_apply_dirichlet(y2.data_ptr(), bc_work.data_ptr(), c, wp.n0,
                                              wp.n1, wp.ndim, n, sp))
            if on_fine:
                nat.check(lib.pml_prolong2x(nat.PML_F64, y2.data_ptr(), yf.data_ptr(), n * c,
                                            wp.n0, wp.n1, sp))
                nat.check(lib.pml_apply_dirichlet(yf.data_ptr(), bc_fine.data_ptr(), c, fp.n0,
                                                  fp.n1, fp.ndim, n, sp))
                inputs[t].copy_(yf.view(n, c, h, w))
            else:
                inputs[t].copy_(y2.view(n, c, h, w))
    torch.cuda.current_stream().wait_stream(stream)
    return inputs.reshape(-1, c, h, w), snaps[1:].reshape(-1, c, h, w)


def gaussian_mixture_states(cp, n: int, seed: int, components: int = 4,
                            mean_range=(0.2, 0.8), var_range=(0.005, 0.02),
                            noise: float = 0.0) -> np.ndarray:
    """Seeded sums of ``components`` normalised isotropic Gaussian densities
    (random means and variances) on the problem's vertices, plus white noise
    of standard deviation ``noise`` times each state's maximum magnitude,
    with the static Dirichlet values imposed; shape (n, *vertices_shape,
    y_dim)."""
    from pararealml_amd.constraint import apply_constraints_along_last_axis

    rng = np.random.RandomState(seed)
    grids = cp.mesh.vertex_coordinate_grids
    lo = [iv[0] for iv in cp.mesh.x_intervals]
    span = [iv[1] - iv[0] for iv in cp.mesh.x_intervals]
    y_dim = cp.differential_equation.y_dimension
    out = np.zeros((n,) + cp.y_vertices_shape)
    for k in range(n):
        for f in range(y_dim):
            acc = np.zeros(cp.mesh.vertices_shape)
            for _ in range(components):
                mean = [lo[a] + span[a] * rng.uniform(*mean_range) for a in range(2)]
                var = rng.uniform(*var_range) * span[0] * span[1]
                r2 = sum((grids[a] - mean[a]) ** 2 for a in range(2))
                acc += np.exp(-r2 / (2.0 * var)) / (2.0 * np.pi * var)
            out[k, ..., f] = acc
        if noise > 0.0:
            out[k] += noise * np.abs(out[k]).max() * rng.standard_normal(out[k].shape)
        apply_constraints_along_last_axis(cp.static_y_vertex_constraints, out[k])
    return out


def generate_pairs(fdm_op, cp, y0s: np.ndarray, t_span: float, pair_dt: float,
                   coarsen: bool = True, stream=None):
    """Trajectories of ``fdm_op`` (an FDMOperator; all initial states in one
    batch) sampled every ``pair_dt`` over ``t_span``, restricted to the
    working grid.  Returns an FP32 device tensor (n_snap, n, C, H, W),
    n_snap = t_span / pair_dt + 1 (the initial states included)."""
    import torch

    stream = stream if stream is not None else dev.solver_stream()
    problem = fdm_op.device_problem(cp)
    n = len(y0s)
    steps = int(round(pair_dt / fdm_op.d_t))
    n_pairs = int(round(t_span / pair_dt))
    wp = dev.device_problem(coarse_problem(cp)) if coarsen else problem
    out = torch.empty((n_pairs + 1, n, wp.n_fields, wp.n0, wp.n1), dtype=torch.float32,
                      device="cuda")
    lib = nat.device_lib()
    sp = ctypes.c_void_p(stream.cuda_stream)
    with torch.cuda.stream(stream):
        batch = dev.FDMBatch([problem], [0] * n, fdm_op.precision,
                             dev.integrator_id(fdm_op.integrator), stream)
        batch.load_host(list(y0s))
        f64 = torch.empty(batch.total, dtype=torch.float64, device="cuda")
        tmp = torch.empty((n * wp.n_fields, wp.n0, wp.n1), dtype=torch.float64, device="cuda")
        for k in range(n_pairs + 1):
            if k:
                batch.advance(fdm_op.d_t, steps, tile_config=fdm_op.tile_config_for(problem, n),
                              use_graph=True)
            batch.state_f64(f64)
            if coarsen:
                nat.check(lib.pml_restrict2x(nat.PML_F64, f64.data_ptr(), tmp.data_ptr(),
                                             n * wp.n_fields, wp.n0, wp.n1, sp))
                out[k].copy_(tmp.view(n, wp.n_fields, wp.n0, wp.n1))
            else:
                out[k].copy_(f64.view(n, wp.n_fields, wp.n0, wp.n1))
    stream.synchronize()
    return out
