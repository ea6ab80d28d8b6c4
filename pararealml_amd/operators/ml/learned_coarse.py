"""The convolutional surrogate as a Parareal coarse propagator, its training
data generation from the device FDM solver, and the FDM fallback.

Two forms: ``LearnedCoarseOperator`` replaces the coarse solve by repeated
surrogate steps; ``CorrectedCoarseOperator`` runs a cheap FDM solve on a
coarsened grid and applies the surrogate once per slice as a learned
correction (stable under the Parareal sweep's repeated application).

``LearnedCoarseOperator(surrogate, d_t, coarsen=True, fallback=None)`` steps
the state by ``d_t`` per network application (the reference's learned
operators step autoregressively by their ``d_t``,
pararealml/operators/ml/supervised/supervised_ml_operator.py:91-128).  With
``coarsen`` the network works on the 2x-coarsened vertex grid.  After every
application the problem's Dirichlet values are imposed.  When the output of
a slice is not finite or its magnitude exceeds ``guard`` times the input's,
that slice is recomputed by the ``fallback`` operator (normally the plain FDM
coarse operator) and counted in ``fallbacks``.

Grid transfers (``GridLevels``): injection down, cubic (default) or
bilinear prolongation up, Dirichlet values imposed on every grid.
"""

import ctypes
from typing import Optional, Union

import numpy as np

from pararealml_amd import _device as dev
from pararealml_amd import _native as nat
from pararealml_amd.operators.ml.conv_surrogate import ConvSurrogate
from pararealml_amd.operators.parareal.propagators import (
    FDMSlicePropagator,
    GridLevels,
    coarse_problem,
)

__all__ = ["coarse_problem", "LearnedCoarseOperator", "CorrectedCoarseOperator",
           "corrected_pairs", "gaussian_mixture_states", "generate_pairs"]


class _FallbackGuard:
    """The fallback test of a learned propagator: a slice whose prediction
    is not finite, or whose magnitude exceeds ``op.guard`` times its
    input's, is recomputed by the fallback propagator (and counted in
    ``op.fallbacks``).

    By default each slice is tested as it is predicted, which costs one host
    synchronisation per slice.  With ``deferred`` set (the Parareal driver
    sets it for its coarse sweeps) a test only records a device flag;
    ``first_bad()`` reads the flags once per sweep, and the driver re-runs
    the sweep from the first flagged slice after ``force_fallback`` on it,
    which gives the result of testing every slice as it is made."""

    supports_deferral = True

    def _guard_init(self, n_slices: int) -> None:
        self.deferred = False
        self._n_slices = n_slices
        self._flags = None  # int32 per slice, set by pml_slice_guard
        self._one = None    # int32, the immediate test's flag
        self._guard_work = None
        self._forced = set()

    def _run_fallback(self, u_f64, out_f64, k: int, first: int) -> None:
        self.op.fallbacks += 1
        self.fallback.run(u_f64[k : k + 1], first + k, first + k + 1, out_f64[k : k + 1])

    def _guard(self, u_f64, out_f64, k: int, first: int) -> None:
        """Runs on the propagator's stream (the caller's context): one
        ``pml_slice_guard`` (non-finite prediction, or magnitude above
        ``op.guard`` times the input's) per slice."""
        if self.fallback is None:
            return
        torch = self.torch
        if self._guard_work is None:
            self._guard_work = torch.empty(4 * nat.PML_REDUCE_SLOTS, dtype=torch.float64,
                                           device="cuda")
            self._flags = torch.zeros(self._n_slices, dtype=torch.int32, device="cuda")
            self._one = torch.zeros(1, dtype=torch.int32, device="cuda")
        flag = self._flags[first + k] if self.deferred else self._one[0]
        if not self.deferred:
            flag.zero_()
        self._test(u_f64, out_f64, k, first, flag)
        if not self.deferred and int(flag.item()):
            self._run_fallback(u_f64, out_f64, k, first)

    def _test(self, u_f64, out_f64, k: int, first: int, flag) -> None:
        """Sets ``flag`` (a device int32) when slice first + k's prediction
        needs the fallback."""
        nat.check(nat.device_lib().pml_slice_guard(
            u_f64[k].data_ptr(), out_f64[k].data_ptr(), u_f64[k].numel(),
            float(self.op.guard), self._guard_work.data_ptr(), flag.data_ptr(),
            ctypes.c_void_p(self.stream.cuda_stream)))

    def first_bad(self):
        """Index of the first slice flagged since the last call (None if
        none); clears the flags.  One host synchronisation."""
        if self._flags is None:
            return None
        with self.torch.cuda.stream(self.stream):
            host = self._flags.cpu().numpy()
            self._flags.zero_()
        idx = np.flatnonzero(host)
        return int(idx[0]) if idx.size else None

    def force_fallback(self, j: int) -> None:
        """Slice j is computed by the fallback until ``clear_forced``."""
        self._forced.add(int(j))

    def clear_forced(self) -> None:
        self._forced.clear()


class LearnedCoarseOperator:
    """Parareal coarse operator backed by a ``ConvSurrogate``."""

    def __init__(self, surrogate: ConvSurrogate, d_t: float, coarsen: bool = True,
                 fallback=None, guard: float = 10.0, prolongation: str = "cubic"):
        if d_t <= 0.0:
            raise ValueError("time step size must be greater than 0")
        if prolongation not in ("cubic", "bilinear"):
            raise ValueError("prolongation is 'cubic' or 'bilinear'")
        self.surrogate = surrogate
        self._d_t = d_t
        self.coarsen = coarsen
        self.fallback = fallback
        self.guard = guard
        self.prolongation = prolongation
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
        return LearnedPropagator(self, cp, slice_steps, stream, fb, n_slices)


class LearnedPropagator(_FallbackGuard):
    def __init__(self, op: LearnedCoarseOperator, cp, steps: int, stream, fallback_prop,
                 n_slices: int = 1):
        import torch

        self.op = op
        self.steps = steps
        self.stream = stream
        self.fallback = fallback_prop
        self.torch = torch
        self.level = 1 if op.coarsen else 0
        self.grids = GridLevels(cp, self.level, op.prolongation)
        self.work = self.grids.probs[self.level]
        self.buf = torch.empty(
            (1, self.work.n_fields, self.work.n0, self.work.n1), dtype=torch.float64, device="cuda"
        )
        self._guard_init(n_slices)

    def run(self, u_f64, first: int, last: int, out_f64):
        torch = self.torch
        sp = ctypes.c_void_p(self.stream.cuda_stream)
        net = self.op.surrogate
        g = self.grids
        # every torch operation below (copies, the fallback test) runs on
        # the propagator's stream, in order with its kernels
        with torch.cuda.stream(self.stream):
            for k in range(last - first):
                if first + k in self._forced:
                    self._run_fallback(u_f64, out_f64, k, first)
                    continue
                if self.level:
                    g.down(u_f64[k].data_ptr(), self.buf.data_ptr(), 1, 0, self.level, sp)
                else:
                    self.buf.view(-1).copy_(u_f64[k])
                    g._dirichlet(self.buf.data_ptr(), 0, 1, sp)
                for _ in range(self.steps):
                    net.predict(self.buf, out=self.buf, stream=self.stream)
                    g._dirichlet(self.buf.data_ptr(), self.level, 1, sp)
                if self.level:
                    g.up(self.buf.data_ptr(), out_f64[k].data_ptr(), 1, 0, self.level, sp)
                else:
                    out_f64[k].copy_(self.buf.view(-1))
                self._guard(u_f64, out_f64, k, first)


def _level(correct_on: Union[str, int], base_levels: int) -> int:
    if isinstance(correct_on, str):
        try:
            lv = {"fine": 0, "coarse": 1, "base": base_levels}[correct_on]
        except KeyError:
            raise ValueError("correct_on is 'fine', 'coarse', 'base' or a grid level") from None
    else:
        lv = int(correct_on)
    if not 0 <= lv <= base_levels:
        raise ValueError(f"the correction level must lie in [0, base_levels={base_levels}]")
    return lv


class CorrectedCoarseOperator:
    """Parareal coarse operator: a cheap FDM solve on a coarsened grid
    corrected by one surrogate application per slice.

    Per slice: inject the state down ``base_levels`` 2x levels, advance it
    over the slice with ``base`` (an ``FDMOperator`` with a step that is
    stable on its grid but large, so the solve is cheap and inaccurate),
    prolong the prediction up to the correction grid (``correct_on``:
    "fine" = the problem's grid, "coarse" = the 2x grid, "base" = the base
    grid itself, or a level number), apply the surrogate once there
    (``u + s net(u / s)``, trained by ``corrected_pairs`` to map the base
    prediction to the fine solution on that grid), and prolong to the
    problem's grid; the Dirichlet values are imposed on every grid.
    ``prolongation`` is "cubic" (default; the bilinear interpolation error
    is the largest error of a coarse prediction on smooth states) or
    "bilinear".  A non-finite slice, or one whose magnitude exceeds
    ``guard`` times its input's, is recomputed by ``fallback`` and counted
    in ``fallbacks``.

    With ``freeze`` each slice's correction (the corrected minus the base
    prediction on the correction grid) is evaluated once, at the slice's
    first coarse evaluation (Parareal's initial sweep), and reused: the
    operator is then affine per slice and the correction cancels in
    Parareal's G(U_{k+1}) - G(U_k), so the iteration contracts like the base
    solve from the network-corrected initial guess.  Without it the network
    runs in every evaluation and enters the contraction itself."""

    def __init__(self, base, surrogate: ConvSurrogate, fallback=None, guard: float = 10.0,
                 base_levels: int = 2, freeze: bool = True,
                 correct_on: Union[str, int] = "coarse", prolongation: str = "cubic"):
        if not 1 <= base_levels <= 4:
            raise ValueError("the base solve runs 1 to 4 2x levels below the problem's grid")
        if prolongation not in ("cubic", "bilinear"):
            raise ValueError("prolongation is 'cubic' or 'bilinear'")
        self.correct_level = _level(correct_on, base_levels)
        self.correct_on = correct_on
        self.base = base
        self.surrogate = surrogate
        self.fallback = fallback
        self.guard = guard
        self.base_levels = base_levels
        self.freeze = freeze
        self.prolongation = prolongation
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


class CorrectedPropagator(_FallbackGuard):
    def __init__(self, op: CorrectedCoarseOperator, cp, n_slices: int, steps: int, stream,
                 fallback_prop):
        import torch

        self.op = op
        self.stream = stream
        self.fallback = fallback_prop
        self.torch = torch
        self.grids = GridLevels(cp, op.base_levels, op.prolongation)
        self.lb, self.lc = op.base_levels, op.correct_level
        self.basep = self.grids.probs[self.lb]
        self.inner = FDMSlicePropagator(op.base, self.grids.cps[self.lb], 1, steps, stream)
        cpb = self.grids.probs[self.lc]
        f64 = dict(dtype=torch.float64, device="cuda")
        self.cbuf = torch.empty((1, cpb.n_fields, cpb.n0, cpb.n1), **f64)
        self.b_in = torch.empty((1, self.basep.elems), **f64)
        self.b_out = torch.empty((1, self.basep.elems), **f64)
        # frozen per-slice corrections on the correction grid
        self.corr = self.base_pred = None
        if op.freeze:
            self.corr = torch.zeros((n_slices, cpb.elems), **f64)
            self.base_pred = torch.empty(cpb.elems, **f64)
            self.have = np.zeros(n_slices, dtype=bool)
        self._guard_init(n_slices)

    def _correct(self, x, n: int, sp):
        """Applies slice n's correction to x ((1, C, H, W) on the correction
        grid, Dirichlet values imposed): the network, or its frozen
        result."""
        g, lc = self.grids, self.lc
        if self.corr is None:
            self.op.surrogate.predict(x, out=x, stream=self.stream)
            g._dirichlet(x.data_ptr(), lc, 1, sp)
        elif not self.have[n]:
            self.base_pred.copy_(x.view(-1))
            self.op.surrogate.predict(x, out=x, stream=self.stream)
            g._dirichlet(x.data_ptr(), lc, 1, sp)
            self.torch.sub(x.view(-1), self.base_pred, out=self.corr[n])
            self.have[n] = True
        else:
            x.view(-1).add_(self.corr[n])

    def run(self, u_f64, first: int, last: int, out_f64):
        torch = self.torch
        sp = ctypes.c_void_p(self.stream.cuda_stream)
        g, lb, lc = self.grids, self.lb, self.lc
        # every torch operation below (the corrections, the fallback test)
        # runs on the propagator's stream, in order with its kernels
        with torch.cuda.stream(self.stream):
            for k in range(last - first):
                if first + k in self._forced:
                    self._run_fallback(u_f64, out_f64, k, first)
                    continue
                g.down(u_f64[k].data_ptr(), self.b_in.data_ptr(), 1, 0, lb, sp)
                self.inner.run(self.b_in, 0, 1, self.b_out)
                g._dirichlet(self.b_out.data_ptr(), lb, 1, sp)
                if lc < lb:
                    g.up(self.b_out.data_ptr(), self.cbuf.data_ptr(), 1, lc, lb, sp)
                else:
                    self.cbuf.view(-1).copy_(self.b_out.view(-1))
                self._correct(self.cbuf, first + k, sp)
                if lc > 0:
                    g.up(self.cbuf.data_ptr(), out_f64[k].data_ptr(), 1, 0, lc, sp)
                else:
                    out_f64[k].copy_(self.cbuf.view(-1))
                self._guard(u_f64, out_f64, k, first)


def corrected_pairs(base_op, cp, snaps, slice_dt: float, stream=None, base_levels: int = 2,
                    correct_on: Union[str, int] = "coarse", prolongation: str = "cubic"):
    """Training pairs for ``CorrectedCoarseOperator`` from trajectories
    (``snaps``: FP32 device tensor (T, N, C, H, W), states one slice apart on
    the correction grid, as ``generate_pairs(..., level=...)`` returns
    them).  The input is the base operator's prediction over one slice from
    state t (injected down to the base grid, stepped, prolonged back to the
    correction grid; Dirichlet values imposed on every grid), the target is
    state t + 1.  Returns (inputs, targets), each ((T-1) N, C, H, W)
    FP32."""
    import torch

    stream = stream if stream is not None else dev.solver_stream()
    n_t, n, c, h, w = snaps.shape
    lc = _level(correct_on, base_levels)
    grids = GridLevels(cp, base_levels, prolongation, batch=n)
    cpb = grids.probs[lc]
    if (cpb.n0, cpb.n1) != (h, w):
        raise ValueError("the trajectories are not on the correction grid")
    bp = base_op.device_problem(grids.cps[base_levels])
    steps = int(round(slice_dt / base_op.d_t))
    sp = ctypes.c_void_p(stream.cuda_stream)
    inputs = torch.empty((n_t - 1, n, c, h, w), dtype=torch.float32, device="cuda")
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        batch = dev.FDMBatch([bp], [0] * n, base_op.precision,
                             dev.integrator_id(base_op.integrator), stream)
        xb = torch.empty(n * bp.elems, dtype=torch.float64, device="cuda")
        yc = torch.empty((n * c, h, w), dtype=torch.float64, device="cuda")
        for t in range(n_t - 1):
            x = snaps[t].to(torch.float64).contiguous()
            if lc < base_levels:
                grids.down(x.data_ptr(), xb.data_ptr(), n, lc, base_levels, sp)
                batch.load_f64(xb)
            else:
                batch.load_f64(x.view(-1))
            batch.advance(base_op.d_t, steps, tile_config=base_op.tile_config_for(bp, n),
                          use_graph=True, constrain_input=True)
            yb = batch.state_f64()
            grids._dirichlet(yb.data_ptr(), base_levels, n, sp)
            if lc < base_levels:
                grids.up(yb.data_ptr(), yc.data_ptr(), n, lc, base_levels, sp)
            else:
                yc.view(-1).copy_(yb)
            inputs[t].copy_(yc.view(n, c, h, w))
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
                   coarsen: bool = True, stream=None, level: Optional[int] = None):
    """Trajectories of ``fdm_op`` (an FDMOperator; all initial states in one
    batch) sampled every ``pair_dt`` over ``t_span``, injected down to grid
    ``level`` (default: 1 with ``coarsen``, else 0).  Returns an FP32 device
    tensor (n_snap, n, C, H, W), n_snap = t_span / pair_dt + 1 (the initial
    states included)."""
    import torch

    stream = stream if stream is not None else dev.solver_stream()
    problem = fdm_op.device_problem(cp)
    lv = (1 if coarsen else 0) if level is None else int(level)
    n = len(y0s)
    steps = int(round(pair_dt / fdm_op.d_t))
    n_pairs = int(round(t_span / pair_dt))
    grids = GridLevels(cp, lv, batch=n) if lv else None
    wp = grids.probs[lv] if lv else problem
    out = torch.empty((n_pairs + 1, n, wp.n_fields, wp.n0, wp.n1), dtype=torch.float32,
                      device="cuda")
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
            if lv:
                grids.down(f64.data_ptr(), tmp.data_ptr(), n, 0, lv, sp)
                out[k].copy_(tmp.view(n, wp.n_fields, wp.n0, wp.n1))
            else:
                out[k].copy_(f64.view(n, wp.n_fields, wp.n0, wp.n1))
    stream.synchronize()
    return out
