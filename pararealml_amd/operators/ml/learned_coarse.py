"""The convolutional surrogate as a Parareal coarse propagator, its training
data generation from the device FDM solver, and the FDM fallback.

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

    def slice_propagator(self, cp, n_slices, slice_steps, stream, record_trajectory=False):
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
                net.predict(self.buf, out=self.buf)
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
