"""Device slice propagators used by the one-GPU Parareal driver.

A propagator advances initial values given as FP64 device tensors in the
fine problem's field-major layout over whole time slices, for a range of
slices at once:

* ``FDMSlicePropagator`` -- an ``FDMOperator`` on the problem's own grid;
  all requested slices form one batch advanced by a single
  ``pml_fdm_solve`` call (the fine solves of every slice in one launch per
  step).  Optionally records every step's state for the trajectory output.
* ``CoarseGridPropagator`` -- an ``FDMOperator`` on a 2x-coarsened grid
  (one or more levels): injection to the coarse grid, stepping there,
  cubic (or bilinear) prolongation back (``pml_restrict2x`` /
  ``pml_prolong2x_cubic`` / ``pml_prolong2x``).  The grid transfer has no
  counterpart in the reference; injection followed by prolongation is exact
  on the coarse points, so a chain of per-slice coarse solves equals one
  continuous coarse solve, as the reference's initial coarse sweep assumes.
  The cubic prolongation matters for Parareal: the bilinear one's
  interpolation error is the largest part of the coarse operator's error
  on smooth states and costs iterations.

* ``ODESlicePropagator`` -- an ``FDMOperator`` on a system without a mesh:
  all requested slices advance together in one launch of the ODE kernel
  (``pml_ode_solve``, one thread per slice, the whole slice in the launch),
  each from its own start time.

Every slice's initial value has the problem's static Dirichlet values
imposed before its first step, as ``DiscreteInitialCondition`` does for
each sub-problem of the reference's Parareal (parareal_operator.py:153-163).
"""

import ctypes
from typing import Optional

import numpy as np

from pararealml_amd import _device as dev
from pararealml_amd import _native as nat


class FDMSlicePropagator:
    """Slices of one FDM problem, batched."""

    def __init__(self, op, cp, n_slices: int, slice_steps: int, stream,
                 record_trajectory: bool = False):
        import torch

        self.op = op
        self.problem = op.device_problem(cp)
        self.n_slices = n_slices
        self.steps = slice_steps
        self.stream = stream
        self.batch = dev.FDMBatch(
            [self.problem], [0] * n_slices, op.precision, dev.integrator_id(op.integrator),
            stream,
        )
        self.elems = self.problem.elems
        self.snaps = None
        if record_trajectory:
            self.snaps = self.batch.new_snapshots(slice_steps)
        self._esz = 8 if self.batch.prec == nat.PML_F64 else 4
        self.torch = torch

    def _sptr(self):
        return ctypes.c_void_p(self.stream.cuda_stream)

    def run(self, u_f64, first: int, last: int, out_f64):
        """Advances slices [first, last) from ``u_f64[k]`` (k = 0 .. last -
        first - 1, FP64, shape (count, elems)) and writes the end states to
        ``out_f64`` (same shape)."""
        count = last - first
        if count <= 0:
            return
        b = self.batch
        off = first * self.elems
        lib = b.lib
        hi = b.hi.data_ptr() + off * self._esz
        lo = None if b.lo is None else b.lo.data_ptr() + off * self._esz
        nat.check(lib.pml_from_f64(b.prec, u_f64.data_ptr(), hi, lo, count * self.elems,
                                   self._sptr()))
        b.advance(
            self.op.d_t,
            self.steps,
            snapshots=self.snaps,
            snapshot_every=1 if self.snaps is not None else 0,
            tile_config=self.op.tile_config_for(self.problem, count),
            use_graph=True,
            constrain_input=True,
            first_grid=first,
            n_grids=count,
        )
        nat.check(lib.pml_to_f64(b.prec, hi, lo, out_f64.data_ptr(), count * self.elems,
                                 self._sptr()))

    def trajectory_f64(self):
        """Recorded states, FP64, shape (steps, n_slices, elems)."""
        if self.snaps is None:
            raise RuntimeError("trajectory was not recorded")
        flat = self.batch.snapshots_f64(self.snaps, self.steps)
        return flat.view(self.steps, self.n_slices, self.elems)


class ODEProblem:
    """The device-layout description of an ODE state (the interface
    ``PararealOperator`` uses for grids): ``n`` fields of one point."""

    def __init__(self, n: int):
        self.n_fields = n
        self.n0 = self.n1 = 1

    @property
    def plane(self) -> int:
        return 1

    @property
    def elems(self) -> int:
        return self.n_fields

    def to_device_layout(self, y) -> np.ndarray:
        return np.asarray(y, dtype=np.float64).reshape(self.n_fields, 1, 1)

    def from_device_layout(self, y, spatial_shape) -> np.ndarray:
        lead = y.shape[:-3]
        return np.ascontiguousarray(y.reshape(lead + (self.n_fields,)))


class ODESlicePropagator:
    """Slices of an ODE system, all advanced by one kernel launch.  Slice j
    starts at ``borders[j]`` and takes ``slice_steps`` steps on the
    reference's time grid of the sub-problem, linspace(T_j, T_j + n d_t,
    n + 1) (the time of sub-problem solves in parareal_operator.py:
    153-163)."""

    def __init__(self, op, cp, n_slices: int, slice_steps: int, stream,
                 record_trajectory: bool = False, borders=None):
        import torch

        from pararealml_amd.operators.fdm.fdm_operator import ode_time_step

        if borders is None or len(borders) != n_slices + 1:
            raise ValueError("an ODE slice propagator needs the n_slices + 1 slice borders")
        self.op = op
        self.prec = op._ode_precision()
        self.integ = dev.integrator_id(op.integrator)
        self.program = op._ode_program(cp.differential_equation)
        n = cp.differential_equation.y_dimension
        self.problem = ODEProblem(n)
        self.elems = n
        self.n_slices = n_slices
        self.steps = slice_steps
        self.stream = stream
        self.dtype = torch.float64 if self.prec == nat.PML_F64 else torch.float32
        f64 = dict(dtype=torch.float64, device="cuda")
        self.t0 = torch.tensor([float(b) for b in borders[:-1]], **f64)
        self.t_step = torch.tensor(
            [ode_time_step(float(b), slice_steps, op.d_t) for b in borders[:-1]], **f64)
        self.y = torch.empty((n_slices, n), dtype=self.dtype, device="cuda")
        self.snaps = None
        if record_trajectory:
            self.snaps = torch.empty((slice_steps, n_slices, n), dtype=self.dtype, device="cuda")

    def run(self, u_f64, first: int, last: int, out_f64):
        """Advances slices [first, last) from ``u_f64`` (FP64, (count, n))
        into ``out_f64``."""
        import torch

        count = last - first
        if count <= 0:
            return
        n = self.elems
        y = self.y[first:last]
        y.copy_(u_f64.reshape(count, n))
        tmp = None
        if self.snaps is not None:
            tmp = torch.empty((self.steps, count, n), dtype=self.dtype, device="cuda")
        lib = nat.device_lib()
        nat.check(lib.pml_ode_solve(
            self.prec, self.program.handle(), self.integ, count,
            self.t0.data_ptr() + 8 * first, self.t_step.data_ptr() + 8 * first,
            float(self.op.d_t), int(self.steps), y.data_ptr(),
            None if tmp is None else tmp.data_ptr(), 1, self.stream.cuda_stream))
        if tmp is not None:
            self.snaps[:, first:last].copy_(tmp)
        out_f64.reshape(count, n).copy_(y)

    def trajectory_f64(self):
        """Recorded states, FP64, shape (steps, n_slices, n)."""
        if self.snaps is None:
            raise RuntimeError("trajectory was not recorded")
        return self.snaps.double()


def coarse_problem(cp, levels: int = 1):
    """The constrained problem on the vertex grid coarsened ``levels`` times
    by 2 (every other vertex kept, so each axis needs 2m - 1 vertices at
    every level)."""
    from pararealml_amd.constrained_problem import ConstrainedProblem
    from pararealml_amd.mesh import Mesh

    for _ in range(levels):
        mesh = cp.mesh
        if mesh.dimensions != 2:
            raise NotImplementedError("grid coarsening covers 2-D problems")
        if any((n - 1) % 2 or n < 5 for n in mesh.vertices_shape):
            raise ValueError(
                "coarsening needs an odd number of vertices (2m - 1, m >= 3) along every axis"
            )
        cmesh = Mesh(mesh.x_intervals, [2.0 * d for d in mesh.d_x], mesh.coordinate_system_type)
        cp = ConstrainedProblem(cp.differential_equation, cmesh, cp.boundary_conditions)
    return cp


class GridLevels:
    """A problem's vertex grid (level 0) and its 2x coarsenings down to
    ``levels``, with the transfers between them on FP64 device states:
    injection downwards, bilinear or cubic (``pml_prolong2x_cubic``, where
    the coarser grid has at least 4 vertices per axis) prolongation upwards,
    the problem's Dirichlet values imposed on every grid a transfer lands
    on.  Holds one FP64 buffer per intermediate level for ``batch``
    states."""

    def __init__(self, cp, levels: int, prolongation: str = "cubic", batch: int = 1):
        import torch

        if prolongation not in ("cubic", "bilinear"):
            raise ValueError("prolongation is 'cubic' or 'bilinear'")
        self.prolongation = prolongation
        self.levels = levels
        self.batch = batch
        self.cps = [cp]
        for _ in range(levels):
            self.cps.append(coarse_problem(self.cps[-1]))
        self.probs = [dev.device_problem(c) for c in self.cps]
        self.bcs = [torch.from_numpy(p.bc).to("cuda") for p in self.probs]
        self.bufs = [torch.empty(batch * p.elems, dtype=torch.float64, device="cuda")
                     for p in self.probs]
        self.lib = nat.device_lib()

    def _dirichlet(self, t_ptr, level, count, sp):
        p = self.probs[level]
        nat.check(self.lib.pml_apply_dirichlet(t_ptr, self.bcs[level].data_ptr(), p.n_fields,
                                               p.n0, p.n1, p.ndim, count, sp))

    def down(self, src_ptr, dst_ptr, count: int, lo: int, hi: int, sp) -> None:
        """Injection of ``count`` states from level ``lo`` to level ``hi``
        (> lo)."""
        if count > self.batch:
            raise ValueError("more states than the transfer buffers hold")
        cur = src_ptr
        for lv in range(lo + 1, hi + 1):
            p = self.probs[lv]
            dst = dst_ptr if lv == hi else self.bufs[lv].data_ptr()
            nat.check(self.lib.pml_restrict2x(nat.PML_F64, cur, dst, count * p.n_fields, p.n0,
                                              p.n1, sp))
            self._dirichlet(dst, lv, count, sp)
            cur = dst

    def up(self, src_ptr, dst_ptr, count: int, lo: int, hi: int, sp) -> None:
        """Prolongation of ``count`` states from level ``hi`` to level ``lo``
        (< hi)."""
        if count > self.batch:
            raise ValueError("more states than the transfer buffers hold")
        cur = src_ptr
        for lv in range(hi, lo, -1):
            p = self.probs[lv]
            dst = dst_ptr if lv - 1 == lo else self.bufs[lv - 1].data_ptr()
            cubic = self.prolongation == "cubic" and p.n0 >= 4 and p.n1 >= 4
            fn = self.lib.pml_prolong2x_cubic if cubic else self.lib.pml_prolong2x
            nat.check(fn(nat.PML_F64, cur, dst, count * p.n_fields, p.n0, p.n1, sp))
            self._dirichlet(dst, lv - 1, count, sp)
            cur = dst


class CoarseGridPropagator:
    """An FDM operator on the ``levels`` times 2x-coarsened vertex grid of
    the problem: each slice's initial value is injected down, stepped there
    and prolonged back (cubic by default)."""

    def __init__(self, op, cp, n_slices: int, slice_steps: int, stream, levels: int = 1,
                 prolongation: str = "cubic"):
        import torch

        self.grids = GridLevels(cp, levels, prolongation)
        self.levels = levels
        self.coarse_cp = self.grids.cps[-1]
        self.inner = FDMSlicePropagator(op, self.coarse_cp, 1, slice_steps, stream)
        self.stream = stream
        self.c_in = torch.empty(self.inner.elems, dtype=torch.float64, device="cuda")
        self.c_out = torch.empty(self.inner.elems, dtype=torch.float64, device="cuda")
        self.elems = self.grids.probs[0].elems

    def run(self, u_f64, first: int, last: int, out_f64):
        sp = ctypes.c_void_p(self.stream.cuda_stream)
        for k in range(last - first):
            self.grids.down(u_f64[k].data_ptr(), self.c_in.data_ptr(), 1, 0, self.levels, sp)
            self.inner.run(self.c_in.view(1, -1), 0, 1, self.c_out.view(1, -1))
            self.grids.up(self.c_out.data_ptr(), out_f64[k].data_ptr(), 1, 0, self.levels, sp)


class CoarseGridFDMOperator:
    """Marks an ``FDMOperator`` to be run on a coarsened grid when used as a
    Parareal coarse propagator (a capability the reference does not have:
    its coarse operators share the fine grid): ``levels`` 2x coarsenings
    (injection) and back by ``prolongation`` ("cubic", the default, or
    "bilinear")."""

    def __init__(self, fdm_operator, levels: int = 1, prolongation: str = "cubic"):
        if levels < 1:
            raise ValueError("levels must be at least 1")
        if prolongation not in ("cubic", "bilinear"):
            raise ValueError("prolongation is 'cubic' or 'bilinear'")
        self.fdm = fdm_operator
        self.levels = levels
        self.prolongation = prolongation

    @property
    def d_t(self) -> float:
        return self.fdm.d_t

    @property
    def vertex_oriented(self) -> Optional[bool]:
        return True

    def slice_propagator(self, cp, n_slices, slice_steps, stream, record_trajectory=False,
                         borders=None):
        return CoarseGridPropagator(self.fdm, cp, n_slices, slice_steps, stream, self.levels,
                                    self.prolongation)
