"""Device-side plumbing between the problem model and ``libpml.so``.

* ``DeviceProblem``: what the fused kernels need to know about one
  constrained problem -- equation family and coefficients, vertex counts,
  spacings and the FP64 boundary edge values in the ``pml_grid`` layout of
  include/pml.h.
* ``FDMBatch``: a (possibly ragged) batch of grids of one equation held in
  HBM in field-major SoA layout, advanced in place by ``pml_fdm_solve``.

The reference layout is channel-last ``(*vertices_shape, y_dim)``; the device
layout is ``[field][n0][n1]`` per grid with axis 1 contiguous, and a 1-D
problem is laid out as ``n0 = 1`` with the line along axis 1.
"""

from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import ctypes
import numpy as np

from pararealml_amd import _native as nat
from pararealml_amd.mesh import CoordinateSystem

PRECISIONS = {
    "float64": nat.PML_F64,
    "float32": nat.PML_F32,
    "float32c": nat.PML_F32C,
}


def precision_id(precision: str) -> int:
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise ValueError(
            f"unknown precision {precision!r}; expected one of {sorted(PRECISIONS)}"
        ) from None


def _slab_values(constraint, shape) -> np.ndarray:
    """Values of a boundary-slab constraint on its slab, NaN where free."""
    out = np.full(shape, np.nan)
    if constraint is not None:
        out[..., constraint.mask] = constraint.values
    return out


@dataclass
class DeviceProblem:
    eq_name: str
    params: Tuple[float, ...]
    n_fields: int
    ndim: int
    n0: int
    n1: int
    dx0: float
    dx1: float
    bc: np.ndarray  # FP64, n_fields * 4 * (n0 + n1)
    # LHS.Y_LAPLACIAN equations (field, source field, c): lap(field) = c *
    # source, re-solved after every step by FDMOperator.solve
    y_laplacian: Tuple[Tuple[int, int, float], ...] = ()

    @property
    def plane(self) -> int:
        return self.n0 * self.n1

    @property
    def elems(self) -> int:
        return self.n_fields * self.plane

    def to_device_layout(self, y: np.ndarray) -> np.ndarray:
        """Channel-last reference array -> contiguous field-major array."""
        return np.ascontiguousarray(np.moveaxis(y, -1, 0)).reshape(
            self.n_fields, self.n0, self.n1
        )

    def from_device_layout(self, y: np.ndarray, spatial_shape) -> np.ndarray:
        """Field-major ``(..., F, n0, n1)`` -> channel-last ``(..., *shape, F)``."""
        lead = y.shape[:-3]
        y = y.reshape(lead + (self.n_fields,) + tuple(spatial_shape))
        return np.ascontiguousarray(np.moveaxis(y, len(lead), -1))


def unsupported_reason(cp) -> Optional[str]:
    """Why a constrained problem cannot run on the fused device kernels, or
    None when it can."""
    diff_eq = cp.differential_equation
    dev = diff_eq.device_parameters()
    if dev is None:
        return f"{type(diff_eq).__name__} has no fused device kernel"
    if diff_eq.x_dimension not in (1, 2):
        return "device kernels cover 1-D and 2-D problems"
    if cp.mesh.coordinate_system_type != CoordinateSystem.CARTESIAN:
        return "device kernels cover Cartesian meshes"
    if not cp.are_all_boundary_conditions_static:
        return "device kernels cover static boundary conditions"
    if any(n < 3 for n in cp.mesh.vertices_shape):
        return "every axis needs at least 3 vertices"
    return None


def device_problem(cp) -> DeviceProblem:
    """Builds the kernel-side description of a constrained problem."""
    reason = unsupported_reason(cp)
    if reason is not None:
        raise NotImplementedError(reason)
    diff_eq = cp.differential_equation
    name, params = diff_eq.device_parameters()
    mesh = cp.mesh
    n_fields = diff_eq.y_dimension
    if diff_eq.x_dimension == 1:
        ndim, n0, n1 = 1, 1, mesh.vertices_shape[0]
        dx0, dx1 = 1.0, mesh.d_x[0]
    else:
        ndim = 2
        n0, n1 = mesh.vertices_shape
        dx0, dx1 = mesh.d_x
    y_cons = cp.static_y_vertex_constraints
    d_cons = cp.static_boundary_vertex_constraints[1]
    blocks = []
    for f in range(n_fields):
        full = np.full(mesh.vertices_shape + (1,), np.nan)
        if y_cons is not None and y_cons[f] is not None:
            full[..., y_cons[f].mask] = y_cons[f].values
        full = full[..., 0]
        if ndim == 2:
            dirichlet = [full[0, :], full[-1, :], full[:, 0], full[:, -1]]
            pairs = [d_cons[0, f], d_cons[1, f]]
            slab_shapes = [(1, n1, 1), (n0, 1, 1)]
            neumann = []
            for axis in range(2):
                pair = pairs[axis] if pairs[axis] is not None else (None, None)
                for side in range(2):
                    neumann.append(_slab_values(pair[side], slab_shapes[axis]).ravel())
        else:
            nan_row = np.full(n1, np.nan)
            dirichlet = [nan_row, nan_row, full[:1], full[-1:]]
            pair = d_cons[0, f] if d_cons[0, f] is not None else (None, None)
            neumann = [
                nan_row,
                nan_row,
                _slab_values(pair[0], (1, 1)).ravel(),
                _slab_values(pair[1], (1, 1)).ravel(),
            ]
        blocks.append(np.concatenate(dirichlet + neumann).astype(np.float64))
    bc = np.concatenate(blocks)
    assert bc.size == n_fields * 4 * (n0 + n1)
    return DeviceProblem(name, tuple(params), n_fields, ndim, n0, n1, dx0, dx1, bc,
                         tuple(diff_eq.device_y_laplacian()))


_streams = {}


def solver_stream():
    """A dedicated (default-priority) stream per device for the solvers, so
    that step loops can be captured in hipGraphs (capture is not possible on
    the null stream)."""
    import torch

    dev_idx = torch.cuda.current_device()
    if dev_idx not in _streams:
        _streams[dev_idx] = torch.cuda.Stream(device=dev_idx)
    return _streams[dev_idx]


def integrator_id(integrator) -> int:
    from pararealml_amd.operators.fdm.numerical_integrator import (
        RK4,
        ExplicitMidpointMethod,
        ForwardEulerMethod,
    )

    if isinstance(integrator, ForwardEulerMethod):
        return nat.PML_INT_FE
    if isinstance(integrator, ExplicitMidpointMethod):
        return nat.PML_INT_MIDPOINT
    if isinstance(integrator, RK4):
        return nat.PML_INT_RK4
    raise NotImplementedError(
        f"{type(integrator).__name__} has no fused device kernel; the device "
        "path covers ForwardEulerMethod, ExplicitMidpointMethod and RK4"
    )


class FDMBatch:
    """A batch of grids of one equation resident in HBM.

    ``layout`` is a list of ``(problem_index, count)``: grid k of the batch
    uses problem ``problems[problem_index]``; grids of the same problem share
    one boundary block, so e.g. all Parareal slices of a problem share it.
    """

    def __init__(
        self,
        problems: Sequence[DeviceProblem],
        grid_problem: Sequence[int],
        precision: str,
        integrator: int,
        stream=None,
        poisson_phase: bool = False,
    ):
        import torch

        self.lib = nat.device_lib()
        first = problems[0]
        if first.y_laplacian and not poisson_phase:
            raise NotImplementedError(
                "equations with LHS.Y_LAPLACIAN components run through FDMOperator.solve, "
                "which solves their Poisson equations after every step"
            )
        for p in problems:
            if (p.eq_name, p.params, p.ndim, p.n_fields) != (
                first.eq_name,
                first.params,
                first.ndim,
                first.n_fields,
            ):
                raise ValueError("a batch holds grids of one equation")
        self.problems = list(problems)
        self.grid_problem = list(grid_problem)
        self.precision = precision
        self.prec = precision_id(precision)
        self.integrator = integrator
        self.eq_id = nat.PML_EQ[first.eq_name]
        self.params = first.params
        self.ndim = first.ndim
        self.n_fields = first.n_fields
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.stream = stream if stream is not None else torch.cuda.current_stream()

        bc_offsets, off = [], 0
        for p in problems:
            bc_offsets.append(off)
            off += p.bc.size
        self.bc = torch.from_numpy(np.concatenate([p.bc for p in problems])).to(
            self.device
        )
        self.grids = (nat.PmlGrid * len(grid_problem))()
        self.offsets = []
        total = 0
        for k, pi in enumerate(grid_problem):
            p = problems[pi]
            g = self.grids[k]
            g.n0, g.n1, g.dx0, g.dx1 = p.n0, p.n1, p.dx0, p.dx1
            g.state_offset = total
            g.bc_offset = bc_offsets[pi]
            self.offsets.append(total)
            total += p.elems
        self.total = total
        dtype = torch.float64 if self.prec == nat.PML_F64 else torch.float32
        self.dtype = dtype
        mk = lambda: torch.empty(total, dtype=dtype, device=self.device)  # noqa: E731
        self.hi, self.work_hi = mk(), mk()
        if self.prec == nat.PML_F32C:
            self.lo, self.work_lo = mk(), mk()
        else:
            self.lo = self.work_lo = None
        self._consume(self.bc, self.hi, self.work_hi, self.lo, self.work_lo)
        self._params = (ctypes.c_double * 6)(*self.params, *([0.0] * (6 - len(self.params))))

    # ------------------------------------------------------------------ I/O
    @staticmethod
    def _ptr(t):
        return None if t is None else t.data_ptr()

    def _sptr(self):
        return ctypes.c_void_p(self.stream.cuda_stream)

    # The batch's kernels run on ``self.stream``; tensors handed in or out
    # are ordered against the caller's current stream (a no-op when the two
    # are the same stream).
    def _consume(self, *tensors) -> None:
        import torch

        cur = torch.cuda.current_stream()
        if cur != self.stream:
            self.stream.wait_stream(cur)
            for t in tensors:
                if t is not None:
                    t.record_stream(self.stream)

    def _produce(self, *tensors) -> None:
        import torch

        cur = torch.cuda.current_stream()
        if cur != self.stream:
            cur.wait_stream(self.stream)
            for t in tensors:
                if t is not None:
                    t.record_stream(self.stream)

    def load_f64(self, flat_f64) -> None:
        """Sets the state from a device FP64 tensor of ``total`` elements."""
        self._consume(flat_f64)
        nat.check(
            self.lib.pml_from_f64(
                self.prec, flat_f64.data_ptr(), self._ptr(self.hi), self._ptr(self.lo),
                self.total, self._sptr(),
            )
        )

    def state_f64(self, out=None):
        """The state as a device FP64 tensor (exact for every precision)."""
        import torch

        if out is None:
            out = torch.empty(self.total, dtype=torch.float64, device=self.device)
        self._consume(out)
        nat.check(
            self.lib.pml_to_f64(
                self.prec, self._ptr(self.hi), self._ptr(self.lo), out.data_ptr(),
                self.total, self._sptr(),
            )
        )
        self._produce(out)
        return out

    def load_host(self, ys: Sequence[np.ndarray]) -> None:
        """Sets grid k from the channel-last host array ``ys[k]``."""
        import torch

        flat = np.concatenate(
            [
                self.problems[pi].to_device_layout(np.asarray(y, dtype=np.float64)).ravel()
                for pi, y in zip(self.grid_problem, ys)
            ]
        )
        dev = torch.from_numpy(flat).to(self.device)
        self.load_f64(dev)

    # ----------------------------------------------------------------- steps
    def advance(
        self,
        d_t: float,
        n_steps: int,
        snapshots: Optional[Tuple] = None,
        snapshot_every: int = 0,
        tile_config: int = 0,
        use_graph: bool = False,
        constrain_input: bool = False,
        first_grid: int = 0,
        n_grids: Optional[int] = None,
        prepare: bool = False,
    ) -> None:
        """Advances grids ``[first_grid, first_grid + n_grids)`` (default: all)
        by ``n_steps`` steps in place.  ``snapshots`` is an optional
        ``(hi, lo)`` pair of device tensors receiving the batch state after
        every ``snapshot_every``-th step; ``constrain_input`` imposes the
        Dirichlet values on the state before the first step.  ``prepare``
        builds (and caches) the work-unit tables and the step-loop graph of
        this call without launching anything."""
        snap_hi = snap_lo = None
        if snapshots is not None:
            snap_hi, snap_lo = snapshots
            self._consume(snap_hi, snap_lo)
        nat.check(
            self.lib.pml_fdm_solve(
                self.eq_id,
                self._params,
                len(self.params),
                self.integrator,
                self.prec,
                self.ndim,
                ctypes.byref(self.grids[first_grid]),
                len(self.grid_problem) - first_grid if n_grids is None else n_grids,
                self.bc.data_ptr(),
                self._ptr(self.hi),
                self._ptr(self.lo),
                self._ptr(self.work_hi),
                self._ptr(self.work_lo),
                self.total,
                float(d_t),
                int(n_steps),
                self._ptr(snap_hi),
                self._ptr(snap_lo),
                int(snapshot_every),
                int(tile_config),
                (nat.PML_FLAG_GRAPH if use_graph else 0)
                | (nat.PML_FLAG_CONSTRAIN_INPUT if constrain_input else 0)
                | (nat.PML_FLAG_PREPARE if prepare else 0),
                self._sptr(),
            )
        )

    def new_snapshots(self, n: int):
        import torch

        hi = torch.empty(n * self.total, dtype=self.dtype, device=self.device)
        lo = (
            torch.empty(n * self.total, dtype=self.dtype, device=self.device)
            if self.prec == nat.PML_F32C
            else None
        )
        return hi, lo

    def snapshots_f64(self, snaps, n: int):
        import torch

        hi, lo = snaps
        out = torch.empty(n * self.total, dtype=torch.float64, device=self.device)
        self._consume(out)
        nat.check(
            self.lib.pml_to_f64(
                self.prec, hi.data_ptr(), self._ptr(lo), out.data_ptr(), n * self.total,
                self._sptr(),
            )
        )
        self._produce(out)
        return out.view(n, self.total)

    def grid_view(self, flat, k: int):
        """Slice of grid k, shaped ``(..., F, n0, n1)``, of a flat tensor whose
        last axis spans the batch state."""
        p = self.problems[self.grid_problem[k]]
        off = self.offsets[k]
        return flat[..., off : off + p.elems].reshape(
            flat.shape[:-1] + (p.n_fields, p.n0, p.n1)
        )
