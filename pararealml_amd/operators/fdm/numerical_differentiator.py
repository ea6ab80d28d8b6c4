"""Three-point finite-difference operators evaluated by gfx950 kernels.

The public methods and their validation follow
pararealml/operators/fdm/numerical_differentiator.py:14-996 (Cartesian
branches) and the stencil semantics follow
``ThreePointCentralDifferenceMethod`` (numerical_differentiator.py:
999-1242):

* first derivatives: zero halos, boundary values replaced by the derivative
  constraints (``pml_stencil_first``);
* same-axis second derivatives: ghosts ``y[1] - 2dx g`` / ``y[n-2] + 2dx g``
  where constrained, 0 elsewhere (``pml_stencil_second``);
* mixed second derivatives: a constrained first derivative followed by an
  unconstrained one;
* the anti-Laplacian: Jacobi sweeps (``pml_jacobi_sweep``) with the y
  constraints re-imposed after each sweep until the 2-norm of the update is
  at most ``tol``.

Inputs may be NumPy arrays (copied to HBM, results copied back) or torch
tensors already on the GPU (results stay there).  Arrays are in the
reference layout, channel-last ``(*vertices_shape, y_dim)``.  Polar,
cylindrical and spherical meshes follow the reference's curvilinear
branches: each operator is a sum of stencil results and field components
weighted by coordinate factors (1/r, 1/(r sin phi), cos(phi)/sin(phi), ...),
every term evaluated on the device (``pml_weighted_add``); the factors are
computed once per call from the mesh coordinates (for the anti-Laplacian:
per-vertex Jacobi weights, ``pml_jacobi_solve_weighted``).
"""

from abc import ABC
from typing import Optional, Sequence, Union

import ctypes
import numpy as np

from pararealml_amd import _native as nat
from pararealml_amd.constraint import Constraint
from pararealml_amd.mesh import CoordinateSystem, Mesh


# --------------------------------------------------------------------------
# array plumbing


class _Arr:
    """A contiguous device copy of an input plus how to hand results back."""

    def __init__(self, y):
        import torch

        self.is_numpy = isinstance(y, np.ndarray)
        if self.is_numpy:
            arr = np.ascontiguousarray(y)
            if arr.dtype not in (np.float32, np.float64):
                arr = arr.astype(np.float64)
            self.t = torch.from_numpy(arr).to("cuda")
        else:
            t = y if y.is_cuda else y.to("cuda")
            if t.dtype not in (torch.float32, torch.float64):
                t = t.to(torch.float64)
            self.t = t.contiguous()
        self.prec = nat.PML_F64 if self.t.dtype == torch.float64 else nat.PML_F32

    def new(self, shape):
        import torch

        return torch.empty(shape, dtype=self.t.dtype, device=self.t.device)

    def give(self, t):
        return t.cpu().numpy() if self.is_numpy else t


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _slab_array(pairs, side: int, shape, axis: int):
    """FP64 NaN-encoded (outer, inner) boundary values of the derivative
    constraints on one side of ``axis`` for an array of ``shape``, or None
    when nothing is constrained."""
    if pairs is None:
        return None
    slab_shape = list(shape)
    slab_shape[axis] = 1
    slab = np.full(slab_shape, np.nan)
    any_set = False
    for f, pair in enumerate(pairs):
        if pair is None or pair[side] is None:
            continue
        c = pair[side]
        slab[..., f : f + 1][..., c.mask] = c.values
        any_set = True
    if not any_set:
        return None
    import torch

    return torch.from_numpy(np.ascontiguousarray(slab).ravel()).to("cuda")


def _view(shape, axis):
    outer = int(np.prod(shape[:axis], dtype=np.int64))
    inner = int(np.prod(shape[axis + 1 :], dtype=np.int64))
    return outer, int(shape[axis]), inner


def _first(src: "_Arr", t, out, axis: int, d_x: float, pairs, mode: int):
    lib = nat.device_lib()
    lo = _slab_array(pairs, 0, t.shape, axis)
    hi = _slab_array(pairs, 1, t.shape, axis)
    outer, n, inner = _view(t.shape, axis)
    nat.check(
        lib.pml_stencil_first(
            src.prec, t.data_ptr(), out.data_ptr(), outer, n, inner, float(d_x),
            None if lo is None else lo.data_ptr(),
            None if hi is None else hi.data_ptr(),
            mode, _stream(),
        )
    )


def _second(src: "_Arr", t, out, axis: int, d_x: float, d_x2: float, pairs, mode: int):
    lib = nat.device_lib()
    lo = _slab_array(pairs, 0, t.shape, axis)
    hi = _slab_array(pairs, 1, t.shape, axis)
    outer, n, inner = _view(t.shape, axis)
    nat.check(
        lib.pml_stencil_second(
            src.prec, t.data_ptr(), out.data_ptr(), outer, n, inner, float(d_x),
            float(d_x2),
            None if lo is None else lo.data_ptr(),
            None if hi is None else hi.data_ptr(),
            mode, _stream(),
        )
    )


def _weights(src: "_Arr", coef, shape):
    import torch

    dtype = np.float64 if src.prec == nat.PML_F64 else np.float32
    arr = np.array(np.broadcast_to(coef, tuple(shape)), dtype=dtype)
    return torch.from_numpy(arr).to("cuda")


class _Sum:
    """``out`` = a sum of terms, each a stencil result or a field component
    times an optional per-vertex factor; every term runs on the device (the
    first term writes, the rest accumulate)."""

    def __init__(self, src: "_Arr", shape):
        self.a = src
        self.shape = tuple(shape)
        self.out = src.new(self.shape)
        self.mode = 0
        self._tmp = None

    def _scratch(self):
        if self._tmp is None:
            self._tmp = self.a.new(self.shape)
        return self._tmp

    def _weighted(self, x, coef):
        c = _weights(self.a, coef, self.shape)
        nat.check(
            nat.device_lib().pml_weighted_add(
                self.a.prec, x.data_ptr(), c.data_ptr(), self.out.data_ptr(), self.out.numel(),
                self.mode, _stream(),
            )
        )
        self.mode = 1

    def first(self, y, axis: int, d_x: float, pairs, coef=None):
        if coef is None or (isinstance(coef, float) and coef == -1.0 and self.mode == 1):
            _first(self.a, y, self.out, axis, d_x, pairs, -1 if coef is not None else self.mode)
            self.mode = 1
        else:
            t = self._scratch()
            _first(self.a, y, t, axis, d_x, pairs, 0)
            self._weighted(t, coef)

    def second(self, y, axis: int, d_x: float, pairs, coef=None):
        if coef is None:
            _second(self.a, y, self.out, axis, d_x, d_x, pairs, self.mode)
            self.mode = 1
        else:
            t = self._scratch()
            _second(self.a, y, t, axis, d_x, d_x, pairs, 0)
            self._weighted(t, coef)

    def mixed(self, y, axis1: int, axis2: int, d_x1: float, d_x2: float, pairs, coef=None):
        inner = self.a.new(self.shape)
        _first(self.a, y, inner, axis1, d_x1, pairs, 0)
        self.first(inner, axis2, d_x2, [None] * self.shape[-1], coef)

    def value(self, y, coef):
        self._weighted(y, coef)


def _jacobi_weights(mesh: Mesh):
    """Per-vertex weights of the curvilinear Jacobi update
    (numerical_differentiator.py:1097-1186): the new estimate is
    (sum_a w_sum[a] (prev + next) + w_diff[a] (next - prev) - laplacian)
    / w_den.  Lists of three FP64 vertex arrays (None: no term) and the
    denominator array."""
    r, sin_phi, cos_phi = _polar_factors(mesh)
    r, sin_phi, cos_phi = (None if a is None else a[..., 0] for a in (r, sin_phi, cos_phi))
    dx = list(mesh.d_x) + [1.0] * (3 - mesh.dimensions)
    sq = [d * d for d in dx]
    r2 = r * r
    ones = np.ones(mesh.vertices_shape)
    w_sum = [ones / sq[0], None, None]
    w_diff = [None, None, None]
    if mesh.coordinate_system_type == CoordinateSystem.SPHERICAL:
        r2s2 = r2 * sin_phi * sin_phi
        w_diff[0] = 1.0 / (dx[0] * r)
        w_sum[1] = 1.0 / (sq[1] * r2s2)
        w_sum[2] = 1.0 / (sq[2] * r2)
        w_diff[2] = cos_phi / (2.0 * dx[2] * sin_phi * r2)
        w_den = 2.0 / sq[0] + 2.0 / (sq[1] * r2s2) + 2.0 / (sq[2] * r2)
    else:
        w_diff[0] = 1.0 / (2.0 * dx[0] * r)
        w_sum[1] = 1.0 / (sq[1] * r2)
        w_den = 2.0 / sq[0] + 2.0 / (sq[1] * r2)
        if mesh.coordinate_system_type == CoordinateSystem.CYLINDRICAL:
            w_sum[2] = ones / sq[2]
            w_den = w_den + 2.0 / sq[2]
    return w_sum, w_diff, w_den


def _polar_factors(mesh: Mesh):
    """r, sin(phi), cos(phi) on the vertices as (*vertices, 1) arrays; the
    angle terms are None outside spherical coordinates."""
    grids = mesh.vertex_coordinate_grids
    r = grids[0][..., np.newaxis]
    if mesh.coordinate_system_type == CoordinateSystem.SPHERICAL:
        phi = grids[2][..., np.newaxis]
        return r, np.sin(phi), np.cos(phi)
    return r, None, None


# --------------------------------------------------------------------------


class NumericalDifferentiator(ABC):
    """Validation and public interface (numerical_differentiator.py:14-996)."""

    def __init__(self, tol: float = 1e-3, *, max_sweeps: int = 10_000_000):
        if tol < 0.0:
            raise ValueError("tolerance must be non-negative")
        if max_sweeps < 1:
            raise ValueError("max_sweeps must be positive")
        self._tol = tol
        # The reference's Jacobi loop has no bound; this one raises
        # RuntimeError after max_sweeps sweeps without reaching tol.
        self._max_sweeps = int(max_sweeps)
        self.last_jacobi_sweeps = 0

    # ---- validation helpers (numerical_differentiator.py:929-996) ---------
    @staticmethod
    def _verify_input_shape_matches_mesh(arr, mesh: Mesh, name: str = "y"):
        if tuple(arr.shape[:-1]) != tuple(mesh.vertices_shape):
            raise ValueError(
                f"{name} shape up to second to last axis {tuple(arr.shape[:-1])} "
                f"must match mesh vertices shape {mesh.vertices_shape}"
            )

    @staticmethod
    def _verify_input_is_a_vector_field(arr, mesh: Mesh):
        NumericalDifferentiator._verify_input_shape_matches_mesh(arr, mesh)
        if arr.shape[-1] != mesh.dimensions:
            raise ValueError(
                f"y value vector length ({arr.shape[-1]}) must match number of "
                f"x dimensions ({mesh.dimensions})"
            )

    @staticmethod
    def _verify_and_get_derivative_boundary_constraints(dbc, x_axes, y_elements):
        if dbc is None:
            return np.empty((x_axes, y_elements), dtype=object)
        if dbc.shape != (x_axes, y_elements):
            raise ValueError(
                "expected derivative boundary constraints shape to be "
                f"{(x_axes, y_elements)} but got {dbc.shape}"
            )
        return dbc

    @staticmethod
    def _require_points(shape, axis):
        if shape[axis] <= 2:
            raise ValueError(f"y must contain at least 3 points along x-axis ({axis})")


class ThreePointCentralDifferenceMethod(NumericalDifferentiator):
    """Second-order central differences on gfx950
    (numerical_differentiator.py:999-1242)."""

    def __init__(self, tol: float = 1e-3, *, max_sweeps: int = 10_000_000):
        super().__init__(tol, max_sweeps=max_sweeps)

    # ---------------------------------------------------------------- public
    def gradient(self, y, mesh: Mesh, x_axis: int, derivative_boundary_constraints=None):
        """numerical_differentiator.py:114-173."""
        self._verify_input_shape_matches_mesh(y, mesh)
        if not (0 <= x_axis < mesh.dimensions):
            raise ValueError(
                f"x-axis ({x_axis}) must be non-negative and less than number "
                f"of x dimensions ({mesh.dimensions})"
            )
        dbc = self._verify_and_get_derivative_boundary_constraints(
            derivative_boundary_constraints, mesh.dimensions, y.shape[-1]
        )
        self._require_points(y.shape, x_axis)
        a = _Arr(y)
        out = _Sum(a, a.t.shape)
        cs = mesh.coordinate_system_type
        coef = None
        if cs != CoordinateSystem.CARTESIAN:
            r, sin_phi, _ = _polar_factors(mesh)
            if cs == CoordinateSystem.SPHERICAL:
                coef = {0: None, 1: 1.0 / (r * sin_phi), 2: 1.0 / r}[x_axis]
            elif x_axis == 1:
                coef = 1.0 / r
        out.first(a.t, x_axis, mesh.d_x[x_axis], dbc[x_axis], coef)
        return a.give(out.out)

    def hessian(
        self, y, mesh: Mesh, x_axis1: int, x_axis2: int, derivative_boundary_constraints=None
    ):
        """numerical_differentiator.py:175-308."""
        self._verify_input_shape_matches_mesh(y, mesh)
        if not (0 <= x_axis1 < mesh.dimensions) or not (0 <= x_axis2 < mesh.dimensions):
            raise ValueError(
                f"both first x-axis ({x_axis1}) and second x-axis ({x_axis2}) "
                "must be non-negative and less than number of x dimensions "
                f"({mesh.dimensions})"
            )
        dbc = self._verify_and_get_derivative_boundary_constraints(
            derivative_boundary_constraints, mesh.dimensions, y.shape[-1]
        )
        self._require_points(y.shape, x_axis1)
        self._require_points(y.shape, x_axis2)
        a = _Arr(y)
        out = _Sum(a, a.t.shape)
        d = mesh.d_x

        def second(coef=None):
            if x_axis1 == x_axis2:
                out.second(a.t, x_axis1, d[x_axis1], dbc[x_axis1], coef)
            else:
                out.mixed(a.t, x_axis1, x_axis2, d[x_axis1], d[x_axis2], dbc[x_axis1], coef)

        def first(axis, coef):
            self._require_points(y.shape, axis)
            out.first(a.t, axis, d[axis], dbc[axis], coef)

        cs = mesh.coordinate_system_type
        axes = {x_axis1, x_axis2}
        if cs == CoordinateSystem.CARTESIAN:
            second()
        elif cs == CoordinateSystem.SPHERICAL:
            r, sin_phi, cos_phi = _polar_factors(mesh)
            if axes == {0}:
                second()
            elif axes == {1}:
                second(1.0 / (r * sin_phi) ** 2)
                first(0, 1.0 / r)
                first(2, cos_phi / (r * r * sin_phi))
            elif axes == {2}:
                second(1.0 / (r * r))
                first(0, 1.0 / r)
            elif axes == {0, 1}:
                second(1.0 / (r * sin_phi))
                first(1, -1.0 / (r * r * sin_phi))
            elif axes == {0, 2}:
                second(1.0 / r)
                first(2, -1.0 / (r * r))
            else:
                second(1.0 / (r * r * sin_phi))
                first(1, -cos_phi / (r * sin_phi) ** 2)
        else:
            r, _, _ = _polar_factors(mesh)
            if axes <= {0, 2}:
                second()
            elif axes == {1}:
                second(1.0 / (r * r))
                first(0, 1.0 / r)
            elif axes == {0, 1}:
                second(1.0 / r)
                first(1, -1.0 / (r * r))
            else:
                second(1.0 / r)
        return a.give(out.out)

    def divergence(self, y, mesh: Mesh, derivative_boundary_constraints=None):
        """numerical_differentiator.py:310-400 (Cartesian: the component
        derivatives accumulated in component order)."""
        self._verify_input_is_a_vector_field(y, mesh)
        dbc = self._verify_and_get_derivative_boundary_constraints(
            derivative_boundary_constraints, mesh.dimensions, y.shape[-1]
        )
        for i in range(y.shape[-1]):
            self._require_points(y.shape, i)
        a = _Arr(y)
        out = _Sum(a, tuple(a.t.shape[:-1]) + (1,))
        comp = [a.t[..., i : i + 1].contiguous() for i in range(y.shape[-1])]

        def d(i, coef=None):
            out.first(comp[i], i, mesh.d_x[i], dbc[i, i : i + 1], coef)

        cs = mesh.coordinate_system_type
        if cs == CoordinateSystem.CARTESIAN:
            for i in range(y.shape[-1]):
                d(i)
        elif cs == CoordinateSystem.SPHERICAL:
            r, sin_phi, cos_phi = _polar_factors(mesh)
            d(0)
            d(2, 1.0 / r)
            out.value(comp[0], 2.0 / r)
            d(1, 1.0 / (r * sin_phi))
            out.value(comp[2], cos_phi / (r * sin_phi))
        else:
            r, _, _ = _polar_factors(mesh)
            d(0)
            out.value(comp[0], 1.0 / r)
            d(1, 1.0 / r)
            if cs == CoordinateSystem.CYLINDRICAL:
                d(2)
        return a.give(out.out)

    def curl(self, y, mesh: Mesh, curl_ind: int = 0, derivative_boundary_constraints=None):
        """numerical_differentiator.py:402-590."""
        self._verify_input_is_a_vector_field(y, mesh)
        if not (2 <= mesh.dimensions <= 3):
            raise ValueError(f"number of x dimensions ({mesh.dimensions}) must be 2 or 3")
        if mesh.dimensions == 2 and curl_ind != 0:
            raise ValueError(f"curl index ({curl_ind}) must be 0 for 2D curl")
        if not (0 <= curl_ind < mesh.dimensions):
            raise ValueError(
                f"curl index ({curl_ind}) must be non-negative and less than "
                f"number of x dimensions ({mesh.dimensions})"
            )
        dbc = self._verify_and_get_derivative_boundary_constraints(
            derivative_boundary_constraints, mesh.dimensions, y.shape[-1]
        )
        a = _Arr(y)
        out = _Sum(a, tuple(a.t.shape[:-1]) + (1,))
        comp = [a.t[..., i : i + 1].contiguous() for i in range(y.shape[-1])]

        def d(c, axis, coef=None):
            """coef * d(y_c)/d(x_axis), the constraints of (axis, c)."""
            self._require_points(y.shape, axis)
            out.first(comp[c], axis, mesh.d_x[axis], dbc[axis, c : c + 1], coef)

        cs = mesh.coordinate_system_type
        planar = mesh.dimensions == 2 or curl_ind == 2
        if cs == CoordinateSystem.CARTESIAN:
            if planar:
                (pc, pa), (mc, ma) = (1, 0), (0, 1)
            elif curl_ind == 0:
                (pc, pa), (mc, ma) = (2, 1), (1, 2)
            else:
                (pc, pa), (mc, ma) = (0, 2), (2, 0)
            d(pc, pa)
            d(mc, ma, -1.0)
        elif cs == CoordinateSystem.SPHERICAL:
            r, sin_phi, cos_phi = _polar_factors(mesh)
            if curl_ind == 0:
                d(1, 2, 1.0 / r)
                out.value(comp[1], cos_phi / (r * sin_phi))
                d(2, 1, -1.0 / (r * sin_phi))
            elif curl_ind == 1:
                d(2, 0)
                out.value(comp[2], 1.0 / r)
                d(0, 2, -1.0 / r)
            else:
                d(1, 0, -1.0)
                d(0, 1, 1.0 / (r * sin_phi))
                out.value(comp[1], -1.0 / r)
        else:
            r, _, _ = _polar_factors(mesh)
            if cs == CoordinateSystem.POLAR or curl_ind == 2:
                d(1, 0)
                out.value(comp[1], 1.0 / r)
                d(0, 1, -1.0 / r)
            elif curl_ind == 0:
                d(2, 1, 1.0 / r)
                d(1, 2, -1.0)
            else:
                d(0, 2)
                d(2, 0, -1.0)
        return a.give(out.out)

    def laplacian(self, y, mesh: Mesh, derivative_boundary_constraints=None):
        """Element-wise scalar Laplacian (numerical_differentiator.py:592-725;
        Cartesian: accumulated over axes in order)."""
        self._verify_input_shape_matches_mesh(y, mesh)
        dbc = self._verify_and_get_derivative_boundary_constraints(
            derivative_boundary_constraints, mesh.dimensions, y.shape[-1]
        )
        a = _Arr(y)
        return a.give(self._laplacian_dev(a, a.t, mesh, dbc))

    def _laplacian_dev(self, a, t, mesh, dbc):
        out = _Sum(a, t.shape)
        self._laplacian_terms(out, t, mesh, dbc)
        return out.out

    def _laplacian_terms(self, out, t, mesh, dbc):
        for axis in range(t.dim() - 1):
            self._require_points(t.shape, axis)
        d = mesh.d_x
        cs = mesh.coordinate_system_type
        if cs == CoordinateSystem.CARTESIAN:
            for axis in range(t.dim() - 1):
                out.second(t, axis, d[axis], dbc[axis, :])
        elif cs == CoordinateSystem.SPHERICAL:
            r, sin_phi, cos_phi = _polar_factors(mesh)
            out.second(t, 0, d[0], dbc[0, :])
            out.first(t, 0, d[0], dbc[0, :], 2.0 / r)
            out.second(t, 2, d[2], dbc[2, :], 1.0 / (r * r))
            out.first(t, 2, d[2], dbc[2, :], cos_phi / (r * r * sin_phi))
            out.second(t, 1, d[1], dbc[1, :], 1.0 / (r * sin_phi) ** 2)
        else:
            r, _, _ = _polar_factors(mesh)
            out.second(t, 0, d[0], dbc[0, :])
            out.second(t, 1, d[1], dbc[1, :], 1.0 / (r * r))
            out.first(t, 0, d[0], dbc[0, :], 1.0 / r)
            if cs == CoordinateSystem.CYLINDRICAL:
                out.second(t, 2, d[2], dbc[2, :])

    def vector_laplacian(
        self, y, mesh: Mesh, vector_laplacian_ind: int, derivative_boundary_constraints=None
    ):
        """numerical_differentiator.py:727-870: the scalar Laplacian of the
        selected component plus, in curvilinear coordinates, the coupling
        terms of the other components."""
        self._verify_input_is_a_vector_field(y, mesh)
        if not (0 <= vector_laplacian_ind < mesh.dimensions):
            raise ValueError(
                f"vector Laplacian index ({vector_laplacian_ind}) must be "
                "non-negative and less than number of x dimensions "
                f"({mesh.dimensions})"
            )
        dbc = self._verify_and_get_derivative_boundary_constraints(
            derivative_boundary_constraints, mesh.dimensions, y.shape[-1]
        )
        k = vector_laplacian_ind
        a = _Arr(y)
        comp = [a.t[..., i : i + 1].contiguous() for i in range(y.shape[-1])]
        out = _Sum(a, comp[k].shape)
        self._laplacian_terms(out, comp[k], mesh, dbc[:, k : k + 1])
        cs = mesh.coordinate_system_type
        if cs == CoordinateSystem.CARTESIAN:
            return a.give(out.out)

        def d(c, axis, coef):
            self._require_points(y.shape, axis)
            out.first(comp[c], axis, mesh.d_x[axis], dbc[axis, c : c + 1], coef)

        r, sin_phi, cos_phi = _polar_factors(mesh)
        r2 = r * r
        if cs == CoordinateSystem.SPHERICAL:
            if k == 1:
                out.value(comp[0], -2.0 / r2)
                d(2, 2, -2.0 / r2)
                out.value(comp[2], -2.0 * cos_phi / (r2 * sin_phi))
                d(1, 1, -2.0 / (r2 * sin_phi))
            elif k == 2:
                d(0, 1, 2.0 / (r2 * sin_phi))
                d(2, 1, 2.0 * cos_phi / (r2 * sin_phi * sin_phi))
                out.value(comp[1], -1.0 / (r2 * sin_phi * sin_phi))
            else:
                d(0, 2, 2.0 / r2)
                out.value(comp[2], -1.0 / (r2 * sin_phi * sin_phi))
                d(1, 1, -2.0 * cos_phi / (r2 * sin_phi * sin_phi))
        else:
            if k == 0:
                out.value(comp[0], -1.0 / r2)
                d(1, 1, -2.0 / r2)
            elif k == 1:
                out.value(comp[1], -1.0 / r2)
                d(0, 1, 2.0 / r2)
        return a.give(out.out)

    def anti_laplacian(
        self,
        laplacian,
        mesh: Mesh,
        y_constraints: Union[Sequence[Optional[Constraint]], np.ndarray],
        derivative_boundary_constraints=None,
        y_init=None,
    ):
        """Jacobi solve of the Poisson problem with the given Laplacian
        (numerical_differentiator.py:872-927, 1097-1186)."""
        import torch

        self._verify_input_shape_matches_mesh(laplacian, mesh, "Laplacian")
        dbc = self._verify_and_get_derivative_boundary_constraints(
            derivative_boundary_constraints, mesh.dimensions, laplacian.shape[-1]
        )
        if mesh.dimensions > 3:
            raise NotImplementedError("the device Jacobi sweep covers up to 3 axes")
        if not np.all(np.array(laplacian.shape[:-1]) > 2):
            raise ValueError("y must contain at least 3 points along all x axes")
        if y_init is None:
            y_host = np.random.random(laplacian.shape)
        else:
            if y_init.shape != laplacian.shape:
                raise ValueError
            y_host = y_init
        lap = _Arr(laplacian)
        shape = tuple(laplacian.shape)
        n_fields = shape[-1]
        spatial = list(shape[:-1]) + [1] * (3 - (len(shape) - 1))
        dxs = list(mesh.d_x) + [1.0] * (3 - mesh.dimensions)

        # full-grid y constraints, NaN where free
        ycons = None
        if y_constraints is not None:
            full = np.full(shape, np.nan)
            for f, c in enumerate(y_constraints):
                if c is not None:
                    full[..., f : f + 1][..., c.mask] = c.values
            ycons = torch.from_numpy(full.ravel()).to("cuda")
        # ghost slabs per axis in "axis removed" layout
        glo, ghi, keep = [], [], []
        for axis in range(3):
            if axis < mesh.dimensions:
                for side, dst in ((0, glo), (1, ghi)):
                    s = _slab_array(dbc[axis], side, shape, axis)
                    keep.append(s)
                    dst.append(None if s is None else s.data_ptr())
            else:
                glo.append(None)
                ghi.append(None)

        y = _Arr(np.asarray(y_host, dtype=np.float64) if lap.is_numpy else y_host)
        cur = y.t.to(lap.t.dtype).clone()
        if ycons is not None:
            mask = ~torch.isnan(ycons)
            cur.view(-1)[mask] = ycons[mask].to(cur.dtype)
        work = torch.empty_like(cur)
        state = torch.zeros(nat.PML_JACOBI_STATE_BYTES // 8, dtype=torch.float64,
                            device=cur.device)
        lib = nat.device_lib()
        c_shape = (ctypes.c_int64 * 3)(*spatial)
        c_dx = (ctypes.c_double * 3)(*dxs)
        c_lo = (ctypes.c_void_p * 3)(*glo)
        c_hi = (ctypes.c_void_p * 3)(*ghi)
        sweeps = ctypes.c_int64(0)
        diff = ctypes.c_double(0.0)
        # the loop and its stopping test run on the device; the host looks
        # at the loop state once per 64 sweeps
        args = (
            lap.prec, cur.data_ptr(), work.data_ptr(), lap.t.data_ptr(), c_shape,
            n_fields, c_dx, c_lo, c_hi,
            None if ycons is None else ycons.data_ptr(),
            float(self._tol), int(self._max_sweeps), 64, state.data_ptr(),
            ctypes.byref(sweeps), ctypes.byref(diff),
        )
        if mesh.coordinate_system_type == CoordinateSystem.CARTESIAN:
            nat.check(lib.pml_jacobi_solve(*args, _stream()))
        else:
            w_sum, w_diff, w_den = _jacobi_weights(mesh)
            dev = [None if w is None else torch.from_numpy(np.ascontiguousarray(w).ravel())
                   .to("cuda") for w in w_sum + w_diff + [w_den]]
            ptr = [None if w is None else w.data_ptr() for w in dev]
            nat.check(
                lib.pml_jacobi_solve_weighted(
                    *args, (ctypes.c_void_p * 3)(*ptr[0:3]), (ctypes.c_void_p * 3)(*ptr[3:6]),
                    ptr[6], _stream(),
                )
            )
        if sweeps.value < 0:
            raise RuntimeError(
                f"the Jacobi iteration did not reach tol={self._tol} in "
                f"{-sweeps.value} sweeps (last update norm {diff.value:.3e})"
            )
        self.last_jacobi_sweeps = int(sweeps.value)
        return lap.give(cur)
