"""Finite-difference operator backed by fused gfx950 kernels.

``FDMOperator(integrator, differentiator, d_t).solve(ivp)`` keeps the
reference's interface and result (pararealml/operators/fdm/
fdm_operator.py:27-77): the solution at every time step after t0, as a
vertex-oriented ``Solution``.  The time loop runs on the GPU:
``pml_fdm_solve`` advances the state held in HBM with one fused kernel per
step (or per several steps), writing each requested snapshot straight from
the kernel.  Right-hand sides with ``LHS.Y`` components are recomputed from
the pre-step state after each step as in fdm_operator.py:127-139.

Device coverage: Cartesian 1-D and 2-D problems with static boundary
conditions, the equations with fused kernels (diffusion, convection-
diffusion, wave, Cahn-Hilliard, Burgers, shallow water, Navier-Stokes with
its stream function re-solved by device Jacobi sweeps), the explicit
integrators and ``ThreePointCentralDifferenceMethod``.  Polar,
cylindrical and spherical meshes, 3-D Cartesian ones, time-dependent
boundary conditions and the implicit integrators run through a composed
device loop (``_solve_composed``: the differentiator's device operators per
stage) for the diffusion, convection-diffusion, wave, shallow-water,
Burgers, Cahn-Hilliard and Navier-Stokes equations.
Anything else raises ``NotImplementedError``; there is no host fallback.

Any other equation -- every ODE system and any user-defined
``DifferentialEquation`` -- runs its symbolic right-hand sides through a
device program (``pararealml_amd.operators.device_program``): ODEs in one
launch of the batched ODE kernel for the whole time loop
(``pml_ode_solve``), PDEs through the composed loop with the derivative
symbols computed by the differentiator's device operators and the
expressions by ``pml_expr_eval``.

Extensions over the reference (keyword-only, defaults reproduce it):
``precision`` ("float64", "float32" or "float32c" = compensated FP32),
``snapshot_every`` (keep every k-th step only), ``tile_config`` (kernel
variant; None = the tuned or default one, see ``tune``), ``use_graph`` and
``symbolic_rhs`` (evaluate the symbolic right-hand side through the device
program even when a fused kernel exists).
"""

import warnings
from typing import Optional

import numpy as np

from pararealml_amd import _device as dev
from pararealml_amd import _native as nat
from pararealml_amd.initial_value_problem import InitialValueProblem
from pararealml_amd.mesh import CoordinateSystem
from pararealml_amd.operator import Operator, discretize_time_domain
from pararealml_amd.operators.fdm.numerical_differentiator import (
    NumericalDifferentiator,
    ThreePointCentralDifferenceMethod,
)
from pararealml_amd.operators.fdm.numerical_integrator import ImplicitMethod, NumericalIntegrator
from pararealml_amd.solution import DeviceSolution, Solution


class _SymbolicDeviceRHS:
    """Device evaluation of some of a system's right-hand sides on a
    channel-last state: the FDMSymbolMapper's symbol map
    (fdm_symbol_mapper.py:45-158) with every derivative symbol computed by
    the differentiator's device operators (constraints of the symbol's own
    components at the evaluation time) and the expressions by one
    ``pml_expr_eval`` launch of their compiled program
    (symbol_mapper.py:222-271)."""

    def __init__(self, cp, diff, indices, dtype):
        import sympy as sp
        import torch

        from pararealml_amd.operators.device_program import compile_expressions

        rhs = cp.differential_equation.symbolic_equation_system.rhs
        exprs = [sp.sympify(rhs[i]) for i in indices]
        syms = sorted(set().union(*[e.free_symbols for e in exprs]), key=lambda s: s.name)
        self.inputs = [s for s in syms if s.name != "t"]
        index = {s: k for k, s in enumerate(self.inputs)}
        self.program = compile_expressions(exprs, index.__getitem__, self.inputs)
        self.specs = []
        for sym in self.inputs:
            tokens = sym.name.split("_")
            self.specs.append((tokens[0], [int(v) for v in tokens[1:]]))
        self.mesh = cp.mesh
        self.diff = diff
        self.dtype = dtype
        self.prec = nat.PML_F64 if dtype == torch.float64 else nat.PML_F32
        self.x = {}
        for prefix, idx in self.specs:
            if prefix == "x" and idx[0] not in self.x:
                grid = np.array(cp.mesh.vertex_coordinate_grids[idx[0]])
                self.x[idx[0]] = torch.from_numpy(grid).to(device="cuda", dtype=dtype)

    def _derivative(self, prefix, idx, y, dbc):
        diff, mesh = self.diff, self.mesh
        if prefix == "y-gradient":
            i, axis = idx
            return diff.gradient(y[..., i:i + 1].contiguous(), mesh, axis, dbc[:, i:i + 1])
        if prefix == "y-hessian":
            i, a1, a2 = idx
            return diff.hessian(y[..., i:i + 1].contiguous(), mesh, a1, a2, dbc[:, i:i + 1])
        if prefix == "y-laplacian":
            i = idx[0]
            return diff.laplacian(y[..., i:i + 1].contiguous(), mesh, dbc[:, i:i + 1])
        if prefix == "y-divergence":
            return diff.divergence(y[..., idx].contiguous(), mesh, dbc[:, idx])
        if prefix == "y-curl":
            if mesh.dimensions == 2:
                return diff.curl(y[..., idx].contiguous(), mesh, 0, dbc[:, idx])
            comps = idx[:-1]
            return diff.curl(y[..., comps].contiguous(), mesh, idx[-1], dbc[:, comps])
        if prefix == "y-vector-laplacian":
            comps = idx[:-1]
            return diff.vector_laplacian(y[..., comps].contiguous(), mesh, idx[-1],
                                         dbc[:, comps])
        raise ValueError(f"unknown symbol prefix {prefix}")

    def __call__(self, y, t, dbc, out, cols) -> None:
        """Writes right-hand side k into column cols[k] of the channel-last
        tensor ``out`` (same vertices as ``y``)."""
        import ctypes

        import torch

        y_dim = y.shape[-1]
        esz = y.element_size()
        ptrs, strides, keep = [], [], []
        for prefix, idx in self.specs:
            if prefix == "y":
                ptrs.append(y.data_ptr() + idx[0] * esz)
                strides.append(y_dim)
            elif prefix == "x":
                ptrs.append(self.x[idx[0]].data_ptr())
                strides.append(1)
            else:
                arr = self._derivative(prefix, idx, y, dbc).contiguous()
                keep.append(arr)
                ptrs.append(arr.data_ptr())
                strides.append(1)
        n_out = out.shape[-1]
        outs = [out.data_ptr() + int(c) * out.element_size() for c in cols]
        n_in, n_o = len(ptrs), len(outs)
        nat.check(nat.device_lib().pml_expr_eval(
            self.prec, self.program.handle(), (ctypes.c_void_p * max(n_in, 1))(*ptrs),
            (ctypes.c_int64 * max(n_in, 1))(*strides), (ctypes.c_void_p * n_o)(*outs),
            (ctypes.c_int64 * n_o)(*([n_out] * n_o)), float(t), y.numel() // y_dim,
            torch.cuda.current_stream().cuda_stream))
        del keep


def ode_time_step(t0: float, n_steps: int, d_t: float) -> float:
    """The spacing of the reference's time grid ``linspace(t0, t0 + n d_t,
    n + 1)`` (operator.py:60-74), as NumPy computes it: ``(stop - start) /
    n``; step s starts at ``s * spacing + t0``."""
    return ((t0 + n_steps * d_t) - t0) / n_steps


class FDMOperator(Operator):
    """A finite difference method based differential equation solver."""

    def __init__(
        self,
        integrator: NumericalIntegrator,
        differentiator: NumericalDifferentiator,
        d_t: float,
        *,
        precision: str = "float64",
        snapshot_every: int = 1,
        tile_config: Optional[int] = None,
        use_graph: bool = True,
        symbolic_rhs: bool = False,
    ):
        super().__init__(d_t, True)
        dev.precision_id(precision)
        if snapshot_every < 1:
            raise ValueError("snapshot_every must be at least 1")
        self._integrator = integrator
        self._differentiator = differentiator
        self._precision = precision
        self._snapshot_every = int(snapshot_every)
        self._tile_config = tile_config
        self._use_graph = use_graph
        self._symbolic_rhs = bool(symbolic_rhs)

    @property
    def integrator(self) -> NumericalIntegrator:
        return self._integrator

    @property
    def differentiator(self) -> NumericalDifferentiator:
        return self._differentiator

    @property
    def precision(self) -> str:
        return self._precision

    def device_problem(self, cp) -> "dev.DeviceProblem":
        """Validates that ``cp`` runs on the fused kernels and describes it."""
        if not isinstance(self._differentiator, ThreePointCentralDifferenceMethod):
            raise NotImplementedError(
                "the device path implements ThreePointCentralDifferenceMethod"
            )
        dev.integrator_id(self._integrator)
        return dev.device_problem(cp)

    def tile_config_for(self, problem, n_grids: int = 1) -> int:
        """The kernel variant for a batch of ``n_grids`` grids of
        ``problem``: the one given to the constructor, else the tuned one for
        the batch's size class (``tune``), else the default."""
        if self._tile_config is not None:
            return int(self._tile_config)
        from pararealml_amd.autotune import lookup

        return lookup(problem, dev.integrator_id(self._integrator), self._precision,
                      cells=problem.plane * int(n_grids))

    def tune(self, cp, n_grids: int = 1, n_steps: int = 64, seed: int = 0) -> int:
        """Times every kernel variant on a scratch batch of ``n_grids``
        copies of ``cp``'s grid (random state, this operator's time step)
        and records the fastest for the batch's size class; returns it."""
        import torch

        from pararealml_amd.autotune import tune

        problem = self.device_problem(cp)
        rng = np.random.RandomState(seed)
        y = 1e-3 * rng.standard_normal(tuple(cp.mesh.vertices_shape) + (problem.n_fields,))
        stream = dev.solver_stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            batch = dev.FDMBatch([problem], [0] * int(n_grids), self._precision,
                                 dev.integrator_id(self._integrator), stream)
            batch.load_host([y] * int(n_grids))
            best, _ = tune(batch, self.d_t, n_steps=n_steps)
        torch.cuda.current_stream().wait_stream(stream)
        return best

    def slice_propagator(self, cp, n_slices, slice_steps, stream, record_trajectory=False,
                         borders=None):
        """Device propagator over Parareal time slices (see
        ``pararealml_amd.operators.parareal.propagators``); ``borders`` are
        the slice boundary times (needed by systems whose right-hand side
        depends on t)."""
        from pararealml_amd.operators.parareal.propagators import (
            FDMSlicePropagator,
            ODESlicePropagator,
        )

        if not cp.differential_equation.x_dimension:
            return ODESlicePropagator(self, cp, n_slices, slice_steps, stream,
                                      record_trajectory, borders)
        return FDMSlicePropagator(self, cp, n_slices, slice_steps, stream, record_trajectory)

    def solve(self, ivp: InitialValueProblem, parallel_enabled: bool = True) -> Solution:
        cp = ivp.constrained_problem
        if not cp.differential_equation.x_dimension:
            return self._solve_ode(ivp)
        if (
            cp.mesh.coordinate_system_type != CoordinateSystem.CARTESIAN
            or cp.mesh.dimensions > 2
            or not cp.are_all_boundary_conditions_static
            or isinstance(self._integrator, ImplicitMethod)
            or self._symbolic_rhs
            or cp.differential_equation.device_parameters() is None
        ):
            return self._solve_composed(ivp)
        problem = self.device_problem(cp)
        integ = dev.integrator_id(self._integrator)
        t = discretize_time_domain(ivp.t_interval, self._d_t)
        n_steps = len(t) - 1
        every = self._snapshot_every
        if n_steps % every:
            raise ValueError(
                f"snapshot_every ({every}) must divide the number of steps ({n_steps})"
            )
        n_snap = n_steps // every
        y0 = ivp.initial_condition.discrete_y_0(True)
        if n_snap == 0:
            return Solution(
                ivp, t[1:], np.empty((0,) + cp.y_vertices_shape), vertex_oriented=True,
                d_t=self._d_t,
            )

        import torch

        if problem.y_laplacian:
            return self._solve_with_poisson(ivp, problem, integ, t, n_steps, every, y0)
        stream = dev.solver_stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            batch = dev.FDMBatch([problem], [0], self._precision, integ, stream)
            batch.load_host([y0])
            snaps = batch.new_snapshots(n_snap)
            batch.advance(
                self._d_t,
                n_steps,
                snaps,
                every,
                tile_config=self.tile_config_for(problem),
                use_graph=self._use_graph,
            )
            dev_y = batch.snapshots_f64(snaps, n_snap)
            flat = dev_y.cpu().numpy()
        torch.cuda.current_stream().wait_stream(stream)
        flat = flat.reshape(n_snap, problem.n_fields, problem.n0, problem.n1)
        y = problem.from_device_layout(flat, cp.mesh.vertices_shape)
        return DeviceSolution(
            ivp,
            t[every::every],
            y,
            dev_y,
            vertex_oriented=True,
            d_t=self._d_t * every,
            field_major=True,
        )

    def _ode_precision(self) -> int:
        if self._precision not in ("float64", "float32"):
            raise NotImplementedError("ODE systems run in float64 or float32")
        return nat.PML_F64 if self._precision == "float64" else nat.PML_F32

    def _ode_program(self, diff_eq):
        """The compiled right-hand side of an ODE system (cached per
        equation object)."""
        from pararealml_amd.operators.device_program import ode_program

        key = id(diff_eq)
        cache = self.__dict__.setdefault("_programs", {})
        if key not in cache or cache[key][0] is not diff_eq:
            cache[key] = (diff_eq, ode_program(diff_eq))
        return cache[key][1]

    def _solve_ode(self, ivp: InitialValueProblem) -> Solution:
        """FDMOperator.solve on a system without a mesh (fdm_operator.py:
        48-77; the constraint functions return None, 167-172): the compiled
        right-hand side and the explicit integrator run the whole time loop
        in one launch of the ODE kernel, writing every snapshot_every-th
        state."""
        import torch

        if isinstance(self._integrator, ImplicitMethod):
            raise NotImplementedError(
                "ODE systems run the explicit integrators (ForwardEulerMethod, "
                "ExplicitMidpointMethod, RK4) on the device"
            )
        integ = dev.integrator_id(self._integrator)
        prec = self._ode_precision()
        cp = ivp.constrained_problem
        n = cp.differential_equation.y_dimension
        t = discretize_time_domain(ivp.t_interval, self._d_t)
        n_steps = len(t) - 1
        every = self._snapshot_every
        if n_steps % every:
            raise ValueError(
                f"snapshot_every ({every}) must divide the number of steps ({n_steps})"
            )
        n_snap = n_steps // every
        if n_snap == 0:
            return Solution(ivp, t[1:], np.empty((0, n)), vertex_oriented=True, d_t=self._d_t)
        program = self._ode_program(cp.differential_equation)
        dtype = torch.float64 if prec == nat.PML_F64 else torch.float32
        y0 = np.asarray(ivp.initial_condition.discrete_y_0(True), dtype=np.float64)
        stream = dev.solver_stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            y = torch.from_numpy(y0.reshape(1, n)).to(device="cuda", dtype=dtype)
            snaps = torch.empty((n_snap, 1, n), dtype=dtype, device="cuda")
            t0 = torch.tensor([t[0]], dtype=torch.float64, device="cuda")
            step = torch.tensor([ode_time_step(t[0], n_steps, self._d_t)],
                                dtype=torch.float64, device="cuda")
            lib = nat.device_lib()
            nat.check(lib.pml_ode_solve(prec, program.handle(), integ, 1, t0.data_ptr(),
                                        step.data_ptr(), float(self._d_t), int(n_steps),
                                        y.data_ptr(), snaps.data_ptr(), int(every),
                                        stream.cuda_stream))
            dev_y = snaps.view(n_snap, n).to(torch.float64)
            out = dev_y.cpu().numpy()
        torch.cuda.current_stream().wait_stream(stream)
        return DeviceSolution(ivp, t[every::every], out, dev_y, vertex_oriented=True,
                              d_t=self._d_t * every)

    def _solve_composed(self, ivp: InitialValueProblem) -> Solution:
        """Polar, cylindrical and spherical meshes, 3-D Cartesian ones and
        time-dependent boundary conditions (fdm_operator.py:48-165, 199-231
        with the differentiator's curvilinear branches), which the fused
        kernels do not cover.  The state stays on
        the device in the reference's channel-last layout; each stage's
        right-hand side is composed from the device operators of
        ThreePointCentralDifferenceMethod (the Laplacian of component k with
        the derivative constraints of component k, as
        fdm_symbol_mapper.py:120-125), the stage sums run through
        ``pml_weighted_add`` and the static y constraints through
        ``pml_set_values``, with the stage arithmetic of
        numerical_integrator.py:42-132.  Covers equations whose right-hand
        side the loop composes (diffusion, convection-diffusion, wave,
        shallow water, Burgers, Cahn-Hilliard with its LHS.Y potential,
        Navier-Stokes with its Poisson phase) in FP64 or FP32; other
        equations raise ``NotImplementedError``.  Called directly, it also
        runs Cartesian meshes (the tests compare it with the fused kernels)."""
        import torch

        cp = ivp.constrained_problem
        eq = cp.differential_equation
        if not isinstance(self._differentiator, ThreePointCentralDifferenceMethod):
            raise NotImplementedError(
                "the device path implements ThreePointCentralDifferenceMethod"
            )
        params = eq.device_parameters()
        family = params[0] if params else None
        if self._symbolic_rhs or family is None:
            family = "symbolic"
        if family not in ("diffusion", "wave", "shallow_water", "burgers", "symbolic",
                          "convection_diffusion", "cahn_hilliard", "navier_stokes"):
            raise NotImplementedError(
                "curvilinear and 3-D meshes run the diffusion, convection-diffusion, "
                "wave, shallow-water, Burgers, Cahn-Hilliard and Navier-Stokes equations; "
                f"got {type(eq).__name__}"
            )
        if self._precision not in ("float64", "float32"):
            raise NotImplementedError("curvilinear meshes run in float64 or float32")
        implicit = isinstance(self._integrator, ImplicitMethod)
        if implicit and self._precision != "float64":
            # the secant iteration's absolute tolerance (1.48e-8 by default)
            # is below FP32 rounding for O(1) values
            raise NotImplementedError("implicit integrators run in float64")
        integ = None if implicit else dev.integrator_id(self._integrator)
        t = discretize_time_domain(ivp.t_interval, self._d_t)
        n_steps = len(t) - 1
        every = self._snapshot_every
        if n_steps % every:
            raise ValueError(
                f"snapshot_every ({every}) must divide the number of steps ({n_steps})"
            )
        n_snap = n_steps // every
        if n_snap == 0:
            return Solution(
                ivp, t[1:], np.empty((0,) + cp.y_vertices_shape), vertex_oriented=True,
                d_t=self._d_t,
            )
        mesh = cp.mesh
        diff = self._differentiator
        dtype = torch.float64 if self._precision == "float64" else torch.float32
        prec = nat.PML_F64 if dtype == torch.float64 else nat.PML_F32
        shape = tuple(cp.y_vertices_shape)
        count = int(np.prod(shape))
        lib = nat.device_lib()

        dynamic = not cp.are_all_boundary_conditions_static

        def encode(y_constraints):
            """Full-grid y constraints as a device FP64 array, NaN where free."""
            full = np.full(shape, np.nan)
            for f, c in enumerate(y_constraints if y_constraints is not None else []):
                if c is not None:
                    full[..., f : f + 1][..., c.mask] = c.values
            return torch.from_numpy(full.ravel()).to("cuda")

        def no_dbc(d):
            return d if d is not None else np.empty((mesh.dimensions, shape[-1]), dtype=object)

        static = None
        if not dynamic:
            static = (encode(cp.static_y_vertex_constraints),
                      no_dbc(cp.static_boundary_vertex_constraints[1]),
                      cp.static_y_vertex_constraints)
        cache = {}

        def constraints_at(t_c):
            """(device y constraints, derivative constraints, host y
            constraints) at time t_c; dynamic conditions are evaluated once per
            time and step (fdm_operator.py:199-231)."""
            if not dynamic:
                return static
            if t_c not in cache:
                y_bc, d_bc = cp.create_boundary_constraints(True, t_c)
                host = cp.create_y_vertex_constraints(y_bc)
                cache[t_c] = (encode(host), no_dbc(d_bc), host)
            return cache[t_c]

        def sptr():
            return torch.cuda.current_stream().cuda_stream

        def full_of(v):
            return torch.full(shape, float(v), dtype=dtype, device="cuda")

        def axpy(out, x, coef, mode):
            """out (=, +=) coef * x on the device (coef: a full tensor)."""
            nat.check(lib.pml_weighted_add(prec, x.data_ptr(), coef.data_ptr(), out.data_ptr(),
                                           count, mode, sptr()))

        def constrain(y, t_c):
            ycons = constraints_at(t_c)[0]
            nat.check(lib.pml_set_values(prec, y.data_ptr(), ycons.data_ptr(), count, sptr()))
            return y

        if family != "symbolic":
            coef_p = full_of(params[1][0])
            coef_f = torch.full(shape[:-1] + (1,), float(params[1][0]), dtype=dtype,
                                device="cuda")

        def comp(y, k):
            return y[..., k : k + 1].contiguous()

        def field_full(v):
            return torch.full(shape[:-1] + (1,), float(v), dtype=dtype, device="cuda")

        def sub_sum(out, terms):
            """out = sum of sign * coef * x over (coef, x, sign): coef is a
            field-sized tensor (a constant or a field), the first term is
            positive and writes, the rest add or subtract."""
            for i, (c, x, sign) in enumerate(terms):
                mode = 0 if i == 0 else sign
                nat.check(lib.pml_weighted_add(prec, x.data_ptr(), c.data_ptr(), out.data_ptr(),
                                               out.numel(), mode, sptr()))

        def rhs_shallow_water(y, dbc):
            """differential_equation.py:797-819 (h, b, v, f, g)."""
            ys = [comp(y, k) for k in range(3)]
            grad = [[diff.gradient(ys[k], mesh, a, dbc[:, k : k + 1]) for a in range(2)]
                    for k in range(3)]
            div = diff.divergence(y[..., 1:3].contiguous(), mesh, dbc[:, 1:3])
            lap = [None] + [diff.laplacian(ys[k], mesh, dbc[:, k : k + 1]) for k in (1, 2)]
            outs = [torch.empty_like(ys[0]) for _ in range(3)]
            sub_sum(outs[0], [(c_mh, div, 1), (ys[0], grad[1][0], -1), (ys[1], grad[0][0], -1),
                              (ys[0], grad[2][1], -1), (ys[2], grad[0][1], -1)])
            sub_sum(outs[1], [(c_v, lap[1], 1), (ys[1], grad[1][0], -1), (ys[2], grad[1][1], -1),
                              (c_mg, grad[0][0], 1), (c_mb, ys[1], 1), (c_f, ys[2], 1)])
            sub_sum(outs[2], [(c_v, lap[2], 1), (ys[1], grad[2][0], -1), (ys[2], grad[2][1], -1),
                              (c_mg, grad[0][1], 1), (c_mb, ys[2], 1), (c_mf, ys[1], 1)])
            return torch.cat(outs, dim=-1).contiguous()

        if family == "shallow_water":
            h_, b_, v_, f_, g_ = params[1]
            c_mh, c_v, c_mg, c_mb, c_f, c_mf = (field_full(x) for x in
                                                (-h_, v_, -g_, -b_, f_, -f_))

        def rhs_burgers(y, dbc):
            """differential_equation.py:757-765: (1/re) lap(y_i) - y . grad(y_i)."""
            n = mesh.dimensions
            ys = [comp(y, k) for k in range(n)]
            outs = []
            for i in range(n):
                terms = [(c_re, diff.laplacian(ys[i], mesh, dbc[:, i : i + 1]), 1)]
                for j in range(n):
                    terms.append((ys[j], diff.gradient(ys[i], mesh, j, dbc[:, i : i + 1]), -1))
                out = torch.empty_like(ys[0])
                sub_sum(out, terms)
                outs.append(out)
            return torch.cat(outs, dim=-1).contiguous()

        if family == "burgers":
            c_re = field_full(params[1][0])

        def rhs_convection_diffusion(y, dbc):
            """differential_equation.py:633-670: d lap(y) - velocity . grad(y)."""
            terms = [(c_d, diff.laplacian(y, mesh, dbc), 1)]
            for j, c_vj in enumerate(c_vel):
                terms.append((c_vj, diff.gradient(y, mesh, j, dbc), -1))
            out = torch.empty_like(y)
            sub_sum(out, terms)
            return out

        if family == "convection_diffusion":
            c_d = field_full(params[1][0])
            c_vel = [field_full(v) for v in eq._velocity]

        def rhs_cahn_hilliard(y, dbc):
            """differential_equation.py:723-732: c_t = d lap(mu); mu (LHS.Y)
            has no time derivative inside the step."""
            lap_mu = diff.laplacian(comp(y, 1), mesh, dbc[:, 1:2])
            dc = torch.empty_like(lap_mu)
            sub_sum(dc, [(c_chd, lap_mu, 1)])
            return torch.cat([dc, torch.zeros_like(dc)], dim=-1).contiguous()

        def with_mu(nxt, y_pre, t_pre, t_next):
            """fdm_operator.py:127-139: mu_next = C_{t_next}(c^3 - c - gamma
            lap(c)) from the pre-step c, derivative constraints at t_pre; the
            other components are left as the integrator produced them."""
            c = comp(y_pre, 0)
            lap_c = diff.laplacian(c, mesh, constraints_at(t_pre)[1][:, :1])
            c2 = torch.empty_like(c)
            sub_sum(c2, [(c, c, 1)])
            mu = torch.empty_like(c)
            sub_sum(mu, [(c2, c, 1), (c_ones_f, c, -1), (c_gamma, lap_c, -1)])
            tmp = nxt.clone()
            tmp[..., 1:2].copy_(mu)
            constrain(tmp, t_next)
            nxt[..., 1:2].copy_(tmp[..., 1:2])
            return nxt

        def rhs_navier_stokes(y, dbc):
            """differential_equation.py:835-850, vorticity row: (1/re) lap(w)
            - (u, v) . grad(w); psi (LHS.Y_LAPLACIAN), u and v (LHS.Y) have
            no time derivative inside the step."""
            w = comp(y, 0)
            terms = [(c_re_ns, diff.laplacian(w, mesh, dbc[:, 0:1]), 1)]
            for j in range(2):
                terms.append((comp(y, 2 + j), diff.gradient(w, mesh, j, dbc[:, 0:1]), -1))
            dw = torch.empty_like(w)
            sub_sum(dw, terms)
            return torch.cat([dw] + [torch.zeros_like(dw)] * 3, dim=-1).contiguous()

        def with_psi_uv(nxt, y_pre, t_pre, t_next):
            """fdm_operator.py:127-161 for Navier-Stokes: psi solves
            lap(psi) = -w_pre (Jacobi to the differentiator's tol, constraints
            at t_next, started from the pre-step psi), u = d psi / d x1 and
            v = -d psi / d x0 from the pre-step psi (derivative constraints at
            t_pre), all constrained at t_next."""
            d_pre = constraints_at(t_pre)[1]
            d_next = constraints_at(t_next)[1]
            host_next = constraints_at(t_next)[2]
            psi = comp(y_pre, 1)
            u = diff.gradient(psi, mesh, 1, d_pre[:, 1:2])
            v = torch.empty_like(psi)
            sub_sum(v, [(c_mone, diff.gradient(psi, mesh, 0, d_pre[:, 1:2]), 1)])
            src = torch.empty_like(psi)
            sub_sum(src, [(c_mone, comp(y_pre, 0), 1)])
            psi_next = diff.anti_laplacian(
                src, mesh, None if host_next is None else host_next[1:2], d_next[:, 1:2],
                y_init=psi)
            tmp = nxt.clone()
            tmp[..., 1:2].copy_(psi_next)
            tmp[..., 2:3].copy_(u)
            tmp[..., 3:4].copy_(v)
            constrain(tmp, t_next)
            nxt[..., 1:4].copy_(tmp[..., 1:4])
            return nxt

        if family == "navier_stokes":
            c_re_ns = field_full(params[1][0])
            c_mone = field_full(-1.0)

        if family == "cahn_hilliard":
            c_chd = field_full(params[1][0])
            c_gamma = field_full(params[1][1])
            c_ones_f = field_full(1.0)

        if family == "symbolic":
            from pararealml_amd.differential_equation import LHS

            system = eq.symbolic_equation_system
            d_idx = list(system.equation_indices_by_type(LHS.D_Y_OVER_D_T))
            y_idx = list(system.equation_indices_by_type(LHS.Y))
            lap_idx = list(system.equation_indices_by_type(LHS.Y_LAPLACIAN))
            sym_d = _SymbolicDeviceRHS(cp, diff, d_idx, dtype)
            sym_y = _SymbolicDeviceRHS(cp, diff, y_idx, dtype) if y_idx else None
            sym_l = _SymbolicDeviceRHS(cp, diff, lap_idx, dtype) if lap_idx else None

        def rhs_symbolic(y, t_c, dbc):
            """d y / d t: zero except at the D_Y_OVER_D_T components
            (fdm_operator.py:114-120)."""
            out = torch.zeros_like(y)
            sym_d(y, t_c, dbc, out, d_idx)
            return out

        def with_symbolic_lhs(nxt, y_pre, t_pre, t_next):
            """fdm_operator.py:127-161: LHS.Y components = C_{t_next}(their
            right-hand sides at the pre-step state and time); LHS.Y_LAPLACIAN
            components solve lap(y_k) = rhs_k (pre-step state) by the device
            Jacobi iteration with the constraints at t_next, started from
            the pre-step values."""
            dbc_pre = constraints_at(t_pre)[1]
            if sym_y is not None:
                tmp = nxt.clone()
                sym_y(y_pre, t_pre, dbc_pre, tmp, y_idx)
                constrain(tmp, t_next)
                nxt[..., y_idx] = tmp[..., y_idx]
            if sym_l is not None:
                lap = torch.empty(shape[:-1] + (len(lap_idx),), dtype=dtype, device="cuda")
                sym_l(y_pre, t_pre, dbc_pre, lap, range(len(lap_idx)))
                d_next = constraints_at(t_next)[1]
                host_next = constraints_at(t_next)[2]
                ycons = None if host_next is None else \
                    np.asarray(host_next, dtype=object)[lap_idx]
                nxt[..., lap_idx] = diff.anti_laplacian(
                    lap, mesh, ycons, d_next[:, lap_idx],
                    y_init=y_pre[..., lap_idx].contiguous())
            return nxt

        def rhs(y, t_c):
            dbc = constraints_at(t_c)[1]
            if family == "symbolic":
                return rhs_symbolic(y, t_c, dbc)
            if family == "shallow_water":
                return rhs_shallow_water(y, dbc)
            if family == "burgers":
                return rhs_burgers(y, dbc)
            if family == "convection_diffusion":
                return rhs_convection_diffusion(y, dbc)
            if family == "cahn_hilliard":
                return rhs_cahn_hilliard(y, dbc)
            if family == "navier_stokes":
                return rhs_navier_stokes(y, dbc)
            if family == "diffusion":
                lap = diff.laplacian(y, mesh, dbc)
                out = torch.empty_like(y)
                axpy(out, lap, coef_p, 0)
                return out
            # wave: (y1, c^2 lap(y0))
            lap = diff.laplacian(y[..., :1].contiguous(), mesh, dbc[:, :1])
            acc = torch.empty_like(lap)
            nat.check(lib.pml_weighted_add(prec, lap.data_ptr(), coef_f.data_ptr(),
                                           acc.data_ptr(), lap.numel(), 0, sptr()))
            return torch.cat([y[..., 1:2], acc], dim=-1).contiguous()

        dt = self._d_t
        c_dt, c_half, c_one, c_two, c_sixth = (full_of(v) for v in
                                                (dt, 0.5, 1.0, 2.0, 1.0 / 6.0))
        y = torch.from_numpy(np.ascontiguousarray(ivp.initial_condition.discrete_y_0(True))) \
            .to(device="cuda", dtype=dtype)
        if dynamic:
            y = constrain(y, t[0])  # fdm_operator.py:55-63
        half_dt = dt / 2.0
        if implicit:
            im = self._integrator
            c_adt = full_of(im.a * dt)
            c_b = full_of(1.0 - im.a)
            counters = torch.zeros(2, dtype=torch.int64, device="cuda")

            def secant(p, residual):
                """scipy.optimize.newton's array secant branch (no
                derivative): the reference's ImplicitMethod._solve
                (numerical_integrator.py:154-176) on the device, with SciPy's
                exits: all residual differences zero -> midpoint and stop;
                no element with a nonzero difference still moving by tol or
                more -> stop; afterwards every element whose persistent
                failure flag is still set counts as not converged (all ->
                RuntimeError, some -> RuntimeWarning, as SciPy)."""
                p1 = torch.empty_like(p)
                nat.check(lib.pml_secant_step(prec, 2, p.data_ptr(), None, None, None,
                                              p1.data_ptr(), None, None, count,
                                              float(np.finfo(float).eps ** 0.33), None, sptr()))
                q0, q1 = residual(p), residual(p1)
                active = torch.ones(shape, dtype=torch.uint8, device="cuda")
                failing = torch.ones(shape, dtype=torch.uint8, device="cuda")
                for _ in range(im.max_iterations):
                    counters.zero_()
                    p_new = torch.empty_like(p)
                    nat.check(lib.pml_secant_step(
                        prec, 0, p.data_ptr(), p1.data_ptr(), q0.data_ptr(), q1.data_ptr(),
                        p_new.data_ptr(), active.data_ptr(), failing.data_ptr(), count,
                        float(im.tol), counters.data_ptr(), sptr()))
                    n_nz, n_fail_nz = (int(v) for v in counters.tolist())
                    if n_nz == 0:
                        nat.check(lib.pml_secant_step(prec, 1, p.data_ptr(), p1.data_ptr(),
                                                      None, None, p_new.data_ptr(), None, None,
                                                      count, 0.0, None, sptr()))
                        p = p_new
                        break
                    p = p_new
                    if n_fail_nz == 0:
                        break
                    p1, p = p, p1
                    q0 = q1
                    q1 = residual(p1)
                n_failed = int(failing.sum())
                if n_failed:
                    msg = (f"{'all' if n_failed == count else 'some'} failed to converge "
                           f"after {im.max_iterations} iterations")
                    if n_failed == count:
                        raise RuntimeError(msg)
                    warnings.warn(msg, RuntimeWarning, stacklevel=2)
                return p

        snaps = []
        for step in range(n_steps):
            t_i = t[step]
            t_h, t_n = t_i + half_dt, t_i + dt
            k1 = torch.empty_like(y)
            axpy(k1, rhs(y, t_i), c_dt, 0)
            if implicit:
                # numerical_integrator.py:194-270: start from the explicit
                # step, solve y_next = C(y + a dt f(t_n, y_next) + (1-a) k1)
                init = y.clone()
                axpy(init, k1, c_one, 1)
                init = constrain(init, t_n)

                def residual(y_next, y=y, k1=k1, t_n=t_n):
                    v = y.clone()
                    axpy(v, rhs(y_next, t_n), c_adt, 1)
                    if im.a != 1.0:
                        axpy(v, k1, c_b, 1)
                    constrain(v, t_n)
                    r = y_next.clone()
                    axpy(r, v, c_one, -1)
                    return r

                nxt = secant(init, residual)
                if family == "cahn_hilliard":
                    nxt = with_mu(nxt, y, t_i, t_n)
                if family == "navier_stokes":
                    nxt = with_psi_uv(nxt, y, t_i, t_n)
                if family == "symbolic":
                    nxt = with_symbolic_lhs(nxt, y, t_i, t_n)
                y = nxt
                cache.clear()
                if (step + 1) % every == 0:
                    snaps.append(y.clone())
                continue
            if integ == nat.PML_INT_FE:
                nxt = y.clone()
                axpy(nxt, k1, c_one, 1)
            elif integ == nat.PML_INT_MIDPOINT:
                half = y.clone()
                axpy(half, k1, c_half, 1)
                k2 = torch.empty_like(y)
                axpy(k2, rhs(constrain(half, t_h), t_h), c_dt, 0)
                nxt = y.clone()
                axpy(nxt, k2, c_one, 1)
            else:
                st = y.clone()
                axpy(st, k1, c_half, 1)
                k2 = torch.empty_like(y)
                axpy(k2, rhs(constrain(st, t_h), t_h), c_dt, 0)
                st = y.clone()
                axpy(st, k2, c_half, 1)
                k3 = torch.empty_like(y)
                axpy(k3, rhs(constrain(st, t_h), t_h), c_dt, 0)
                st = y.clone()
                axpy(st, k3, c_one, 1)
                k4 = torch.empty_like(y)
                axpy(k4, rhs(constrain(st, t_n), t_n), c_dt, 0)
                acc = k1.clone()
                axpy(acc, k2, c_two, 1)
                axpy(acc, k3, c_two, 1)
                axpy(acc, k4, c_one, 1)
                nxt = y.clone()
                axpy(nxt, acc, c_sixth, 1)
            if family == "cahn_hilliard":
                nxt = with_mu(nxt, y, t_i, t_n)
            if family == "navier_stokes":
                nxt = with_psi_uv(nxt, y, t_i, t_n)
            if family == "symbolic":
                nxt = with_symbolic_lhs(nxt, y, t_i, t_n)
            y = constrain(nxt, t_n)
            cache.clear()
            if (step + 1) % every == 0:
                snaps.append(y.clone())
        dev_y = torch.stack(snaps).to(torch.float64)
        out = dev_y.cpu().numpy()
        return DeviceSolution(ivp, t[every::every], out, dev_y, vertex_oriented=True,
                              d_t=dt * every)

    def _solve_with_poisson(self, ivp, problem, integ, t, n_steps, every, y0) -> Solution:
        """The time loop for equations with LHS.Y_LAPLACIAN components
        (fdm_operator.py:141-161): each step runs the fused kernel (the
        D_Y_OVER_D_T fields advance, the LHS.Y fields are recomputed from the
        pre-step state, the Y_LAPLACIAN fields stay), then every
        Y_LAPLACIAN field is re-solved from the pre-step state's right-hand
        side by the differentiator's device Jacobi iteration, with the
        field's own constraints.  The iteration starts from the field's
        current value; the reference starts from uniform random values, so
        its result is defined only up to the Jacobi tolerance, which bounds
        the difference between the two."""
        import torch

        if self._precision == "float32c":
            raise NotImplementedError(
                "LHS.Y_LAPLACIAN equations run in float32 or float64"
            )
        cp = ivp.constrained_problem
        mesh = cp.mesh
        plane = problem.plane
        shape = (problem.n0, problem.n1, 1)
        y_cons = cp.static_y_vertex_constraints
        d_cons = cp.static_boundary_vertex_constraints[1]
        n_snap = n_steps // every
        stream = dev.solver_stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream):
            batch = dev.FDMBatch([problem], [0], self._precision, integ, stream,
                                 poisson_phase=True)
            batch.load_host([y0])
            traj = torch.empty((n_snap, problem.elems), dtype=torch.float64, device="cuda")
            tile = self.tile_config_for(problem)
            state = batch.hi  # field-major, the batch precision
            for i in range(n_steps):
                sources = {
                    f: (c * state[src * plane:(src + 1) * plane]).view(shape)
                    for f, src, c in problem.y_laplacian
                }
                batch.advance(self._d_t, 1, tile_config=tile, use_graph=False)
                for f, lap in sources.items():
                    cur = state[f * plane:(f + 1) * plane].view(shape)
                    new = self._differentiator.anti_laplacian(
                        lap, mesh,
                        [None if y_cons is None else y_cons[f]],
                        None if d_cons is None else d_cons[:, f:f + 1],
                        y_init=cur,
                    )
                    cur.copy_(new)
                if (i + 1) % every == 0:
                    batch.state_f64(traj[(i + 1) // every - 1])
            flat = traj.cpu().numpy()
        torch.cuda.current_stream().wait_stream(stream)
        flat = flat.reshape(n_snap, problem.n_fields, problem.n0, problem.n1)
        y = problem.from_device_layout(flat, mesh.vertices_shape)
        return DeviceSolution(ivp, t[every::every], y, traj, vertex_oriented=True,
                              d_t=self._d_t * every, field_major=True)
