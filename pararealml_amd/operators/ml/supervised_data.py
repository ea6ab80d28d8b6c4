"""The supervised-learning data protocol of PararealML, with the oracle's
solves batched on the device.

``SupervisedDataProtocol`` carries the flags of SupervisedMLOperator
(pararealml/operators/ml/supervised/supervised_ml_operator.py:20-52: step,
orientation, auto-regression, time as a predictor, step as an input) and its
``generate_data`` (:152-245, worker :381-499, :501-551): sub-IVPs of one
operator step solved by an FDM oracle from perturbed initial states, laid
out as the reference's (inputs, targets) rows.  The regression models of the
reference's operator (scikit-learn / Keras) are not part of this package.

On the device every oracle solve of a round is one batched launch sequence
(``FDMBatch``): with ``isolate_perturbations``, or without auto-regression,
all perturbations are drawn first, in the reference's order, and every
sub-IVP of every iteration is solved in one batch; with perturbations that
propagate along the chain, the iterations advance together one operator
step at a time.  The rows are written by ``pml_supervised_rows``.

Random perturbations: the reference runs each job in its own process (its
own copy of NumPy's global random state, seeded by ``seeds[job]`` when
given) and draws iteration by iteration.  Here each job's random state is
kept and restored around its draws, so with one job, or with
``isolate_perturbations``, or without auto-regression, the draws -- and the
data -- are the reference's; with perturbations that propagate along the
chain the draws of a job are taken step by step across its iterations, a
different order of the same distribution.
"""

import warnings
from typing import Callable, List, Optional, Sequence, Tuple

import ctypes
import numpy as np

from pararealml_amd import _device as dev
from pararealml_amd import _native as nat
from pararealml_amd.constraint import apply_constraints_along_last_axis
from pararealml_amd.initial_value_problem import InitialValueProblem
from pararealml_amd.operator import discretize_time_domain


class SupervisedDataProtocol:
    """Data generation of SupervisedMLOperator (see the module docstring)."""

    def __init__(
        self,
        d_t: float,
        vertex_oriented: bool,
        auto_regressive: bool = True,
        time_variant: bool = False,
        input_d_t: bool = False,
    ):
        if not auto_regressive and not time_variant:
            raise ValueError("operator must be time variant if auto-regression is disabled")
        if time_variant and input_d_t:
            raise ValueError("operator must be time invariant to use d_t as an input")
        if d_t <= 0.0:
            raise ValueError("time step size must be greater than 0")
        self._d_t = d_t
        self._vertex_oriented = vertex_oriented
        self._auto_regressive = auto_regressive
        self._time_variant = time_variant
        self._input_d_t = input_d_t

    @property
    def d_t(self) -> float:
        return self._d_t

    @property
    def vertex_oriented(self) -> bool:
        return self._vertex_oriented

    @property
    def auto_regressive(self) -> bool:
        return self._auto_regressive

    @property
    def time_variant(self) -> bool:
        return self._time_variant

    @property
    def input_d_t(self) -> bool:
        return self._input_d_t

    def generate_data(
        self,
        ivp: InitialValueProblem,
        oracle,
        iterations: int,
        perturbation_function: Callable[[float, np.ndarray], np.ndarray],
        isolate_perturbations: bool = False,
        repeat_on_error: bool = False,
        n_jobs: int = 1,
        seeds: Optional[Sequence[int]] = None,
    ) -> Tuple[np.ndarray, np.ndarray]:
        """(inputs, targets) as host arrays (supervised_ml_operator.py:152-245)."""
        inputs, targets = self.generate_data_device(
            ivp, oracle, iterations, perturbation_function, isolate_perturbations,
            repeat_on_error, n_jobs, seeds,
        )
        return inputs.cpu().numpy(), targets.cpu().numpy()

    def generate_data_device(
        self,
        ivp: InitialValueProblem,
        oracle,
        iterations: int,
        perturbation_function: Callable[[float, np.ndarray], np.ndarray],
        isolate_perturbations: bool = False,
        repeat_on_error: bool = False,
        n_jobs: int = 1,
        seeds: Optional[Sequence[int]] = None,
    ):
        """``generate_data`` with the (inputs, targets) rows left in device
        memory (FP64 tensors)."""
        import torch

        if iterations <= 0:
            raise ValueError("number of iterations must be greater than 0")
        if n_jobs < 1:
            raise ValueError("number of jobs must be greater than 0")
        if seeds is not None:
            if len(seeds) != n_jobs:
                raise ValueError(
                    f"number of seeds ({len(seeds)}) must match number of jobs ({n_jobs})"
                )
        else:
            seeds = [None] * n_jobs
        cp = ivp.constrained_problem
        diff_eq = cp.differential_equation
        if not diff_eq.x_dimension:
            raise NotImplementedError("the device data protocol covers PDE problems")
        if not self._vertex_oriented:
            raise NotImplementedError(
                "the device data protocol covers vertex-oriented operators (the FDM oracle's)"
            )
        problem = oracle.device_problem(cp)

        t = discretize_time_domain(ivp.t_interval, self._d_t)
        n_steps = len(t) - 1
        if n_steps < 1:
            raise ValueError("the IVP's time domain holds no operator step")
        y_0 = ivp.initial_condition.discrete_y_0(True)
        cons = cp.static_y_vertex_constraints

        # per job: its iteration count and its random state (a fresh copy of
        # the current one, seeded when a seed is given)
        counts = [len(a) for a in np.array_split(np.arange(iterations), n_jobs)]
        start_state = np.random.get_state()
        states = []
        for seed in seeds:
            np.random.set_state(start_state)
            if seed is not None:
                np.random.seed(seed)
            states.append(np.random.get_state())

        def perturb(t_i: float, y: np.ndarray) -> np.ndarray:
            y_c = np.array(y, dtype=np.float64, copy=True)
            apply_constraints_along_last_axis(cons, y_c)
            p = perturbation_function(t_i, y_c)
            if p.shape != y_c.shape:
                raise ValueError(
                    f"perturbed y shape {p.shape} must match input y shape {y_c.shape}"
                )
            p = np.array(p, dtype=np.float64, copy=True)
            apply_constraints_along_last_axis(cons, p)
            return p

        def draw(job: int, calls):
            """Runs ``calls`` (a list of zero-argument callables) with the
            job's random state."""
            np.random.set_state(states[job])
            try:
                return [c() for c in calls]
            finally:
                states[job] = np.random.get_state()

        n_sub = int(round(self._d_t / oracle.d_t))
        stream = dev.solver_stream()
        stream.wait_stream(torch.cuda.current_stream())
        try:
            with torch.cuda.stream(stream):
                if self._auto_regressive and isolate_perturbations:
                    rec, nxt = self._isolated(ivp, oracle, problem, t, n_sub, y_0, counts,
                                              draw, perturb, repeat_on_error, stream)
                    per_state = 1
                elif self._auto_regressive:
                    rec, nxt = self._chained(cp, oracle, problem, t, n_sub, y_0, counts, draw,
                                             perturb, repeat_on_error, stream)
                    per_state = 1
                else:
                    rec, nxt = self._whole(ivp, oracle, problem, t, y_0, counts, draw,
                                           perturb, repeat_on_error, stream)
                    per_state = n_steps
                inputs, targets = self._rows(cp, t, rec, per_state, nxt, iterations,
                                             n_steps, problem, stream)
        finally:
            # one job runs in the caller's process in the reference (its draws
            # advance the caller's random state); several run in child processes
            np.random.set_state(states[0] if n_jobs == 1 else start_state)
        torch.cuda.current_stream().wait_stream(stream)
        return inputs, targets

    # ------------------------------------------------------------ the solves
    def _solve(self, oracle, problem, ys: List[np.ndarray], n_steps: int, stream,
               repeat_on_error: bool, redraw, every: int = 0, n_snap: int = 0):
        """Solves grids starting from the host states ``ys`` for ``n_steps``
        oracle steps; returns (loaded FP64 states, final FP64 states[,
        snapshots]) as device tensors (grids, elems).  On an error, with
        ``repeat_on_error`` the states are drawn again (``redraw``)."""
        while True:
            try:
                batch = dev.FDMBatch([problem], [0] * len(ys), oracle.precision,
                                     dev.integrator_id(oracle.integrator), stream)
                batch.load_host(ys)
                loaded = batch.state_f64()
                snaps = batch.new_snapshots(n_snap) if n_snap else None
                batch.advance(oracle.d_t, n_steps, snaps, every,
                              tile_config=oracle.tile_config_for(problem, len(ys)),
                              use_graph=True)
                final = batch.state_f64()
                out = (loaded.view(len(ys), -1), final.view(len(ys), -1))
                if n_snap:
                    out += (batch.snapshots_f64(snaps, n_snap),)
                return out
            except Exception as exception:  # noqa: BLE001 -- mirrors the reference
                if not repeat_on_error:
                    raise
                warnings.warn(
                    f"Failed to solve IVP with perturbed initial conditions; {exception}"
                )
                ys = redraw()

    def _isolated(self, ivp, oracle, problem, t, n_sub, y_0, counts, draw, perturb,
                  repeat_on_error, stream):
        """Every sub-IVP of every iteration in one batch; the chain follows
        the unperturbed solution (supervised_ml_operator.py:423-438)."""
        n_steps = len(t) - 1
        starts = [y_0]
        if n_steps > 1:
            cp = ivp.constrained_problem
            head = int(round((t[-2] - t[0]) / oracle.d_t))
            idx = np.rint((t[1:-1] - t[0]) / oracle.d_t).astype(int) - 1
            every, pick = _snapshot_plan(idx, head)
            _, _, snaps = self._solve(oracle, problem, [y_0], head, stream, False, None,
                                      every=every, n_snap=head // every)
            host = snaps[pick].cpu().numpy()
            starts += list(_channel_last(problem, host, cp))

        def all_draws():
            ys = []
            for job, count in enumerate(counts):
                calls = [lambda ti=t_i, y=starts[i]: perturb(ti, y)
                         for _ in range(count) for i, t_i in enumerate(t[:-1])]
                ys += draw(job, calls)
            return ys

        rec, nxt = self._solve(oracle, problem, all_draws(), n_sub, stream, repeat_on_error,
                               all_draws)
        return rec, nxt

    def _chained(self, cp, oracle, problem, t, n_sub, y_0, counts, draw, perturb,
                 repeat_on_error, stream):
        """The iterations advance together, one operator step per round;
        each step starts from the previous step's perturbed solution
        (supervised_ml_operator.py:455-485)."""
        import torch

        n_steps = len(t) - 1
        iterations = sum(counts)
        rec = torch.empty((iterations, n_steps, problem.elems), dtype=torch.float64,
                          device="cuda")
        nxt = torch.empty_like(rec)
        cur = [y_0] * iterations
        for i, t_i in enumerate(t[:-1]):
            def round_draws(i=i, t_i=t_i):
                ys, k = [], 0
                for job, count in enumerate(counts):
                    ys += draw(job, [lambda y=cur[k + j]: perturb(t_i, y)
                                     for j in range(count)])
                    k += count
                return ys

            loaded, final = self._solve(oracle, problem, round_draws(), n_sub, stream,
                                        repeat_on_error, round_draws)
            rec[:, i].copy_(loaded)
            nxt[:, i].copy_(final)
            if i < n_steps - 1:
                cur = list(_channel_last(problem, final.cpu().numpy(), cp))
        return rec.view(iterations * n_steps, -1), nxt.view(iterations * n_steps, -1)

    def _whole(self, ivp, oracle, problem, t, y_0, counts, draw, perturb, repeat_on_error,
               stream):
        """One perturbed solve of the whole IVP per iteration, targets at the
        operator's time points (supervised_ml_operator.py:487-497)."""
        n_steps = len(t) - 1
        total = len(discretize_time_domain(ivp.t_interval, oracle.d_t)) - 1
        idx = np.rint((t[1:] - t[0]) / oracle.d_t).astype(int) - 1
        every, pick = _snapshot_plan(idx, total)

        def all_draws():
            ys = []
            for job, count in enumerate(counts):
                ys += draw(job, [lambda: perturb(t[0], y_0) for _ in range(count)])
            return ys

        rec, _, snaps = self._solve(oracle, problem, all_draws(), total, stream,
                                    repeat_on_error, all_draws, every=every,
                                    n_snap=total // every)
        iterations = sum(counts)
        # snapshots (n_snap, grids x elems) -> (grids, step, elems)
        nxt = snaps[pick].view(n_steps, iterations, problem.elems).transpose(0, 1)
        return rec, nxt.contiguous().view(iterations * n_steps, -1)

    # --------------------------------------------------------------- the rows
    def _rows(self, cp, t, rec, per_state, nxt, iterations, n_steps, problem, stream):
        import torch

        x = cp.mesh.all_index_coordinates(True, flatten=True)
        points, x_dim = x.shape
        c = problem.n_fields
        has_t = self._time_variant or self._input_d_t
        width = c * points + int(has_t) + x_dim
        n_rows = iterations * n_steps * points
        inputs = torch.empty((n_rows, width), dtype=torch.float64, device="cuda")
        targets = torch.empty((n_rows, c), dtype=torch.float64, device="cuda")
        coords = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()
        tcol = None
        if self._time_variant:
            tcol = torch.from_numpy(np.ascontiguousarray(t[1:], dtype=np.float64)).cuda()
        elif self._input_d_t:
            tcol = torch.full((n_steps,), float(self._d_t), dtype=torch.float64, device="cuda")
        rec = rec.contiguous()
        nxt = nxt.contiguous()
        nat.check(nat.device_lib().pml_supervised_rows(
            rec.data_ptr(), per_state, nxt.data_ptr(), iterations * n_steps, n_steps, points,
            c, None if tcol is None else tcol.data_ptr(), coords.data_ptr(), x_dim,
            inputs.data_ptr(), targets.data_ptr(), ctypes.c_void_p(stream.cuda_stream)))
        for tensor in (rec, nxt, coords) + (() if tcol is None else (tcol,)):
            tensor.record_stream(stream)
        return inputs, targets


def _snapshot_plan(idx: np.ndarray, total: int):
    """(snapshot interval, snapshot indices) that sample the oracle steps
    ``idx + 1`` of a ``total``-step solve: every k-th step when they are the
    multiples of k, else every step."""
    k = int(idx[0]) + 1
    if k > 0 and total % k == 0 and np.array_equal(idx, k * np.arange(1, len(idx) + 1) - 1):
        return k, np.arange(len(idx))
    return 1, idx


def _channel_last(problem, host: np.ndarray, cp) -> np.ndarray:
    """(grids, elems) field-major host states -> (grids, *vertices, F)."""
    grids = host.reshape(len(host), problem.n_fields, problem.n0, problem.n1)
    return problem.from_device_layout(grids, cp.mesh.vertices_shape)
