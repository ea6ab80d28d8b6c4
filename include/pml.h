/*
 * pml.h -- C ABI of the MI355X (gfx950) finite-difference + Parareal
 * library (libpml.so).
 *
 * Every entry point returns a status code; on failure a thread-local
 * message is available from pml_last_error().  All device buffers are owned
 * by the caller (plain device pointers, e.g. from hipMalloc or a torch
 * tensor's data_ptr()); the library keeps no pointer to them after a call
 * returns.  Calls are ordered on the stream passed in (NULL = the null
 * stream); entry points that produce host-visible results synchronise that
 * stream before returning, the others do not.
 *
 * Reference interfaces replaced (paths relative to the PararealML source
 * tree):
 *   pml_fdm_solve          <- FDMOperator.solve time loop
 *                             (pararealml/operators/fdm/fdm_operator.py:48-165)
 *                             with ForwardEulerMethod / ExplicitMidpointMethod
 *                             / RK4.integral
 *                             (operators/fdm/numerical_integrator.py:42-132)
 *   pml_stencil_first      <- ThreePointCentralDifferenceMethod._derivative
 *                             (operators/fdm/numerical_differentiator.py:1012-1057)
 *   pml_stencil_second     <- ThreePointCentralDifferenceMethod._second_derivative
 *                             (same axis) + _add_halos_along_axis
 *                             (numerical_differentiator.py:1059-1095, 1188-1242)
 *   pml_weighted_add       <- the coordinate factors of the polar, cylindrical
 *                             and spherical branches of gradient, hessian,
 *                             divergence, curl, laplacian, vector_laplacian
 *                             (numerical_differentiator.py:158-173, 229-308,
 *                             348-400, 482-590, 636-725, 771-870)
 *   pml_set_values         <- apply_constraints_along_last_axis
 *                             (constraint.py:104-131) on a full channel-last
 *                             state (curvilinear stepping)
 *   pml_secant_step        <- scipy.optimize.newton (array secant branch) as
 *                             called by BackwardEulerMethod / CrankNicolsonMethod
 *                             (operators/fdm/numerical_integrator.py:135-270)
 *   pml_jacobi_sweep       <- ThreePointCentralDifferenceMethod.
 *                             _next_anti_laplacian_estimate (Cartesian) and the
 *                             constraint application of anti_laplacian
 *                             (numerical_differentiator.py:872-927, 1097-1186)
 *   pml_jacobi_solve_weighted <- the polar / cylindrical / spherical branches
 *                             of _next_anti_laplacian_estimate
 *                             (numerical_differentiator.py:1097-1186)
 *   pml_jacobi_solve       <- the anti_laplacian loop itself
 *                             (numerical_differentiator.py:904-925)
 *   pml_parareal_update    <- the per-slice update U[j+1] = G[j] + corr[j] and
 *                             the RMS update norms of PararealOperator
 *                             (operators/parareal/parareal_operator.py:53-100,
 *                             164-188)
 *   pml_parareal_correction<- correction = F_end - G_end
 *                             (parareal_operator.py:164)
 *   pml_program_create / pml_expr_eval
 *                          <- SymbolMapper / FDMSymbolMapper: the lambdified
 *                             right-hand sides (operators/symbol_mapper.py:
 *                             160-271, fdm/fdm_symbol_mapper.py:45-158)
 *   pml_ode_solve          <- FDMOperator.solve on a system without a mesh
 *                             (fdm_operator.py:48-77, 190) with the explicit
 *                             integrators (numerical_integrator.py:42-132)
 *   pml_graph_cache_clear  <- no counterpart (the reference has no device
 *                             caches): releases the step-loop graphs and
 *                             work-unit tables pml_fdm_solve keeps
 *   pml_interp_regular / pml_snapshots_time_inner
 *                          <- the scipy.interpolate.interpn call of
 *                             Solution.y and of discrete_y's orientation
 *                             change (solution.py:114-180)
 *   pml_snapshot_diff      <- the matched-time subtraction of Solution.diff
 *                             (solution.py:182-258)
 *   pml_supervised_rows    <- the row layout of
 *                             SupervisedMLOperator._generate_data
 *                             (operators/ml/supervised/
 *                             supervised_ml_operator.py:362-379, 440-499)
 */
#ifndef PML_H
#define PML_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define PML_OK 0
#define PML_E_INVALID 1     /* bad argument   -> Python ValueError        */
#define PML_E_UNSUPPORTED 2 /* not on device  -> Python NotImplementedError */
#define PML_E_HIP 3         /* runtime error  -> Python RuntimeError       */

/* Blocks of the fixed-order reductions (update and Jacobi norms): their
 * workspaces hold this many partial sums per reduced quantity. */
#define PML_REDUCE_SLOTS 2048
/* Bytes of the state buffer of pml_jacobi_solve / pml_jacobi_solve_weighted. */
#define PML_JACOBI_STATE_BYTES (32 + 8 * PML_REDUCE_SLOTS)

/* ---- state precision ---------------------------------------------------- */
#define PML_F32 0  /* FP32 state, FP32 arithmetic                           */
#define PML_F64 1  /* FP64 state, FP64 arithmetic (parity with the oracle)  */
#define PML_F32C 2 /* compensated FP32 state: hi and lo FP32 planes; stencil
                      arithmetic in FP32, state update by error-free sums   */

/* ---- equations (fused right-hand sides) -------------------------------- */
#define PML_EQ_DIFFUSION 0            /* params: d                        */
#define PML_EQ_CAHN_HILLIARD 1        /* params: d, gamma                 */
#define PML_EQ_SHALLOW_WATER 2        /* params: h, b, v, f, g            */
#define PML_EQ_CONVECTION_DIFFUSION 3 /* params: d, vel0, vel1            */
#define PML_EQ_WAVE 4                 /* params: c^2                      */
#define PML_EQ_BURGERS 5              /* params: 1/re                     */
#define PML_EQ_NAVIER_STOKES 6        /* params: 1/re; fields w, psi, u, v;
                                         psi (LHS.Y_LAPLACIAN) is left
                                         unchanged by the step and re-solved
                                         by the caller (pml_jacobi_solve) */

/* ---- integrators ------------------------------------------------------- */
#define PML_INT_FE 0
#define PML_INT_MIDPOINT 1
#define PML_INT_RK4 2

/*
 * One grid of a (possibly ragged) batch.  State is stored field-major:
 * field q of the grid occupies elements
 *   [state_offset + q*n0*n1, state_offset + (q+1)*n0*n1)
 * with axis 1 contiguous.  A 1-D problem is passed as n0 = 1 with the line
 * along axis 1 and ndim = 1 in pml_fdm_solve.
 *
 * Boundary values (always FP64) for the grid start at bc_offset and hold,
 * per field q, 8 consecutive edge arrays:
 *   Dirichlet: axis0-lower[n1], axis0-upper[n1], axis1-lower[n0], axis1-upper[n0]
 *   Neumann:   axis0-lower[n1], axis0-upper[n1], axis1-lower[n0], axis1-upper[n0]
 * i.e. 4*(n0+n1) values per field.  NaN marks an unconstrained point.  The
 * Dirichlet edges are the perimeter of the problem's full-grid y
 * constraints, so corner precedence is resolved before the call.
 */
typedef struct {
  int32_t n0;
  int32_t n1;
  double dx0;
  double dx1;
  int64_t state_offset;
  int64_t bc_offset;
} pml_grid;

/* ---- library ----------------------------------------------------------- */
int pml_init(int device);
const char *pml_last_error(void);
int pml_version(void);
int pml_device_sync(void *stream);

/*
 * Advances every grid of the batch by n_steps steps of size dt.
 *   state_hi/state_lo : in/out state (state_lo only for PML_F32C, else NULL)
 *   work_hi/work_lo   : ping-pong buffers of the same size
 *   snapshots_hi/lo   : optional (NULL = none) output of the whole batch state
 *                       after every snapshot_every-th step; snapshot k (k>=0)
 *                       is written at element offset k*total_elems
 *   total_elems       : elements of one state plane (all grids, all fields)
 *   tile_config       : kernel variant (0 = default; see pml_fdm_configs)
 *   flags             : PML_FLAG_GRAPH -- replay the step loop from a
 *                       captured hipGraph (needs a non-null stream);
 *                       PML_FLAG_CONSTRAIN_INPUT -- impose the Dirichlet
 *                       values on the initial state before the first step
 *                       (what the reference's initial conditions do);
 *                       PML_FLAG_PREPARE -- build (and cache) the unit
 *                       tables and the step-loop graph this call would use,
 *                       and launch nothing
 * After the call the current state is in state_hi/lo.
 *
 * Work-unit tables are built once per distinct grid array and kept (device
 * memory of the library); with PML_FLAG_GRAPH the step loop is a captured
 * hipGraph -- the whole loop for up to 64 launches, else a 32-launch graph
 * replayed -- instantiated once per distinct set of arguments and kept, so
 * a repeated call costs no host-side preparation.  pml_graph_cache_clear()
 * synchronises every device and releases the graphs and the tables.
 */
#define PML_FLAG_GRAPH 1
#define PML_FLAG_CONSTRAIN_INPUT 2
#define PML_FLAG_PREPARE 4

int pml_fdm_solve(int eq_id, const double *eq_params, int n_params,
                  int integrator_id, int precision, int ndim,
                  const pml_grid *grids, int n_grids, const double *bc_values,
                  void *state_hi, void *state_lo, void *work_hi, void *work_lo,
                  int64_t total_elems, double dt, int64_t n_steps,
                  void *snapshots_hi, void *snapshots_lo,
                  int64_t snapshot_every, int tile_config, int flags,
                  void *stream);

int pml_graph_cache_clear(void);

/*
 * pml_fdm_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under hipGraph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as pml_fdm_solve; the work-unit and grid tables
 * go to a file-scope device array of PML_FDM_SCRATCH_BYTES (one per GPU,
 * held until the process ends) instead of cached library allocations, no
 * step-loop graph is used (PML_FLAG_GRAPH is ignored; PML_FLAG_PREPARE is
 * an error), and a batch whose tables exceed the array is refused with
 * PML_E_INVALID.  The array is rewritten by every call: calls that overlap
 * on one GPU (different streams or threads) race.
 */
#define PML_FDM_SCRATCH_BYTES (4 << 20)
int pml_fdm_solve_noalloc(int eq_id, const double *eq_params, int n_params,
                          int integrator_id, int precision, int ndim,
                          const pml_grid *grids, int n_grids,
                          const double *bc_values, void *state_hi,
                          void *state_lo, void *work_hi, void *work_lo,
                          int64_t total_elems, double dt, int64_t n_steps,
                          void *snapshots_hi, void *snapshots_lo,
                          int64_t snapshot_every, int tile_config, int flags,
                          void *stream);

/* Number of kernel variants for (eq, integrator, precision) and a short
 * description of variant i (for the autotuner). */
int pml_fdm_configs(int eq_id, int integrator_id, int precision);
const char *pml_fdm_config_name(int eq_id, int integrator_id, int precision,
                                int config);

/*
 * Standalone stencils on a (outer, n, inner) view of a contiguous array,
 * differentiating along the middle axis with spacing dx.  lo/hi are FP64
 * (outer, inner) arrays of boundary derivative values (NaN = none) or NULL.
 *   pml_stencil_first : D = (y[i+1]-y[i-1])/(2dx), zero halos, D[0]/D[n-1]
 *                       replaced by lo/hi where given.
 *   pml_stencil_second: S = ((y[i+1]-2y[i])+y[i-1])/(dx*dx2), ghosts
 *                       y[1]-2dx*lo / y[n-2]+2dx*hi where given, else 0.
 * mode: 0 -> out = r, 1 -> out += r, -1 -> out -= r.
 */
int pml_stencil_first(int precision, const void *y, void *out, int64_t outer,
                      int64_t n, int64_t inner, double dx, const double *lo,
                      const double *hi, int mode, void *stream);
int pml_stencil_second(int precision, const void *y, void *out, int64_t outer,
                       int64_t n, int64_t inner, double dx, double dx2,
                       const double *lo, const double *hi, int mode,
                       void *stream);

/*
 * out (=, +=, -=) coef * x over `count` elements (mode as above); x, coef
 * and out share the precision.  The curvilinear operators of
 * NumericalDifferentiator combine stencil results and field components with
 * coordinate factors (1/r, cos(phi)/sin(phi), ...) through this call.
 */
int pml_weighted_add(int precision, const void *x, const void *coef, void *out,
                     int64_t count, int mode, void *stream);

/*
 * y[i] = values[i] wherever values[i] (FP64) is not NaN, over `count`
 * elements: the static y constraints of a channel-last state in the
 * curvilinear stepping path of FDMOperator.
 */
int pml_set_values(int precision, void *y, const double *values, int64_t count,
                   void *stream);

/*
 * One iteration of the element-wise secant method of
 * scipy.optimize.newton on arrays (no derivative), used by the implicit
 * integrators.  mode 0: where q1 != q0, p_new = p1 - q1 (p1 - p) / (q1 - q0)
 * and failing[i] = |dp| >= tol; elsewhere p_new = (p1 + p) / 2 for elements
 * still active (active[i] != 0), then they become inactive, or p_new = p.
 * counters (2 x uint64, zeroed by the caller) receive the number of
 * elements with q1 != q0 and of those still failing.  mode 1: p_new =
 * (p1 + p) / 2 everywhere (q0, q1, active, failing, counters unused).
 * mode 2: the second start point p_new = p (1 + dx) + (p >= 0 ? dx : -dx)
 * with dx passed in tol (only p and p_new used).
 */
int pml_secant_step(int precision, int mode, const void *p, const void *p1,
                    const void *q0, const void *q1, void *p_new, void *active,
                    void *failing, int64_t count, double tol, void *counters,
                    void *stream);

/*
 * One Jacobi sweep of the Cartesian anti-Laplacian on up to 3 spatial axes
 * (shape[k] == 1 marks an absent axis; the last dimension holds n_fields).
 * ghosts: per axis a, FP64 lower/upper slab arrays of the derivative
 * constraints (NULL = none), laid out as the array with axis a removed.
 * ycons: full-grid FP64 y constraints (NaN = free) or NULL.
 * sumsq_out: device FP64 scalar receiving sum((y_new-y_old)^2), summed in
 * a fixed order (bitwise reproducible); work: device workspace of
 * PML_REDUCE_SLOTS doubles.
 */
int pml_jacobi_sweep(int precision, const void *y_old, void *y_new,
                     const void *lap, const int64_t *shape, int n_fields,
                     const double *dx, const double *const *ghost_lo,
                     const double *const *ghost_hi, const double *ycons,
                     double *sumsq_out, double *work, void *stream);
/*
 * The whole Jacobi loop of anti_laplacian (numerical_differentiator.py:
 * 872-927): sweeps (as pml_jacobi_sweep, y -> work -> y ...) until
 * ||y_new - y_old||_2 <= tol, with the stopping test on the device (a sweep
 * after convergence does nothing) and the host reading the loop state once
 * per `batch` sweeps.  y holds the start (constraints already applied) and
 * receives the result; work is a second array of the same size; state is a
 * device buffer of PML_JACOBI_STATE_BYTES (8-byte aligned).  The update
 * norm is summed in a fixed order, so the sweep count and the result are
 * bitwise reproducible.  *sweeps_out = the
 * sweeps taken, or -(sweeps issued) if max_sweeps ran out first; *diff_out
 * (may be NULL) = the last update norm.
 */
int pml_jacobi_solve(int precision, void *y, void *work, const void *lap,
                     const int64_t *shape, int n_fields, const double *dx,
                     const double *const *ghost_lo, const double *const *ghost_hi,
                     const double *ycons, double tol, int64_t max_sweeps,
                     int64_t batch, void *state, int64_t *sweeps_out,
                     double *diff_out, void *stream);

/*
 * pml_jacobi_solve with per-vertex weights, for polar, cylindrical and
 * spherical meshes (numerical_differentiator.py:1097-1186): each sweep sets
 * y_new = (sum_a w_sum[a] (prev_a + next_a) + w_diff[a] (next_a - prev_a)
 * - lap) / w_den, where prev_a / next_a are the axis-a neighbours (ghosts
 * as in pml_jacobi_sweep).  w_sum[a] (required for every axis with more
 * than one vertex), w_diff[a] (NULL entries or a NULL array: no term) and
 * w_den are FP64 arrays over the vertices (shape[0] x shape[1] x shape[2],
 * C order), shared by all fields.
 */
int pml_jacobi_solve_weighted(int precision, void *y, void *work, const void *lap,
                              const int64_t *shape, int n_fields, const double *dx,
                              const double *const *ghost_lo,
                              const double *const *ghost_hi, const double *ycons,
                              double tol, int64_t max_sweeps, int64_t batch,
                              void *state, int64_t *sweeps_out, double *diff_out,
                              const double *const *w_sum, const double *const *w_diff,
                              const double *w_den, void *stream);

/*
 * Parareal, all FP64.
 *   pml_parareal_correction: corr[s] = f_end[s] - g_end[s] over n elements
 *     per slice, for slices [s0, s1).  f_end may be FP32 or compensated FP32
 *     (f_precision), converted exactly to FP64 before the subtraction.
 *   pml_parareal_update: u_next = g_new + corr (one slice, n elements) and
 *     per field (n_fields planes of n/n_fields) the sum of
 *     (u_next - u_old)^2 into sumsq[n_fields] (accumulated: the caller
 *     zeroes it).  The sums are formed in a fixed order (per-block partials
 *     in work, then added in block order), so norms and iteration counts
 *     are bitwise reproducible; work is a device workspace of
 *     PML_REDUCE_SLOTS * n_fields doubles.
 */
int pml_parareal_correction(int f_precision, const void *f_end_hi,
                            const void *f_end_lo, int64_t f_stride,
                            const double *g_end, double *corr, int64_t n,
                            int64_t n_slices, void *stream);
int pml_parareal_update(const double *g_new, const double *corr,
                        double *u_next, int64_t n, int n_fields,
                        double *sumsq, double *work, void *stream);

/* Imposes the Dirichlet edge values of one grid's boundary block (pml_grid
 * layout, FP64) on `batch` consecutive FP64 field-major copies of the grid
 * (what the reference's constraint application does to a state,
 * pararealml/constraint.py:104-131). */
int pml_apply_dirichlet(double *state, const double *bc_values, int n_fields,
                        int n0, int n1, int ndim, int64_t batch, void *stream);

/* Conversions between FP64 and the state precisions (device to device). */
int pml_from_f64(int precision, const double *src, void *dst_hi, void *dst_lo,
                 int64_t n, void *stream);
int pml_to_f64(int precision, const void *src_hi, const void *src_lo,
               double *dst, int64_t n, void *stream);

/*
 * Fallback test of a learned coarse prediction (no counterpart in the
 * reference): *flag is set to 1 when dst (n FP64 values) holds a non-finite
 * value or, src holding no NaN, max|dst| > guard * max(max|src|, 1e-300),
 * and left as it is otherwise.  work: 4 * PML_REDUCE_SLOTS doubles.
 */
int pml_slice_guard(const double *src, const double *dst, int64_t n,
                    double guard, double *work, int32_t *flag, void *stream);

/*
 * 2x grid transfer between a fine vertex grid (2m-1 points per axis) and a
 * coarse one (m points), batched over `batch` planes, FP64 or FP32 by
 * precision (PML_F32C not accepted).  No counterpart in the reference.
 */
int pml_restrict2x(int precision, const void *fine, void *coarse,
                   int64_t batch, int64_t m0, int64_t m1, void *stream);
int pml_prolong2x(int precision, const void *coarse, void *fine,
                  int64_t batch, int64_t m0, int64_t m1, void *stream);
/* Cubic 2x prolongation (at least 4 coarse points per axis): even fine
 * points copy the coarse ones, odd ones take the 4-point Lagrange midpoint
 * value ((-c[k-1] + 9 c[k]) + 9 c[k+1] - c[k+2]) / 16, one-sided
 * (((5 c[0] + 15 c[1]) - 5 c[2]) + c[3]) / 16 at the two ends; axis 0
 * first, then axis 1.  Exact on cubics in each axis.  No counterpart in
 * the reference. */
int pml_prolong2x_cubic(int precision, const void *coarse, void *fine,
                        int64_t batch, int64_t m0, int64_t m1, void *stream);

/*
 * Symbolic right-hand sides on the device.  The host compiles a system's
 * right-hand sides (SymPy expressions over the reference's symbols: y, t,
 * x, y-gradient, y-hessian, y-laplacian, y-divergence, y-curl,
 * y-vector-laplacian) into a register-machine program: n_instr
 * instructions of four int32 {op, dst, a, b} (ops in pml_expr.hip:
 * constants, input loads, t, + - * /, pow, integer powers, negation,
 * elementary functions, min/max/sign, output stores), n_regs registers
 * (at most PML_EXPR_MAX_REGS), n_inputs per-element inputs and n_outputs
 * outputs.  pml_program_create validates every instruction's operands
 * against those counts and uploads the program to the current device.
 *   pml_expr_eval: evaluates the program at `count` elements; input k of
 *     element e is inputs[k][e * in_strides[k]] and output k goes to
 *     outputs[k][e * out_strides[k]] (element strides; FP32 or FP64 by
 *     precision, consts always FP64); t is the time symbol.  Replaces the
 *     lambdified evaluation of SymbolMapper.map (symbol_mapper.py:222-271)
 *     once the derivative symbols have been computed into arrays.
 *   pml_ode_solve: advances `batch` independent states of an ODE system
 *     (n = n_outputs unknowns; input k is y[k]) by n_steps explicit steps
 *     (PML_INT_FE / MIDPOINT / RK4) of size dt, one thread per state, the
 *     whole time loop in one launch; y is (batch, n) in/out; the time of
 *     step s of state b is s * t_step[b] + t0[b] (the reference's
 *     linspace grid; t0, t_step device FP64 arrays); snapshots (may be
 *     NULL) receive every snapshot_every-th state at
 *     snapshots[(k * batch + b) * n].  Replaces FDMOperator.solve for
 *     systems without a mesh (fdm_operator.py:48-77 with
 *     numerical_integrator.py:42-132).
 */
#define PML_EXPR_MAX_REGS 96
#define PML_EXPR_MAX_IO 48
typedef struct pml_program_s *pml_program;
int pml_program_create(const int32_t *code, int n_instr, const double *consts,
                       int n_consts, int n_regs, int n_inputs, int n_outputs,
                       pml_program *out);
int pml_program_destroy(pml_program program);
int pml_expr_eval(int precision, pml_program program, const void *const *inputs,
                  const int64_t *in_strides, void *const *outputs,
                  const int64_t *out_strides, double t, int64_t count,
                  void *stream);
int pml_ode_solve(int precision, pml_program program, int integrator_id,
                  int64_t batch, const double *t0, const double *t_step,
                  double dt, int64_t n_steps, void *y, void *snapshots,
                  int64_t snapshot_every, void *stream);

/*
 * Solution post-processing on the device.  A snapshot stack holds n_t
 * snapshots of a C-order grid of shape[0..n_dims-1] points with n_fields
 * values per point, in one of three layouts:
 *   PML_LAYOUT_CHANNEL_LAST  [t][point][field] (the reference's layout),
 *   PML_LAYOUT_FIELD_MAJOR   [t][field][point] (the fused solvers' layout),
 *   PML_LAYOUT_TIME_INNER    [point][field][t] (every point's history
 *                            contiguous: the layout scattered queries read
 *                            fastest).
 *   pml_snapshots_time_inner: copies a channel-last or field-major stack
 *     (layout) of `points` grid points into the time-inner layout (out).
 *   pml_interp_regular: values is such a stack (1 <= n_dims <= 3) in
 *     `layout`; coords holds the ascending axis coordinates one axis after
 *     the other (shape[0] + ... values, FP64);
 *     xi the n_points query points (n_points x n_dims, FP64); out receives
 *     out[(t * n_points + p) * n_fields + f] in the precision of values.
 *     Method PML_INTERP_LINEAR or PML_INTERP_NEAREST with SciPy's
 *     RegularGridInterpolator semantics (scipy 1.15) under bounds_error
 *     False and fill_value None: outside the grid the boundary interval
 *     extrapolates, NaN coordinates give NaN; FP64 results equal SciPy's.
 *     Replaces the interpn call of Solution.y (solution.py:114-145), which
 *     Solution.discrete_y's orientation change also uses (solution.py:147-180).
 *   pml_snapshot_diff: out[k] = b[ib[k]] - a[ia[k]] for k < n_pairs, two
 *     stacks of one layout (field_major: PML_LAYOUT_FIELD_MAJOR, else
 *     channel-last) and `points` x n_fields values per snapshot
 *     (FP32 or FP64), out channel-last; ia, ib are device int64 index
 *     arrays, an index outside [0, n_a) or [0, n_b) gives NaNs for its pair.
 *     Replaces the matched-time subtraction of Solution.diff
 *     (solution.py:241-254).
 */
#define PML_INTERP_LINEAR 0
#define PML_INTERP_NEAREST 1
#define PML_LAYOUT_CHANNEL_LAST 0
#define PML_LAYOUT_FIELD_MAJOR 1
#define PML_LAYOUT_TIME_INNER 2
int pml_snapshots_time_inner(int precision, int layout, const void *values,
                             int64_t n_t, int64_t points, int n_fields,
                             void *out, void *stream);
int pml_interp_regular(int precision, int method, int layout,
                       const void *values, int64_t n_t, int n_dims,
                       const int64_t *shape, int n_fields,
                       const double *coords, const double *xi,
                       int64_t n_points, void *out, void *stream);
int pml_snapshot_diff(int precision, int field_major, const void *a,
                      const int64_t *ia, int64_t n_a, const void *b,
                      const int64_t *ib, int64_t n_b, void *out,
                      int64_t n_pairs, int64_t points, int n_fields,
                      void *stream);

/*
 * Rows of the supervised-learning data set (FP64, device arrays), as
 * SupervisedMLOperator._generate_data lays them out
 * (operators/ml/supervised/supervised_ml_operator.py:362-379, 440-499).
 * n_blocks = iterations x n_steps blocks (iteration-major); block b's
 * recorded state is states[b / steps_per_state] and its next state
 * nexts[b], both field-major (n_fields x points).  Row r = b * points + p
 * of `inputs` (width n_fields * points + (tcol != NULL) + x_dim) is the
 * recorded state flattened channel-last, then tcol[b % n_steps] (time or
 * step column, omitted when tcol is NULL), then coords[p] (x_dim values);
 * row r of `targets` (n_fields wide) is the next state at p.
 */
int pml_supervised_rows(const double *states, int64_t steps_per_state,
                        const double *nexts, int64_t n_blocks, int64_t n_steps,
                        int64_t points, int n_fields, const double *tcol,
                        const double *coords, int x_dim, double *inputs,
                        double *targets, void *stream);

/*
 * Convolutional surrogate (no counterpart in the reference; it replaces the
 * coarse propagator).  Tensors are NHWC BF16 unless stated; accumulation is
 * FP32; 3x3 "same" convolutions with zero padding.
 *   pml_conv64          : 64 -> 64 layer on the matrix cores.
 *                         mode 0 (forward): wk = [9][64 co][64 ci] weights,
 *                           bias FP32[64]; writes z_out (pre-activation,
 *                           may be NULL) and a_out = GELU(z).
 *                         mode 1 (input gradient): in = dL/dz of this
 *                           layer, wk = [9][64 ci][64 co] with the taps
 *                           flipped, zprev = the input's pre-activation;
 *                           writes a_out = dL/dz of the previous layer.
 *   pml_wgrad64         : dw[9][64 ci][64 co] += weight gradient and, when
 *                         db is not NULL, db[64] += bias gradient (FP32).
 *   pml_conv_small      : C -> 64 (C in 1..3) or 64 -> C layers, w =
 *                         [9][ci][co]; forward writes z_out/a_out (64
 *                         outputs) or y_out (FP32, C outputs); mode 1 is the
 *                         input gradient of a 64 -> C layer.  A C -> 64
 *                         layer reads its input and its weights in FP32,
 *                         a 64 -> C layer in BF16.
 *   pml_conv_small_wgrad: dw[9][ci][co] and db[co] += gradients (FP32); x is
 *                         FP32 for a C -> 64 layer, BF16 for 64 -> C.
 *                         Both weight-gradient entries take `part`: NULL
 *                         adds each workgroup's sums atomically (the order,
 *                         and so the last bits, vary from run to run);
 *                         a device buffer of at least
 *                         pml_grad_partials_floats(Ci, Co, B, H, W) floats
 *                         receives each workgroup's sums, which are then
 *                         added in workgroup order (bitwise reproducible).
 *   pml_grad_partials_floats: size of that buffer for a Ci -> Co layer
 *                         (64 -> 64, C -> 64 or 64 -> C) on B x H x W.
 *   pml_bias_grad64     : db[64] += sum over pixels of dz.
 *   pml_adam            : Adam step on FP32 masters p/m/v with gradient g
 *                         (step = 1, 2, ...); writes BF16 weights in the
 *                         plain [9][ci][co], forward [9][co][ci] and
 *                         flipped backward layouts (each may be NULL).
 *   pml_surrogate_normalize: per-sample scale s = max|u| and x = u / s,
 *                         u field-major [B][C][HW] FP32/FP64 -> x NHWC FP32.
 *   pml_surrogate_residual : out = u + s y, u and out field-major FP32/FP64
 *                         (the identity part never passes through BF16;
 *                         out may alias u).
 *   pml_surrogate_loss     : loss += sum (y - (t - u)/s)^2 (FP64 scalar) and
 *                         dy = 2 (y - (t - u)/s) / n (BF16 NHWC); u and t
 *                         field-major FP32.
 */
int pml_conv64(int mode, const void *in, const void *wk, const float *bias,
               const void *zprev, void *z_out, void *a_out, int B, int H, int W,
               void *stream);
int pml_wgrad64(const void *x, const void *dz, float *dw, float *db, int B, int H,
                int W, float *part, int64_t part_floats, void *stream);
int pml_conv_small(int mode, const void *in, const void *w, const float *bias,
                   const void *zprev, void *z_out, void *a_out, float *y_out,
                   int Ci, int Co, int B, int H, int W, void *stream);
int pml_conv_small_wgrad(const void *x, const void *dz, float *dw, float *db,
                         int Ci, int Co, int B, int H, int W, float *part,
                         int64_t part_floats, void *stream);
int pml_grad_partials_floats(int Ci, int Co, int B, int H, int W, int64_t *floats);
int pml_bias_grad64(const void *dz, float *db, int64_t pixels, void *stream);
int pml_adam(float *p, float *m, float *v, const float *g, int64_t n, float lr,
             float beta1, float beta2, float eps, int step, void *plain,
             void *wf, void *wb, int ci, int co, void *stream);
int pml_surrogate_normalize(int precision, const void *u, void *x, float *s,
                            int B, int C, int64_t HW, void *stream);
int pml_surrogate_residual(int precision, const void *u, const float *y,
                           const float *s, void *out, int B, int C, int64_t HW,
                           void *stream);
int pml_surrogate_loss(const float *u, const float *y, const float *s,
                       const float *target, void *dy, double *loss, int B,
                       int C, int64_t HW, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PML_H */
