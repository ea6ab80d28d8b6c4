// Standalone three-point stencils and the Jacobi anti-Laplacian sweep for
// gfx950, backing NumericalDifferentiator's public methods on any number of
// dimensions.  Semantics restate
// pararealml/operators/fdm/numerical_differentiator.py:1012-1242.
#include <hip/hip_runtime.h>

#include <cstddef>

#include "pml_common.h"

namespace pml {

static int grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > (1 << 20)) b = 1 << 20;
  return static_cast<int>(b < 1 ? 1 : b);
}

template <typename T>
__device__ __forceinline__ void emit(T *out, int64_t t, T r, int mode) {
  if (mode == 0) out[t] = r;
  else if (mode > 0) out[t] = out[t] + r;
  else out[t] = out[t] - r;
}

// (y[i+1] - y[i-1]) / (2 dx) with zero halos, boundary override by lo/hi
// (numerical_differentiator.py:1012-1057).
template <typename T>
__global__ void k_first(const T *y, T *out, int64_t outer, int64_t n, int64_t inner,
                        double two_dx, const double *lo, const double *hi, int mode) {
  const int64_t total = outer * n * inner;
  const T den = static_cast<T>(two_dx);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = t / (n * inner), rem = t - o * n * inner;
    const int64_t i = rem / inner, q = rem - i * inner;
    const T yp = i > 0 ? y[t - inner] : T(0);
    const T yn = i < n - 1 ? y[t + inner] : T(0);
    T d = (yn - yp) / den;
    if (i == 0 && lo) {
      const double g = lo[o * inner + q];
      if (g == g) d = static_cast<T>(g);
    }
    if (i == n - 1 && hi) {
      const double g = hi[o * inner + q];
      if (g == g) d = static_cast<T>(g);
    }
    emit(out, t, d, mode);
  }
}

// ((y[i+1] - 2 y[i]) + y[i-1]) / (dx dx2) with constraint ghosts
// (numerical_differentiator.py:1083-1095, 1188-1242).
template <typename T>
__global__ void k_second(const T *y, T *out, int64_t outer, int64_t n, int64_t inner,
                         double dx, double den_d, const double *lo, const double *hi,
                         int mode) {
  const int64_t total = outer * n * inner;
  const T den = static_cast<T>(den_d);
  const T m2dx = static_cast<T>(-2.0 * dx), p2dx = static_cast<T>(2.0 * dx);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = t / (n * inner), rem = t - o * n * inner;
    const int64_t i = rem / inner, q = rem - i * inner;
    T yp, yn;
    if (i > 0) {
      yp = y[t - inner];
    } else {
      const double g = lo ? lo[o * inner + q] : NAN;
      yp = (g == g) ? y[t + inner] + m2dx * static_cast<T>(g) : T(0);
    }
    if (i < n - 1) {
      yn = y[t + inner];
    } else {
      const double g = hi ? hi[o * inner + q] : NAN;
      yn = (g == g) ? y[t - inner] + p2dx * static_cast<T>(g) : T(0);
    }
    const T r = ((yn - T(2) * y[t]) + yp) / den;
    emit(out, t, r, mode);
  }
}

struct JacobiGeom {
  int64_t n[3];
  int64_t stride[3];  // element strides of the three spatial axes
  double dx[3];
  double dxsq[3];
  const double *glo[3];
  const double *ghi[3];
  int active[3];
  double coef;  // sum over active axes of 2 / dx^2
  // Curvilinear weights per vertex (NULL: Cartesian): the update is
  // (sum_a wsum[a] (prev + next) + wdiff[a] (next - prev) - lap) / wden.
  const double *wsum[3];
  const double *wdiff[3];
  const double *wden;
};

// One Jacobi update of the Cartesian anti-Laplacian
// (numerical_differentiator.py:1097-1186) followed by the y constraints
// and a sum of squared updates (872-927).
// State of a device-side Jacobi loop (pml_jacobi_solve): the sweep's
// sum of squared updates, a done flag and the number of sweeps taken.
struct JacobiState {
  double sumsq;
  double last_diff;
  long long sweeps;
  int done;
  int pad;
  double partials[PML_REDUCE_SLOTS];  // one per sweep block, summed in order
};
static_assert(sizeof(JacobiState) == PML_JACOBI_STATE_BYTES, "JacobiState layout");

// Each block writes its sum of squared updates to partials[blockIdx.x]
// (grid-stride order, then a fixed tree); k_jacobi_check adds them in block
// order, so the norm and the sweep count do not depend on scheduling.
template <typename T>
__global__ void k_jacobi(const T *y, T *yn, const T *lap, int64_t total, int nf,
                         JacobiGeom g, const double *ycons, double *partials,
                         const int *done) {
  // A loop that has converged turns its remaining sweeps into no-ops.
  if (done && *done) return;
  __shared__ double red[256];
  double acc2 = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = t % nf;
    const int64_t cell = t / nf;
    int64_t idx[3];
    idx[2] = cell % g.n[2];
    idx[1] = (cell / g.n[2]) % g.n[1];
    idx[0] = cell / (g.n[2] * g.n[1]);
    T acc = 0;
    for (int a = 0; a < 3; ++a) {
      if (!g.active[a]) continue;
      const int64_t s = g.stride[a];
      // Slab index: the element with axis a removed.
      int64_t slab = 0;
      for (int b = 0; b < 3; ++b) {
        if (b == a) continue;
        slab = slab * g.n[b] + idx[b];
      }
      slab = slab * nf + f;
      T prev, next;
      if (idx[a] > 0) {
        prev = y[t - s];
      } else {
        const double gv = g.glo[a] ? g.glo[a][slab] : NAN;
        prev = (gv == gv) ? y[t + s] + static_cast<T>(-2.0 * g.dx[a]) * static_cast<T>(gv) : T(0);
      }
      if (idx[a] < g.n[a] - 1) {
        next = y[t + s];
      } else {
        const double gv = g.ghi[a] ? g.ghi[a][slab] : NAN;
        next = (gv == gv) ? y[t - s] + static_cast<T>(2.0 * g.dx[a]) * static_cast<T>(gv) : T(0);
      }
      if (g.wden) {
        acc = acc + static_cast<T>(g.wsum[a][cell]) * (prev + next);
        if (g.wdiff[a]) acc = acc + static_cast<T>(g.wdiff[a][cell]) * (next - prev);
      } else {
        acc = acc + (prev + next) / static_cast<T>(g.dxsq[a]);
      }
    }
    acc = acc - lap[t];
    T v = acc / static_cast<T>(g.wden ? g.wden[cell] : g.coef);
    if (ycons) {
      const double c = ycons[t];
      if (c == c) v = static_cast<T>(c);
    }
    const double d = static_cast<double>(v) - static_cast<double>(y[t]);
    acc2 += d * d;
    yn[t] = v;
  }
  red[threadIdx.x] = acc2;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// sum of parts[0 .. n) in a fixed order into *out (one block).
__device__ void sum_in_order(const double *parts, int n, double *out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) acc += parts[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_jacobi_sum(const double *parts, int n, double *out) {
  sum_in_order(parts, n, out);
}

__global__ void k_jacobi_init(JacobiState *st) {
  if (threadIdx.x == 0) {
    st->sumsq = 0.0;
    st->last_diff = 0.0;
    st->sweeps = 0;
    st->done = 0;
    st->pad = 0;
  }
}

// After a sweep of n_parts blocks: add its partials in order, count it,
// stop once ||y_new - y_old||_2 <= tol (the reference's loop condition).
// A loop that is already done leaves the state as it is.
__global__ void k_jacobi_check(JacobiState *st, int n_parts, double tol) {
  if (st->done) return;
  sum_in_order(st->partials, n_parts, &st->sumsq);
  __syncthreads();
  if (threadIdx.x == 0) {
    st->sweeps += 1;
    st->last_diff = sqrt(st->sumsq);
    if (st->last_diff <= tol) st->done = 1;
  }
}

// out (=, +=, -=) coef * x, element by element: the coordinate-dependent
// factors of the curvilinear operators (numerical_differentiator.py:158-173,
// 229-308, 348-400, 482-590, 636-725, 771-870) applied to stencil results
// and to field components.
template <typename T>
__global__ void k_weighted_add(const T *x, const T *coef, T *out, int64_t count, int mode) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count;
       t += (int64_t)gridDim.x * blockDim.x)
    emit(out, t, coef[t] * x[t], mode);
}

// y[t] = v[t] where v[t] is not NaN: the static y constraints of a
// channel-last state (constraint.py:104-131 over the full grid).
template <typename T>
__global__ void k_set_values(T *y, const double *v, int64_t count) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double c = v[t];
    if (c == c) y[t] = static_cast<T>(c);
  }
}

// One iteration of the element-wise secant method that
// scipy.optimize.newton runs on arrays without a derivative (the implicit
// integrators, numerical_integrator.py:135-270).  mode 0: where q1 != q0,
// p_new = p1 - q1 (p1 - p) / (q1 - q0) and fail = |dp| >= tol; where
// q1 == q0 and the element is still active, p_new = (p1 + p) / 2; elements
// with q1 == q0 become inactive.  counters[0] += elements with q1 != q0,
// counters[1] += those still failing.  mode 1: p_new = (p1 + p) / 2
// everywhere (all residual differences were zero).
template <typename T>
__global__ void k_secant(const T *p, const T *p1, const T *q0, const T *q1, T *p_new,
                         unsigned char *active, unsigned char *fail, int64_t count, double tol,
                         int mode, unsigned long long *counters) {
  unsigned long long n_nz = 0, n_fail = 0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < count;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (mode == 1) {
      p_new[t] = (p1[t] + p[t]) / T(2);
      continue;
    }
    if (mode == 2) {  // the second start point; tol carries dx
      const T dx = static_cast<T>(tol);
      p_new[t] = p[t] * (T(1) + dx) + (p[t] >= T(0) ? dx : -dx);
      continue;
    }
    if (q1[t] != q0[t]) {
      const T dp = (q1[t] * (p1[t] - p[t])) / (q1[t] - q0[t]);
      p_new[t] = p1[t] - dp;
      const bool f = fabs(static_cast<double>(dp)) >= tol;
      fail[t] = f;
      n_nz += 1;
      n_fail += f;
    } else {
      p_new[t] = active[t] ? (p1[t] + p[t]) / T(2) : p[t];
      active[t] = 0;
    }
  }
  if (mode == 0) {
    if (n_nz) atomicAdd(&counters[0], n_nz);
    if (n_fail) atomicAdd(&counters[1], n_fail);
  }
}

}  // namespace pml

using namespace pml;

extern "C" int pml_stencil_first(int precision, const void *y, void *out, int64_t outer,
                                 int64_t n, int64_t inner, double dx, const double *lo,
                                 const double *hi, int mode, void *stream) {
  if (n <= 2) return fail(PML_E_INVALID, "y must contain at least 3 points along the axis");
  if (outer <= 0 || inner <= 0 || !y || !out || !(dx > 0))
    return fail(PML_E_INVALID, "bad stencil arguments");
  hipStream_t st = as_stream(stream);
  const int nb = grid_for(outer * n * inner);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_first<double>, dim3(nb), dim3(256), 0, st, (const double *)y,
                       (double *)out, outer, n, inner, 2.0 * dx, lo, hi, mode);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_first<float>, dim3(nb), dim3(256), 0, st, (const float *)y,
                       (float *)out, outer, n, inner, 2.0 * dx, lo, hi, mode);
  else
    return fail(PML_E_INVALID, "stencils take FP32 or FP64 arrays");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_weighted_add(int precision, const void *x, const void *coef, void *out,
                                int64_t count, int mode, void *stream) {
  if (count <= 0 || !x || !coef || !out) return fail(PML_E_INVALID, "bad weighted-add arguments");
  hipStream_t st = as_stream(stream);
  const int nb = grid_for(count);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_weighted_add<double>, dim3(nb), dim3(256), 0, st, (const double *)x,
                       (const double *)coef, (double *)out, count, mode);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_weighted_add<float>, dim3(nb), dim3(256), 0, st, (const float *)x,
                       (const float *)coef, (float *)out, count, mode);
  else
    return fail(PML_E_INVALID, "weighted add takes FP32 or FP64 arrays");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_set_values(int precision, void *y, const double *values, int64_t count,
                              void *stream) {
  if (count <= 0 || !y || !values) return fail(PML_E_INVALID, "bad set-values arguments");
  hipStream_t st = as_stream(stream);
  const int nb = grid_for(count);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_set_values<double>, dim3(nb), dim3(256), 0, st, (double *)y, values,
                       count);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_set_values<float>, dim3(nb), dim3(256), 0, st, (float *)y, values,
                       count);
  else
    return fail(PML_E_INVALID, "set values takes FP32 or FP64 arrays");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_secant_step(int precision, int mode, const void *p, const void *p1,
                               const void *q0, const void *q1, void *p_new, void *active,
                               void *failing, int64_t count, double tol, void *counters,
                               void *stream) {
  if (count <= 0 || !p || !p_new || (mode != 2 && !p1) ||
      (mode == 0 && (!q0 || !q1 || !active || !failing || !counters)) || mode < 0 || mode > 2)
    return fail(PML_E_INVALID, "bad secant arguments");
  hipStream_t st = as_stream(stream);
  const int nb = grid_for(count);
  auto *act = static_cast<unsigned char *>(active);
  auto *fl = static_cast<unsigned char *>(failing);
  auto *cnt = static_cast<unsigned long long *>(counters);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_secant<double>, dim3(nb), dim3(256), 0, st, (const double *)p,
                       (const double *)p1, (const double *)q0, (const double *)q1,
                       (double *)p_new, act, fl, count, tol, mode, cnt);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_secant<float>, dim3(nb), dim3(256), 0, st, (const float *)p,
                       (const float *)p1, (const float *)q0, (const float *)q1,
                       (float *)p_new, act, fl, count, tol, mode, cnt);
  else
    return fail(PML_E_INVALID, "secant step takes FP32 or FP64 arrays");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_stencil_second(int precision, const void *y, void *out, int64_t outer,
                                  int64_t n, int64_t inner, double dx, double dx2,
                                  const double *lo, const double *hi, int mode, void *stream) {
  if (n <= 2) return fail(PML_E_INVALID, "y must contain at least 3 points along the axis");
  if (outer <= 0 || inner <= 0 || !y || !out || !(dx > 0) || !(dx2 > 0))
    return fail(PML_E_INVALID, "bad stencil arguments");
  hipStream_t st = as_stream(stream);
  const int nb = grid_for(outer * n * inner);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_second<double>, dim3(nb), dim3(256), 0, st, (const double *)y,
                       (double *)out, outer, n, inner, dx, dx * dx2, lo, hi, mode);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_second<float>, dim3(nb), dim3(256), 0, st, (const float *)y,
                       (float *)out, outer, n, inner, dx, dx * dx2, lo, hi, mode);
  else
    return fail(PML_E_INVALID, "stencils take FP32 or FP64 arrays");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

// Geometry of a Jacobi sweep over a (shape[0], shape[1], shape[2], n_fields)
// array; returns PML_OK or the error, and the element count in *total.
static int jacobi_geom(const int64_t *shape, int n_fields, const double *dx,
                       const double *const *ghost_lo, const double *const *ghost_hi,
                       JacobiGeom *out, int64_t *total_out) {
  JacobiGeom g{};
  int64_t total = n_fields;
  g.coef = 0.0;
  for (int a = 0; a < 3; ++a) {
    g.n[a] = shape[a];
    if (shape[a] <= 0) return fail(PML_E_INVALID, "bad Jacobi shape");
    g.active[a] = shape[a] > 1;
    if (g.active[a] && shape[a] <= 2)
      return fail(PML_E_INVALID, "y must contain at least 3 points along all x axes");
    g.dx[a] = dx[a];
    g.dxsq[a] = dx[a] * dx[a];
    g.glo[a] = ghost_lo ? ghost_lo[a] : nullptr;
    g.ghi[a] = ghost_hi ? ghost_hi[a] : nullptr;
    if (g.active[a]) {
      if (!(dx[a] > 0)) return fail(PML_E_INVALID, "spacing must be positive");
      g.coef += 2.0 / g.dxsq[a];
    }
    total *= shape[a];
  }
  g.stride[2] = n_fields;
  g.stride[1] = shape[2] * n_fields;
  g.stride[0] = shape[1] * shape[2] * n_fields;
  *out = g;
  *total_out = total;
  return PML_OK;
}

extern "C" int pml_jacobi_sweep(int precision, const void *y_old, void *y_new, const void *lap,
                                const int64_t *shape, int n_fields, const double *dx,
                                const double *const *ghost_lo, const double *const *ghost_hi,
                                const double *ycons, double *sumsq_out, double *work,
                                void *stream) {
  if (!y_old || !y_new || !lap || !shape || !dx || !sumsq_out || !work || n_fields <= 0)
    return fail(PML_E_INVALID, "bad Jacobi arguments");
  JacobiGeom g{};
  int64_t total = 0;
  if (int rc = jacobi_geom(shape, n_fields, dx, ghost_lo, ghost_hi, &g, &total); rc != PML_OK)
    return rc;
  hipStream_t st = as_stream(stream);
  int nb = grid_for(total);
  if (nb > PML_REDUCE_SLOTS) nb = PML_REDUCE_SLOTS;
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_jacobi<double>, dim3(nb), dim3(256), 0, st, (const double *)y_old,
                       (double *)y_new, (const double *)lap, total, n_fields, g, ycons,
                       work, nullptr);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_jacobi<float>, dim3(nb), dim3(256), 0, st, (const float *)y_old,
                       (float *)y_new, (const float *)lap, total, n_fields, g, ycons,
                       work, nullptr);
  else
    return fail(PML_E_INVALID, "Jacobi sweep takes FP32 or FP64 arrays");
  PML_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_jacobi_sum, dim3(1), dim3(256), 0, st, work, nb, sumsq_out);
  PML_LAUNCH_CHECK();
  return PML_OK;
}

static int jacobi_solve(int precision, void *y, void *work, const void *lap,
                        const int64_t *shape, int n_fields, const double *dx,
                        const double *const *ghost_lo, const double *const *ghost_hi,
                        const double *ycons, double tol, int64_t max_sweeps, int64_t batch,
                        void *state, int64_t *sweeps_out, double *diff_out,
                        const double *const *w_sum, const double *const *w_diff,
                        const double *w_den, void *stream) {
  if (!y || !work || !lap || !shape || !dx || !state || !sweeps_out || n_fields <= 0 ||
      max_sweeps <= 0 || batch <= 0 || !(tol >= 0.0))
    return fail(PML_E_INVALID, "bad Jacobi solve arguments");
  if (precision != PML_F64 && precision != PML_F32)
    return fail(PML_E_INVALID, "Jacobi solve takes FP32 or FP64 arrays");
  JacobiGeom g{};
  int64_t total = 0;
  if (int rc = jacobi_geom(shape, n_fields, dx, ghost_lo, ghost_hi, &g, &total); rc != PML_OK)
    return rc;
  if (w_den) {
    if (!w_sum) return fail(PML_E_INVALID, "weighted Jacobi needs w_sum");
    for (int a = 0; a < 3; ++a) {
      if (g.active[a] && !w_sum[a])
        return fail(PML_E_INVALID, "weighted Jacobi needs w_sum for every axis");
      g.wsum[a] = w_sum[a];
      g.wdiff[a] = w_diff ? w_diff[a] : nullptr;
    }
    g.wden = w_den;
  }
  hipStream_t st = as_stream(stream);
  JacobiState *js = static_cast<JacobiState *>(state);
  int nb = grid_for(total);
  if (nb > PML_REDUCE_SLOTS) nb = PML_REDUCE_SLOTS;
  hipLaunchKernelGGL(k_jacobi_init, dim3(1), dim3(64), 0, st, js);
  JacobiState host;
  host.done = 0;
  host.sweeps = 0;
  host.last_diff = 0.0;
  int64_t issued = 0;
  while (issued < max_sweeps) {
    const int64_t n = std::min<int64_t>(batch, max_sweeps - issued);
    for (int64_t k = 0; k < n; ++k) {
      // sweep i (0-based) reads buffer i % 2 (y, work) and writes the other
      const bool odd = ((issued + k) & 1) != 0;
      const void *src = odd ? work : y;
      void *dst = odd ? y : work;
      if (precision == PML_F64)
        hipLaunchKernelGGL(k_jacobi<double>, dim3(nb), dim3(256), 0, st,
                           static_cast<const double *>(src), static_cast<double *>(dst),
                           static_cast<const double *>(lap), total, n_fields, g, ycons,
                           js->partials, &js->done);
      else
        hipLaunchKernelGGL(k_jacobi<float>, dim3(nb), dim3(256), 0, st,
                           static_cast<const float *>(src), static_cast<float *>(dst),
                           static_cast<const float *>(lap), total, n_fields, g, ycons,
                           js->partials, &js->done);
      hipLaunchKernelGGL(k_jacobi_check, dim3(1), dim3(256), 0, st, js, nb, tol);
    }
    issued += n;
    PML_LAUNCH_CHECK();
    PML_HIP_CHECK(hipMemcpyAsync(&host, js, offsetof(JacobiState, partials),
                                 hipMemcpyDeviceToHost, st));
    PML_HIP_CHECK(hipStreamSynchronize(st));
    if (host.done) break;
  }
  // The converged iterate is the output of sweep `sweeps` (1-based): it is
  // in work when that count is odd.
  const size_t esz = precision == PML_F64 ? 8 : 4;
  if (host.done && (host.sweeps & 1))
    PML_HIP_CHECK(hipMemcpyAsync(y, work, total * esz, hipMemcpyDeviceToDevice, st));
  else if (!host.done && (issued & 1))
    PML_HIP_CHECK(hipMemcpyAsync(y, work, total * esz, hipMemcpyDeviceToDevice, st));
  *sweeps_out = host.done ? host.sweeps : -issued;
  if (diff_out) *diff_out = host.last_diff;
  return PML_OK;
}

extern "C" int pml_jacobi_solve(int precision, void *y, void *work, const void *lap,
                                const int64_t *shape, int n_fields, const double *dx,
                                const double *const *ghost_lo, const double *const *ghost_hi,
                                const double *ycons, double tol, int64_t max_sweeps,
                                int64_t batch, void *state, int64_t *sweeps_out,
                                double *diff_out, void *stream) {
  return jacobi_solve(precision, y, work, lap, shape, n_fields, dx, ghost_lo, ghost_hi, ycons,
                      tol, max_sweeps, batch, state, sweeps_out, diff_out, nullptr, nullptr,
                      nullptr, stream);
}

extern "C" int pml_jacobi_solve_weighted(int precision, void *y, void *work, const void *lap,
                                         const int64_t *shape, int n_fields, const double *dx,
                                         const double *const *ghost_lo,
                                         const double *const *ghost_hi, const double *ycons,
                                         double tol, int64_t max_sweeps, int64_t batch,
                                         void *state, int64_t *sweeps_out, double *diff_out,
                                         const double *const *w_sum,
                                         const double *const *w_diff, const double *w_den,
                                         void *stream) {
  if (!w_sum || !w_den) return fail(PML_E_INVALID, "weighted Jacobi needs w_sum and w_den");
  return jacobi_solve(precision, y, work, lap, shape, n_fields, dx, ghost_lo, ghost_hi, ycons,
                      tol, max_sweeps, batch, state, sweeps_out, diff_out, w_sum, w_diff, w_den,
                      stream);
}
