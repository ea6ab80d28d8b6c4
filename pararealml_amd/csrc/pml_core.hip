// Library state, precision conversions, Parareal FP64 updates and 2x grid
// transfers for gfx950.
#include <hip/hip_runtime.h>

#include <string>

#include <cfloat>

#include "pml_common.h"

namespace pml {

static thread_local std::string t_error;

void set_error(const std::string &msg) { t_error = msg; }

int fail(int code, const std::string &msg) {
  set_error(msg);
  return code;
}

static int blocks_for(int64_t n, int nt) {
  int64_t b = (n + nt - 1) / nt;
  if (b > (1 << 20)) b = 1 << 20;  // grid-stride loops cover the rest
  return static_cast<int>(b < 1 ? 1 : b);
}

// ---- conversions ----------------------------------------------------------
__global__ void k_from_f64_f32(const double *s, float *d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = static_cast<float>(s[i]);
}
__global__ void k_from_f64_f32c(const double *s, float *hi, float *lo, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float h, l;
    split_f64(s[i], h, l);
    hi[i] = h;
    lo[i] = l;
  }
}
__global__ void k_copy_f64(const double *s, double *d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}
__global__ void k_to_f64_f32(const float *s, double *d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = static_cast<double>(s[i]);
}
__global__ void k_to_f64_f32c(const float *hi, const float *lo, double *d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = static_cast<double>(hi[i]) + static_cast<double>(lo[i]);
}

// ---- Parareal ---------------------------------------------------------------
// corr[s][i] = f_end(s, i) - g_end[s][i]   (parareal_operator.py:164)
template <int PREC>
__global__ void k_correction(const void *f_hi, const void *f_lo, int64_t f_stride,
                             const double *g, double *corr, int64_t n, int64_t n_slices) {
  const int64_t total = n * n_slices;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / n, i = t - s * n;
    const int64_t fi = s * f_stride + i;
    double f;
    if (PREC == PML_F64) f = static_cast<const double *>(f_hi)[fi];
    else if (PREC == PML_F32) f = static_cast<const float *>(f_hi)[fi];
    else
      f = static_cast<double>(static_cast<const float *>(f_hi)[fi]) +
          static_cast<double>(static_cast<const float *>(f_lo)[fi]);
    corr[t] = f - g[t];
  }
}

// u_next = g_new + corr; per field sum of squared updates against the
// previous u_next (parareal_operator.py:53-100, 183-185).  Each block
// reduces its share (grid-stride order, then a fixed tree) and writes one
// partial per field to work[f * gridDim.x + block]; k_sum_partials adds the
// partials in block order, so the norm does not depend on scheduling.
__global__ void k_update(const double *g, const double *corr, double *u, int64_t n,
                         int n_fields, double *work) {
  __shared__ double red[256];
  const int64_t per_field = n / n_fields;
  for (int f = 0; f < n_fields; ++f) {
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < per_field;
         i += (int64_t)gridDim.x * blockDim.x) {
      const int64_t e = f * per_field + i;
      const double nv = g[e] + corr[e];
      const double d = nv - u[e];
      acc += d * d;
      u[e] = nv;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) work[static_cast<int64_t>(f) * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

// Fallback test of a learned coarse prediction: per block the largest |src|,
// the largest |dst|, whether dst holds a non-finite value and whether src
// holds a NaN, written to work[4 * block + 0..3]; k_guard_flag merges the
// blocks and sets *flag when dst is not finite or (src has no NaN and)
// max|dst| > guard * max(max|src|, 1e-300) -- a NaN in src makes the ratio
// test false, as a comparison with a NaN bound is.
__device__ __forceinline__ void guard_tree(double *a, double *b, double *c, double *d) {
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      a[threadIdx.x] = fmax(a[threadIdx.x], a[threadIdx.x + s]);
      b[threadIdx.x] = fmax(b[threadIdx.x], b[threadIdx.x + s]);
      c[threadIdx.x] = fmax(c[threadIdx.x], c[threadIdx.x + s]);
      d[threadIdx.x] = fmax(d[threadIdx.x], d[threadIdx.x + s]);
    }
    __syncthreads();
  }
}

__global__ void k_guard_partials(const double *src, const double *dst, int64_t n,
                                 double *work) {
  __shared__ double s_src[256], s_dst[256], s_bad[256], s_nan[256];
  double ms = 0.0, md = 0.0, bad = 0.0, snan = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double a = fabs(src[i]), d = fabs(dst[i]);
    if (a == a) ms = a > ms ? a : ms;
    else snan = 1.0;
    if (d <= DBL_MAX) md = d > md ? d : md;
    else bad = 1.0;  // infinite or NaN
  }
  s_src[threadIdx.x] = ms;
  s_dst[threadIdx.x] = md;
  s_bad[threadIdx.x] = bad;
  s_nan[threadIdx.x] = snan;
  __syncthreads();
  guard_tree(s_src, s_dst, s_bad, s_nan);
  if (threadIdx.x == 0) {
    work[4 * blockIdx.x + 0] = s_src[0];
    work[4 * blockIdx.x + 1] = s_dst[0];
    work[4 * blockIdx.x + 2] = s_bad[0];
    work[4 * blockIdx.x + 3] = s_nan[0];
  }
}

__global__ void k_guard_flag(const double *work, int n_parts, double guard, int32_t *flag) {
  __shared__ double s_src[256], s_dst[256], s_bad[256], s_nan[256];
  double ms = 0.0, md = 0.0, bad = 0.0, snan = 0.0;
  for (int b = threadIdx.x; b < n_parts; b += blockDim.x) {
    ms = fmax(ms, work[4 * b + 0]);
    md = fmax(md, work[4 * b + 1]);
    bad = fmax(bad, work[4 * b + 2]);
    snan = fmax(snan, work[4 * b + 3]);
  }
  s_src[threadIdx.x] = ms;
  s_dst[threadIdx.x] = md;
  s_bad[threadIdx.x] = bad;
  s_nan[threadIdx.x] = snan;
  __syncthreads();
  guard_tree(s_src, s_dst, s_bad, s_nan);
  if (threadIdx.x == 0) {
    const double floor = s_src[0] > 1e-300 ? s_src[0] : 1e-300;
    if (s_bad[0] > 0.0 || (s_nan[0] == 0.0 && s_dst[0] > guard * floor)) *flag = 1;
  }
}

// out[f] += sum over b < n_parts of parts[f * n_parts + b], one block per f,
// in a fixed order (strided per-thread sums, then a fixed tree).
__global__ void k_sum_partials(const double *parts, int n_parts, double *out) {
  __shared__ double red[256];
  const double *p = parts + static_cast<int64_t>(blockIdx.x) * n_parts;
  double acc = 0.0;
  for (int b = threadIdx.x; b < n_parts; b += blockDim.x) acc += p[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] += red[0];
}

// ---- 2x grid transfer ---------------------------------------------------------
template <typename T>
__global__ void k_restrict(const T *fine, T *coarse, int64_t batch, int64_t m0, int64_t m1) {
  const int64_t f0 = 2 * m0 - 1, f1 = 2 * m1 - 1;
  const int64_t total = batch * m0 * m1;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / (m0 * m1), r = t - b * m0 * m1;
    const int64_t i = r / m1, j = r - i * m1;
    coarse[t] = fine[b * f0 * f1 + (2 * i) * f1 + 2 * j];
  }
}

template <typename T>
__global__ void k_prolong(const T *coarse, T *fine, int64_t batch, int64_t m0, int64_t m1) {
  const int64_t f0 = 2 * m0 - 1, f1 = 2 * m1 - 1;
  const int64_t total = batch * f0 * f1;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / (f0 * f1), r = t - b * f0 * f1;
    const int64_t i = r / f1, j = r - i * f1;
    const T *c = coarse + b * m0 * m1;
    const int64_t ci = i >> 1, cj = j >> 1;
    const bool oi = i & 1, oj = j & 1;
    T v;
    if (!oi && !oj) v = c[ci * m1 + cj];
    else if (oi && !oj) v = T(0.5) * (c[ci * m1 + cj] + c[(ci + 1) * m1 + cj]);
    else if (!oi && oj) v = T(0.5) * (c[ci * m1 + cj] + c[ci * m1 + cj + 1]);
    else
      v = T(0.25) * (c[ci * m1 + cj] + c[(ci + 1) * m1 + cj] + c[ci * m1 + cj + 1] +
                     c[(ci + 1) * m1 + cj + 1]);
    fine[t] = v;
  }
}

// Cubic 2x prolongation: fine point 2k along an axis copies coarse point
// k; fine point 2k + 1 is the 4-point Lagrange midpoint value
// ((-c[k-1] + 9 c[k]) + 9 c[k+1] - c[k+2]) / 16, one-sided at the two ends
// (((5 c[0] + 15 c[1]) - 5 c[2]) + c[3]) / 16 and its mirror image, exact
// for cubics; axis 0 first, then axis 1 (a tensor-product stencil of up to
// 4 x 4 coarse points).  Needs at least 4 coarse points per axis.
template <typename T>
__device__ __forceinline__ int cubic_taps(int i, int m, int *idx, T *w) {
  if (!(i & 1)) {
    idx[0] = i >> 1;
    w[0] = T(1);
    return 1;
  }
  const int k = i >> 1;  // midpoint between coarse k and k + 1
  if (k == 0) {
    idx[0] = 0; idx[1] = 1; idx[2] = 2; idx[3] = 3;
    w[0] = T(5); w[1] = T(15); w[2] = T(-5); w[3] = T(1);
  } else if (k + 1 == m - 1) {
    idx[0] = m - 1; idx[1] = m - 2; idx[2] = m - 3; idx[3] = m - 4;
    w[0] = T(5); w[1] = T(15); w[2] = T(-5); w[3] = T(1);
  } else {
    idx[0] = k - 1; idx[1] = k; idx[2] = k + 1; idx[3] = k + 2;
    w[0] = T(-1); w[1] = T(9); w[2] = T(9); w[3] = T(-1);
  }
  return 4;
}

template <typename T>
__device__ __forceinline__ T cubic_combine(int nt, const T *w, const T *v) {
  if (nt == 1) return v[0];
  // left to right, as the written formulas: ((w0 v0 + w1 v1) + w2 v2) + w3 v3
  T acc = w[0] * v[0] + w[1] * v[1];
  acc = acc + w[2] * v[2];
  acc = acc + w[3] * v[3];
  return acc / T(16);
}

template <typename T>
__global__ void k_prolong_cubic(const T *coarse, T *fine, int64_t batch, int64_t m0,
                                int64_t m1) {
  const int64_t f0 = 2 * m0 - 1, f1 = 2 * m1 - 1;
  const int64_t total = batch * f0 * f1;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / (f0 * f1), r = t - b * f0 * f1;
    const int i = static_cast<int>(r / f1), j = static_cast<int>(r - (r / f1) * f1);
    const T *c = coarse + b * m0 * m1;
    int ri[4], cj[4];
    T wi[4], wj[4];
    const int ni = cubic_taps<T>(i, static_cast<int>(m0), ri, wi);
    const int nj = cubic_taps<T>(j, static_cast<int>(m1), cj, wj);
    T col[4];
    for (int q = 0; q < nj; ++q) {
      T v[4];
      for (int p = 0; p < ni; ++p) v[p] = c[static_cast<int64_t>(ri[p]) * m1 + cj[q]];
      col[q] = cubic_combine<T>(ni, wi, v);
    }
    fine[t] = cubic_combine<T>(nj, wj, col);
  }
}

// Dirichlet values of a 2-D grid's edge arrays (the pml_grid boundary block
// layout) imposed on FP64 states of `batch` copies of the grid.
__global__ void k_apply_dirichlet(double *u, const double *bc, int n_fields, int n0, int n1,
                                  int ndim, int64_t batch) {
  const int64_t per = 2 * static_cast<int64_t>(n0 + n1);  // perimeter slots
  const int64_t total = batch * n_fields * per;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / (n_fields * per);
    const int64_t r = t - b * n_fields * per;
    const int f = static_cast<int>(r / per);
    const int64_t k = r - f * per;
    const double *fb = bc + static_cast<int64_t>(f) * 4 * (n0 + n1);
    int gi, gj;
    if (k < n1) { gi = 0; gj = static_cast<int>(k); }
    else if (k < 2 * n1) { gi = n0 - 1; gj = static_cast<int>(k - n1); }
    else if (k < 2 * n1 + n0) { gi = static_cast<int>(k - 2 * n1); gj = 0; }
    else { gi = static_cast<int>(k - 2 * n1 - n0); gj = n1 - 1; }
    if (ndim == 1 && k < 2 * n1) continue;  // 1-D: only the line's two ends
    double v = NAN;
    if (ndim == 2) {
      if (gi == 0) v = fb[gj];
      else if (gi == n0 - 1) v = fb[n1 + gj];
    }
    if (v != v) {
      if (gj == 0) v = fb[2 * n1 + gi];
      else if (gj == n1 - 1) v = fb[2 * n1 + n0 + gi];
    }
    if (v == v)
      u[(b * n_fields + f) * static_cast<int64_t>(n0) * n1 + static_cast<int64_t>(gi) * n1 + gj] = v;
  }
}

}  // namespace pml

using namespace pml;

extern "C" int pml_apply_dirichlet(double *state, const double *bc_values, int n_fields, int n0,
                                   int n1, int ndim, int64_t batch, void *stream) {
  if (!state || !bc_values || n_fields <= 0 || n0 <= 0 || n1 <= 0 || batch <= 0 ||
      (ndim != 1 && ndim != 2))
    return fail(PML_E_INVALID, "bad Dirichlet arguments");
  const int64_t total = batch * n_fields * 2 * static_cast<int64_t>(n0 + n1);
  hipLaunchKernelGGL(k_apply_dirichlet, dim3(blocks_for(total, 256)), dim3(256), 0,
                     as_stream(stream), state, bc_values, n_fields, n0, n1, ndim, batch);
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" const char *pml_last_error(void) { return t_error.c_str(); }

extern "C" int pml_version(void) { return 100; }

extern "C" int pml_init(int device) {
  int count = 0;
  PML_HIP_CHECK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(PML_E_INVALID, "device index out of range");
  PML_HIP_CHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  PML_HIP_CHECK(hipGetDeviceProperties(&prop, device));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
    return fail(PML_E_UNSUPPORTED, std::string("built for gfx950, found ") + prop.gcnArchName);
  return PML_OK;
}

extern "C" int pml_device_sync(void *stream) {
  PML_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
  return PML_OK;
}

extern "C" int pml_from_f64(int precision, const double *src, void *dst_hi, void *dst_lo,
                            int64_t n, void *stream) {
  if (n < 0 || (n && (!src || !dst_hi))) return fail(PML_E_INVALID, "bad conversion buffers");
  if (n == 0) return PML_OK;
  hipStream_t st = as_stream(stream);
  const int nb = blocks_for(n, 256);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_copy_f64, dim3(nb), dim3(256), 0, st, src, (double *)dst_hi, n);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_from_f64_f32, dim3(nb), dim3(256), 0, st, src, (float *)dst_hi, n);
  else if (precision == PML_F32C) {
    if (!dst_lo) return fail(PML_E_INVALID, "compensated conversion needs a lo plane");
    hipLaunchKernelGGL(k_from_f64_f32c, dim3(nb), dim3(256), 0, st, src, (float *)dst_hi,
                       (float *)dst_lo, n);
  } else
    return fail(PML_E_INVALID, "unknown precision");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_to_f64(int precision, const void *src_hi, const void *src_lo, double *dst,
                          int64_t n, void *stream) {
  if (n < 0 || (n && (!src_hi || !dst))) return fail(PML_E_INVALID, "bad conversion buffers");
  if (n == 0) return PML_OK;
  hipStream_t st = as_stream(stream);
  const int nb = blocks_for(n, 256);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_copy_f64, dim3(nb), dim3(256), 0, st, (const double *)src_hi, dst, n);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_to_f64_f32, dim3(nb), dim3(256), 0, st, (const float *)src_hi, dst, n);
  else if (precision == PML_F32C) {
    if (!src_lo) return fail(PML_E_INVALID, "compensated conversion needs a lo plane");
    hipLaunchKernelGGL(k_to_f64_f32c, dim3(nb), dim3(256), 0, st, (const float *)src_hi,
                       (const float *)src_lo, dst, n);
  } else
    return fail(PML_E_INVALID, "unknown precision");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_parareal_correction(int f_precision, const void *f_end_hi,
                                       const void *f_end_lo, int64_t f_stride,
                                       const double *g_end, double *corr, int64_t n,
                                       int64_t n_slices, void *stream) {
  if (n <= 0 || n_slices <= 0 || f_stride < n || !f_end_hi || !g_end || !corr)
    return fail(PML_E_INVALID, "bad correction arguments");
  if (f_precision == PML_F32C && !f_end_lo) return fail(PML_E_INVALID, "missing lo plane");
  hipStream_t st = as_stream(stream);
  const int nb = blocks_for(n * n_slices, 256);
  switch (f_precision) {
    case PML_F64:
      hipLaunchKernelGGL(k_correction<PML_F64>, dim3(nb), dim3(256), 0, st, f_end_hi, f_end_lo,
                         f_stride, g_end, corr, n, n_slices);
      break;
    case PML_F32:
      hipLaunchKernelGGL(k_correction<PML_F32>, dim3(nb), dim3(256), 0, st, f_end_hi, f_end_lo,
                         f_stride, g_end, corr, n, n_slices);
      break;
    case PML_F32C:
      hipLaunchKernelGGL(k_correction<PML_F32C>, dim3(nb), dim3(256), 0, st, f_end_hi, f_end_lo,
                         f_stride, g_end, corr, n, n_slices);
      break;
    default: return fail(PML_E_INVALID, "unknown precision");
  }
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_parareal_update(const double *g_new, const double *corr, double *u_next,
                                   int64_t n, int n_fields, double *sumsq, double *work,
                                   void *stream) {
  if (n <= 0 || n_fields <= 0 || n % n_fields != 0 || !g_new || !corr || !u_next || !sumsq ||
      !work)
    return fail(PML_E_INVALID, "bad update arguments");
  hipStream_t st = as_stream(stream);
  int nb = blocks_for(n / n_fields, 256);
  if (nb > PML_REDUCE_SLOTS) nb = PML_REDUCE_SLOTS;
  hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, st, g_new, corr, u_next, n, n_fields,
                     work);
  PML_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_partials, dim3(n_fields), dim3(256), 0, st, work, nb, sumsq);
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_slice_guard(const double *src, const double *dst, int64_t n, double guard,
                               double *work, int32_t *flag, void *stream) {
  if (n <= 0 || !src || !dst || !work || !flag || !(guard > 0.0))
    return fail(PML_E_INVALID, "bad guard arguments");
  hipStream_t st = as_stream(stream);
  int nb = blocks_for(n, 256);
  if (nb > PML_REDUCE_SLOTS) nb = PML_REDUCE_SLOTS;
  hipLaunchKernelGGL(k_guard_partials, dim3(nb), dim3(256), 0, st, src, dst, n, work);
  PML_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_guard_flag, dim3(1), dim3(256), 0, st, work, nb, guard, flag);
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_restrict2x(int precision, const void *fine, void *coarse, int64_t batch,
                              int64_t m0, int64_t m1, void *stream) {
  if (batch <= 0 || m0 < 2 || m1 < 2 || !fine || !coarse)
    return fail(PML_E_INVALID, "bad restriction arguments");
  hipStream_t st = as_stream(stream);
  const int nb = blocks_for(batch * m0 * m1, 256);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_restrict<double>, dim3(nb), dim3(256), 0, st, (const double *)fine,
                       (double *)coarse, batch, m0, m1);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_restrict<float>, dim3(nb), dim3(256), 0, st, (const float *)fine,
                       (float *)coarse, batch, m0, m1);
  else
    return fail(PML_E_INVALID, "restriction takes FP32 or FP64 planes");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_prolong2x(int precision, const void *coarse, void *fine, int64_t batch,
                             int64_t m0, int64_t m1, void *stream) {
  if (batch <= 0 || m0 < 2 || m1 < 2 || !fine || !coarse)
    return fail(PML_E_INVALID, "bad prolongation arguments");
  hipStream_t st = as_stream(stream);
  const int nb = blocks_for(batch * (2 * m0 - 1) * (2 * m1 - 1), 256);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_prolong<double>, dim3(nb), dim3(256), 0, st, (const double *)coarse,
                       (double *)fine, batch, m0, m1);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_prolong<float>, dim3(nb), dim3(256), 0, st, (const float *)coarse,
                       (float *)fine, batch, m0, m1);
  else
    return fail(PML_E_INVALID, "prolongation takes FP32 or FP64 planes");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_prolong2x_cubic(int precision, const void *coarse, void *fine,
                                   int64_t batch, int64_t m0, int64_t m1, void *stream) {
  if (batch <= 0 || m0 < 4 || m1 < 4 || !fine || !coarse)
    return fail(PML_E_INVALID, "bad cubic prolongation arguments (at least 4 coarse points "
                               "per axis)");
  hipStream_t st = as_stream(stream);
  const int nb = blocks_for(batch * (2 * m0 - 1) * (2 * m1 - 1), 256);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_prolong_cubic<double>, dim3(nb), dim3(256), 0, st,
                       (const double *)coarse, (double *)fine, batch, m0, m1);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_prolong_cubic<float>, dim3(nb), dim3(256), 0, st,
                       (const float *)coarse, (float *)fine, batch, m0, m1);
  else
    return fail(PML_E_INVALID, "prolongation takes FP32 or FP64 planes");
  PML_LAUNCH_CHECK();
  return PML_OK;
}
