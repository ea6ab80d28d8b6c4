// Solution post-processing on the device: snapshot stacks evaluated at
// arbitrary spatial points (Solution.y, pararealml/solution.py:114-145,
// which calls scipy.interpolate.interpn with bounds_error=False and
// fill_value=None) and time-matched snapshot differences (Solution.diff,
// solution.py:182-258).
//
// Interpolation restates SciPy 1.15's RegularGridInterpolator: per axis the
// interval i with x[i] <= v < x[i+1] (0 below the first coordinate, n-2 at
// and above the last, so points outside the mesh extrapolate linearly), the
// normalised distance y = (v - x[i]) / (x[i+1] - x[i]); "linear" sums the
// 2^D corner values times the products of (1 - y) or y over the axes, the
// corners in lexicographic order (axis 0 slowest, lower corner first) and
// the sum started from 0, "nearest" takes the lower corner where y <= 0.5.
// With FP contraction off, FP64 results equal SciPy's bit for bit.  A query
// point with a NaN coordinate gives NaN everywhere, an axis with a single
// coordinate contributes that coordinate with weight 1.
//
// Memory: one thread per query point and a chunk of snapshots; the point's
// corner offsets and weights are computed once and reused over the chunk,
// the output row of a wavefront (64 consecutive points, C values each) is
// one contiguous store.  The kernel is bound by the output stream when the
// queries are regular and by corner gathers through L2 when they are
// scattered.
#include <cstdint>

#include "pml_common.h"

namespace pml {
namespace {

constexpr int kInterpBlock = 256;
constexpr int kMaxDims = 3;
constexpr int kDiffBlock = 256;

struct Axes {
  int64_t n[kMaxDims];       // coordinates per axis
  int64_t off[kMaxDims];     // offset of the axis' coordinates in `coords`
  int64_t stride[kMaxDims];  // element stride of the axis in one snapshot
  int64_t fstride;           // element stride between the fields of a point
};

// The interval of an ascending coordinate array that holds v (see above);
// v is finite.  The first guess assumes uniform spacing and is then moved
// until x[i] <= v < x[i + 1] holds exactly, so non-uniform axes give the
// same interval (after a longer walk).
__device__ __forceinline__ int64_t find_interval(const double *__restrict__ x, int64_t n,
                                                 double v) {
  const double a = x[0], b = x[n - 1];
  if (v < a) return 0;
  if (v >= b) return n - 2;
  int64_t i = (int64_t)((v - a) / (b - a) * (double)(n - 1));
  i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
  while (i > 0 && v < x[i]) --i;
  while (i < n - 2 && v >= x[i + 1]) ++i;
  return i;
}

template <int D, bool kNearest, typename T>
__global__ __launch_bounds__(kInterpBlock) void k_interp(
    const T *__restrict__ values, const double *__restrict__ coords, Axes axes,
    const double *__restrict__ xi, T *__restrict__ out, int64_t n_points, int64_t n_t,
    int64_t t_chunk, int64_t plane, int C) {
  const int64_t p = (int64_t)blockIdx.x * kInterpBlock + threadIdx.x;
  if (p >= n_points) return;
  const int64_t t_begin = (int64_t)blockIdx.y * t_chunk;
  const int64_t t_end = t_begin + t_chunk < n_t ? t_begin + t_chunk : n_t;

  int64_t lo[D], hi[D];
  double y[D];
  bool nan = false;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const double v = xi[p * D + a];
    const int64_t n = axes.n[a];
    const double *x = coords + axes.off[a];
    if (!(v == v)) {
      nan = true;
      lo[a] = hi[a] = 0;
      y[a] = 0.0;
    } else if (n == 1) {
      lo[a] = hi[a] = 0;
      y[a] = 0.0;
    } else {
      const int64_t i = find_interval(x, n, v);
      lo[a] = i;
      hi[a] = i + 1;
      y[a] = (v - x[i]) / (x[i + 1] - x[i]);
    }
  }
  if (nan) {
    const T q = (T)__builtin_nan("");
    for (int64_t t = t_begin; t < t_end; ++t)
      for (int c = 0; c < C; ++c) out[(t * n_points + p) * C + c] = q;
    return;
  }

  if constexpr (kNearest) {
    int64_t idx = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) idx += (y[a] <= 0.5 ? lo[a] : hi[a]) * axes.stride[a];
    for (int64_t t = t_begin; t < t_end; ++t) {
      const T *src = values + t * plane + idx;
      T *dst = out + (t * n_points + p) * C;
      for (int c = 0; c < C; ++c) dst[c] = src[c * axes.fstride];
    }
  } else {
    constexpr int K = 1 << D;
    int64_t off[K];
    double w[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      int64_t o = 0;
      double wk = 1.0;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const bool upper = (k >> (D - 1 - a)) & 1;
        o += (upper ? hi[a] : lo[a]) * axes.stride[a];
        wk = wk * (upper ? y[a] : 1.0 - y[a]);
      }
      off[k] = o;
      w[k] = wk;
    }
    for (int64_t t = t_begin; t < t_end; ++t) {
      const T *src = values + t * plane;
      T *dst = out + (t * n_points + p) * C;
      for (int c = 0; c < C; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k)
          acc = acc + (double)src[off[k] + c * axes.fstride] * w[k];
        dst[c] = (T)acc;
      }
    }
  }
}

// out[k][p][f] = b[ib[k]](p, f) - a[ia[k]](p, f), the output channel-last;
// with kFieldMajor the inputs hold each snapshot as [f][p].  A pair whose
// index lies outside its stack gives NaN (the host validates the indices
// before the launch).
template <bool kFieldMajor, typename T>
__global__ __launch_bounds__(kDiffBlock) void k_snapshot_diff(
    const T *__restrict__ a, const int64_t *__restrict__ ia, int64_t n_a,
    const T *__restrict__ b, const int64_t *__restrict__ ib, int64_t n_b, T *__restrict__ out,
    int64_t n_pairs, int64_t points, int C) {
  const int64_t elems = points * C;
  const int64_t stride = (int64_t)gridDim.x * kDiffBlock;
  for (int64_t k = blockIdx.y; k < n_pairs; k += gridDim.y) {
    const int64_t ja = ia[k], jb = ib[k];
    const bool ok = ja >= 0 && ja < n_a && jb >= 0 && jb < n_b;
    const T *pa = a + (ok ? ja : 0) * elems;
    const T *pb = b + (ok ? jb : 0) * elems;
    T *po = out + k * elems;
    for (int64_t e = (int64_t)blockIdx.x * kDiffBlock + threadIdx.x; e < elems; e += stride) {
      int64_t s = e;
      if constexpr (kFieldMajor) s = (e % C) * points + e / C;
      po[e] = ok ? pb[s] - pa[s] : (T)__builtin_nan("");
    }
  }
}

template <int D, bool kNearest>
void launch_interp(int precision, const void *values, const double *coords, const Axes &axes,
                   const double *xi, void *out, int64_t n_points, int64_t n_t, int64_t plane,
                   int C, hipStream_t st) {
  const int64_t blocks_p = (n_points + kInterpBlock - 1) / kInterpBlock;
  // Split the snapshots into chunks when the points alone give too few
  // workgroups to fill 256 CUs.
  int64_t chunks = (2048 + blocks_p - 1) / blocks_p;
  if (chunks > n_t) chunks = n_t;
  if (chunks < 1) chunks = 1;
  const int64_t t_chunk = (n_t + chunks - 1) / chunks;
  chunks = (n_t + t_chunk - 1) / t_chunk;
  const dim3 grid((unsigned)blocks_p, (unsigned)chunks);
  if (precision == PML_F64)
    hipLaunchKernelGGL((k_interp<D, kNearest, double>), grid, dim3(kInterpBlock), 0, st,
                       (const double *)values, coords, axes, xi, (double *)out, n_points, n_t,
                       t_chunk, plane, C);
  else
    hipLaunchKernelGGL((k_interp<D, kNearest, float>), grid, dim3(kInterpBlock), 0, st,
                       (const float *)values, coords, axes, xi, (float *)out, n_points, n_t,
                       t_chunk, plane, C);
}

}  // namespace
}  // namespace pml

using namespace pml;

extern "C" int pml_interp_regular(int precision, int method, int field_major,
                                  const void *values, int64_t n_t, int n_dims,
                                  const int64_t *shape, int n_fields, const double *coords,
                                  const double *xi, int64_t n_points, void *out, void *stream) {
  if (precision != PML_F64 && precision != PML_F32)
    return fail(PML_E_INVALID, "interpolation takes FP32 or FP64 snapshots");
  if (method != PML_INTERP_LINEAR && method != PML_INTERP_NEAREST)
    return fail(PML_E_UNSUPPORTED, "interpolation method not on the device");
  if (n_dims < 1 || n_dims > kMaxDims || !shape || n_fields < 1 || n_t < 1 || n_points < 0 ||
      !values || !coords || (n_points > 0 && (!xi || !out)))
    return fail(PML_E_INVALID, "bad interpolation arguments");
  if (n_points == 0) return PML_OK;
  if ((n_points + kInterpBlock - 1) / kInterpBlock > 0x7fffffffLL)
    return fail(PML_E_INVALID, "too many interpolation points");
  Axes axes{};
  int64_t points = 1, off = 0;
  for (int a = n_dims - 1; a >= 0; --a) {
    if (shape[a] < 1) return fail(PML_E_INVALID, "empty interpolation axis");
    axes.stride[a] = field_major ? points : points * n_fields;
    points *= shape[a];
  }
  axes.fstride = field_major ? points : 1;
  const int64_t plane = points * n_fields;
  for (int a = 0; a < n_dims; ++a) {
    axes.n[a] = shape[a];
    axes.off[a] = off;
    off += shape[a];
  }
  hipStream_t st = as_stream(stream);
  const bool nearest = method == PML_INTERP_NEAREST;
  switch (n_dims * 2 + (nearest ? 1 : 0)) {
    case 2: launch_interp<1, false>(precision, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    case 3: launch_interp<1, true>(precision, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    case 4: launch_interp<2, false>(precision, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    case 5: launch_interp<2, true>(precision, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    case 6: launch_interp<3, false>(precision, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    default: launch_interp<3, true>(precision, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
  }
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_snapshot_diff(int precision, int field_major, const void *a,
                                 const int64_t *ia, int64_t n_a, const void *b,
                                 const int64_t *ib, int64_t n_b, void *out, int64_t n_pairs,
                                 int64_t points, int n_fields, void *stream) {
  if (precision != PML_F64 && precision != PML_F32)
    return fail(PML_E_INVALID, "differences take FP32 or FP64 snapshots");
  if (n_pairs < 0 || points < 0 || n_fields < 1 || n_a < 0 || n_b < 0)
    return fail(PML_E_INVALID, "bad difference arguments");
  const int64_t elems = points * n_fields;
  if (n_pairs == 0 || elems == 0) return PML_OK;
  if (!a || !b || !ia || !ib || !out) return fail(PML_E_INVALID, "bad difference arguments");
  int64_t bx = (elems + kDiffBlock - 1) / kDiffBlock;
  if (bx > 1024) bx = 1024;
  int64_t by = n_pairs < 65535 ? n_pairs : 65535;
  // At least ~2048 workgroups in flight when there are enough elements.
  while (bx * by < 2048 && bx * kDiffBlock < elems) bx *= 2;
  const dim3 grid((unsigned)bx, (unsigned)by);
  hipStream_t st = as_stream(stream);
  if (precision == PML_F64) {
    if (field_major)
      hipLaunchKernelGGL((k_snapshot_diff<true, double>), grid, dim3(kDiffBlock), 0, st,
                         (const double *)a, ia, n_a, (const double *)b, ib, n_b,
                         (double *)out, n_pairs, points, n_fields);
    else
      hipLaunchKernelGGL((k_snapshot_diff<false, double>), grid, dim3(kDiffBlock), 0, st,
                         (const double *)a, ia, n_a, (const double *)b, ib, n_b,
                         (double *)out, n_pairs, points, n_fields);
  } else {
    if (field_major)
      hipLaunchKernelGGL((k_snapshot_diff<true, float>), grid, dim3(kDiffBlock), 0, st,
                         (const float *)a, ia, n_a, (const float *)b, ib, n_b, (float *)out,
                         n_pairs, points, n_fields);
    else
      hipLaunchKernelGGL((k_snapshot_diff<false, float>), grid, dim3(kDiffBlock), 0, st,
                         (const float *)a, ia, n_a, (const float *)b, ib, n_b, (float *)out,
                         n_pairs, points, n_fields);
  }
  PML_LAUNCH_CHECK();
  return PML_OK;
}
