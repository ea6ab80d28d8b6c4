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

// Corner offsets and weights of one query point (see above).  Offsets are
// in grid points, not elements.  Returns false for a NaN coordinate.
template <int D, bool kNearest>
__device__ __forceinline__ bool point_setup(const double *__restrict__ coords, const Axes &axes,
                                            const double *__restrict__ q, int64_t *off,
                                            double *w) {
  int64_t lo[D], hi[D];
  double y[D];
  bool finite = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const double v = q[a];
    const int64_t n = axes.n[a];
    const double *x = coords + axes.off[a];
    if (!(v == v)) {
      finite = false;
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
  if constexpr (kNearest) {
    int64_t o = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) o += (y[a] <= 0.5 ? lo[a] : hi[a]) * axes.stride[a];
    off[0] = o;
    w[0] = 1.0;
  } else {
#pragma unroll
    for (int k = 0; k < (1 << D); ++k) {
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
  }
  return finite;
}

// Snapshot stacks channel-last or field-major: one thread per query point
// and a chunk of snapshots.
template <int D, bool kNearest, typename T>
__global__ __launch_bounds__(kInterpBlock) void k_interp(
    const T *__restrict__ values, const double *__restrict__ coords, Axes axes,
    const double *__restrict__ xi, T *__restrict__ out, int64_t n_points, int64_t n_t,
    int64_t t_chunk, int64_t plane, int C) {
  constexpr int K = kNearest ? 1 : 1 << D;
  const int64_t p = (int64_t)blockIdx.x * kInterpBlock + threadIdx.x;
  if (p >= n_points) return;
  const int64_t t_begin = (int64_t)blockIdx.y * t_chunk;
  const int64_t t_end = t_begin + t_chunk < n_t ? t_begin + t_chunk : n_t;
  int64_t off[K];
  double w[K];
  if (!point_setup<D, kNearest>(coords, axes, xi + p * D, off, w)) {
    const T q = (T)__builtin_nan("");
    for (int64_t t = t_begin; t < t_end; ++t)
      for (int c = 0; c < C; ++c) out[(t * n_points + p) * C + c] = q;
    return;
  }
  for (int64_t t = t_begin; t < t_end; ++t) {
    const T *src = values + t * plane;
    T *dst = out + (t * n_points + p) * C;
    for (int c = 0; c < C; ++c) {
      if constexpr (kNearest) {
        dst[c] = src[off[0] + c * axes.fstride];
      } else {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) acc = acc + (double)src[off[k] + c * axes.fstride] * w[k];
        dst[c] = (T)acc;
      }
    }
  }
}

// Time-inner snapshot stacks ([point][field][t]): a workgroup takes 32
// query points and a tile of up to 128 snapshots (all of them for stacks
// of up to 128).  Lanes run along t, so every corner read is one contiguous
// run of the point's snapshots; the results pass through LDS (dynamic,
// ts x 33 values) and leave as rows of 32 consecutive points per snapshot.
constexpr int kTiPoints = 32;
constexpr int kTiMaxSnaps = 128;

template <int D, bool kNearest, typename T>
__global__ __launch_bounds__(kInterpBlock) void k_interp_ti(
    const T *__restrict__ vt, const double *__restrict__ coords, Axes axes,
    const double *__restrict__ xi, T *__restrict__ out, int64_t n_points, int64_t n_t, int ts,
    int C) {
  constexpr int K = kNearest ? 1 : 1 << D;
  constexpr int kRow = kTiPoints + 1;
  __shared__ int64_t s_off[kTiPoints][K];
  __shared__ double s_w[kTiPoints][K];
  __shared__ int s_ok[kTiPoints];
  extern __shared__ __align__(16) unsigned char s_raw[];
  T *s_val = reinterpret_cast<T *>(s_raw);  // [ts][kRow]
  const int tid = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * kTiPoints;
  const int64_t t0 = (int64_t)blockIdx.y * ts;
  const int n_loc = (int)(n_points - p0 < kTiPoints ? n_points - p0 : kTiPoints);
  const int t_loc = (int)(n_t - t0 < ts ? n_t - t0 : ts);
  if (tid < n_loc) {
    int64_t off[K];
    double w[K];
    s_ok[tid] = point_setup<D, kNearest>(coords, axes, xi + (p0 + tid) * D, off, w);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      s_off[tid][k] = off[k] * C * n_t;
      s_w[tid][k] = w[k];
    }
  }
  __syncthreads();
  // (pl, tt) of element e = pl * t_loc + tt, advanced by kInterpBlock per
  // iteration without a division
  const int step_pl = kInterpBlock / t_loc, step_tt = kInterpBlock % t_loc;
  const int total = n_loc * t_loc;
  for (int c = 0; c < C; ++c) {
    const T *src = vt + (int64_t)c * n_t + t0;
    int pl = tid / t_loc, tt = tid % t_loc;
    for (int e = tid; e < total; e += kInterpBlock) {
      T v;
      if (!s_ok[pl]) {
        v = (T)__builtin_nan("");
      } else if constexpr (kNearest) {
        v = src[s_off[pl][0] + tt];
      } else {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) acc = acc + (double)src[s_off[pl][k] + tt] * s_w[pl][k];
        v = (T)acc;
      }
      s_val[tt * kRow + pl] = v;
      pl += step_pl;
      tt += step_tt;
      if (tt >= t_loc) {
        tt -= t_loc;
        ++pl;
      }
    }
    __syncthreads();
    for (int e = tid; e < t_loc * kTiPoints; e += kInterpBlock) {
      const int r = e / kTiPoints, q = e % kTiPoints;
      if (q < n_loc) out[((t0 + r) * n_points + p0 + q) * C + c] = s_val[r * kRow + q];
    }
    __syncthreads();
  }
}

// values (channel-last or field-major) -> vt[(p * C + f) * n_t + t], one
// 64-point x 64-snapshot tile per workgroup and field, through LDS.
constexpr int kTrTile = 64;

template <typename T>
__global__ __launch_bounds__(kInterpBlock) void k_time_inner(const T *__restrict__ values,
                                                            T *__restrict__ vt, int64_t n_t,
                                                            int64_t points, int C,
                                                            int field_major) {
  __shared__ T tile[kTrTile][kTrTile + 1];
  const int tid = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * kTrTile;
  const int64_t t0 = (int64_t)blockIdx.y * kTrTile;
  const int c = blockIdx.z;
  const int pl = tid % kTrTile;
  for (int tl = tid / kTrTile; tl < kTrTile; tl += kInterpBlock / kTrTile) {
    const int64_t t = t0 + tl, p = p0 + pl;
    if (t < n_t && p < points)
      tile[tl][pl] = field_major ? values[(t * C + c) * points + p]
                                 : values[(t * points + p) * C + c];
  }
  __syncthreads();
  const int tl = tid % kTrTile;
  for (int q = tid / kTrTile; q < kTrTile; q += kInterpBlock / kTrTile) {
    const int64_t t = t0 + tl, p = p0 + q;
    if (t < n_t && p < points) vt[(p * C + c) * n_t + t] = tile[tl][q];
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
    for (int64_t p = (int64_t)blockIdx.x * kDiffBlock + threadIdx.x; p < points; p += stride) {
      for (int c = 0; c < C; ++c) {
        const int64_t s = kFieldMajor ? c * points + p : p * C + c;
        po[p * C + c] = ok ? pb[s] - pa[s] : (T)__builtin_nan("");
      }
    }
  }
}

template <int D, bool kNearest>
void launch_interp(int precision, int layout, const void *values, const double *coords,
                   const Axes &axes, const double *xi, void *out, int64_t n_points,
                   int64_t n_t, int64_t plane, int C, hipStream_t st) {
  if (layout == PML_LAYOUT_TIME_INNER) {
    // balanced snapshot tiles of at most kTiMaxSnaps
    const int64_t tiles = (n_t + kTiMaxSnaps - 1) / kTiMaxSnaps;
    const int ts = (int)((n_t + tiles - 1) / tiles);
    const dim3 grid((unsigned)((n_points + kTiPoints - 1) / kTiPoints), (unsigned)tiles);
    const size_t lds = (size_t)ts * (kTiPoints + 1) *
                       (precision == PML_F64 ? sizeof(double) : sizeof(float));
    if (precision == PML_F64)
      hipLaunchKernelGGL((k_interp_ti<D, kNearest, double>), grid, dim3(kInterpBlock), lds, st,
                         (const double *)values, coords, axes, xi, (double *)out, n_points,
                         n_t, ts, C);
    else
      hipLaunchKernelGGL((k_interp_ti<D, kNearest, float>), grid, dim3(kInterpBlock), lds, st,
                         (const float *)values, coords, axes, xi, (float *)out, n_points, n_t,
                         ts, C);
    return;
  }
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

extern "C" int pml_interp_regular(int precision, int method, int layout, const void *values,
                                  int64_t n_t, int n_dims, const int64_t *shape, int n_fields,
                                  const double *coords, const double *xi, int64_t n_points,
                                  void *out, void *stream) {
  if (precision != PML_F64 && precision != PML_F32)
    return fail(PML_E_INVALID, "interpolation takes FP32 or FP64 snapshots");
  if (method != PML_INTERP_LINEAR && method != PML_INTERP_NEAREST)
    return fail(PML_E_UNSUPPORTED, "interpolation method not on the device");
  if (layout != PML_LAYOUT_CHANNEL_LAST && layout != PML_LAYOUT_FIELD_MAJOR &&
      layout != PML_LAYOUT_TIME_INNER)
    return fail(PML_E_INVALID, "unknown snapshot layout");
  if (n_dims < 1 || n_dims > kMaxDims || !shape || n_fields < 1 || n_t < 1 || n_points < 0 ||
      !values || !coords || (n_points > 0 && (!xi || !out)))
    return fail(PML_E_INVALID, "bad interpolation arguments");
  if ((n_t + kTiMaxSnaps - 1) / kTiMaxSnaps > 65535)
    return fail(PML_E_INVALID, "too many snapshots");
  if (n_points == 0) return PML_OK;
  if ((n_points + kTiPoints - 1) / kTiPoints > 0x7fffffffLL)
    return fail(PML_E_INVALID, "too many interpolation points");
  Axes axes{};
  int64_t points = 1, off = 0;
  for (int a = n_dims - 1; a >= 0; --a) {
    if (shape[a] < 1) return fail(PML_E_INVALID, "empty interpolation axis");
    // time-inner offsets count grid points (the kernel scales them)
    axes.stride[a] = layout == PML_LAYOUT_CHANNEL_LAST ? points * n_fields : points;
    points *= shape[a];
  }
  axes.fstride = layout == PML_LAYOUT_FIELD_MAJOR ? points : 1;
  const int64_t plane = points * n_fields;
  for (int a = 0; a < n_dims; ++a) {
    axes.n[a] = shape[a];
    axes.off[a] = off;
    off += shape[a];
  }
  hipStream_t st = as_stream(stream);
  const bool nearest = method == PML_INTERP_NEAREST;
  switch (n_dims * 2 + (nearest ? 1 : 0)) {
    case 2: launch_interp<1, false>(precision, layout, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    case 3: launch_interp<1, true>(precision, layout, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    case 4: launch_interp<2, false>(precision, layout, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    case 5: launch_interp<2, true>(precision, layout, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    case 6: launch_interp<3, false>(precision, layout, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
    default: launch_interp<3, true>(precision, layout, values, coords, axes, xi, out, n_points, n_t, plane, n_fields, st); break;
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
  int64_t bx = (points + kDiffBlock - 1) / kDiffBlock;
  if (bx > 1024) bx = 1024;
  int64_t by = n_pairs < 65535 ? n_pairs : 65535;
  // At least ~2048 workgroups in flight when there are enough points.
  while (bx * by < 2048 && bx * kDiffBlock < points) bx *= 2;
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

extern "C" int pml_snapshots_time_inner(int precision, int layout, const void *values,
                                        int64_t n_t, int64_t points, int n_fields, void *out,
                                        void *stream) {
  if (precision != PML_F64 && precision != PML_F32)
    return fail(PML_E_INVALID, "the time-inner copy takes FP32 or FP64 snapshots");
  if (layout != PML_LAYOUT_CHANNEL_LAST && layout != PML_LAYOUT_FIELD_MAJOR)
    return fail(PML_E_INVALID, "the time-inner copy reads channel-last or field-major stacks");
  if (n_t < 0 || points < 0 || n_fields < 1 || n_fields > 65535)
    return fail(PML_E_INVALID, "bad time-inner copy arguments");
  if (n_t == 0 || points == 0) return PML_OK;
  if (!values || !out) return fail(PML_E_INVALID, "bad time-inner copy arguments");
  if ((n_t + kTrTile - 1) / kTrTile > 65535)
    return fail(PML_E_INVALID, "too many snapshots");
  const dim3 grid((unsigned)((points + kTrTile - 1) / kTrTile),
                  (unsigned)((n_t + kTrTile - 1) / kTrTile), (unsigned)n_fields);
  hipStream_t st = as_stream(stream);
  const int fm = layout == PML_LAYOUT_FIELD_MAJOR;
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_time_inner<double>, grid, dim3(kInterpBlock), 0, st,
                       (const double *)values, (double *)out, n_t, points, n_fields, fm);
  else
    hipLaunchKernelGGL(k_time_inner<float>, grid, dim3(kInterpBlock), 0, st,
                       (const float *)values, (float *)out, n_t, points, n_fields, fm);
  PML_LAUNCH_CHECK();
  return PML_OK;
}

// ---------------------------------------------------------------------------
// Rows of the supervised-learning data set (supervised_ml_operator.py:
// 362-379, 440-499): one workgroup per row r = b * points + p of block b
// (iteration-major, time step minor); the row is the recorded state of the
// block flattened channel-last, then the block's time column (when tcol is
// given), then the point's coordinates; its target is the next state at p.
namespace pml {
namespace {

constexpr int kRowBlock = 256;

__global__ __launch_bounds__(kRowBlock) void k_supervised_rows(
    const double *__restrict__ states, int64_t steps_per_state,
    const double *__restrict__ nexts, int64_t n_rows, int64_t n_steps, int64_t points, int C,
    const double *__restrict__ tcol, const double *__restrict__ coords, int x_dim,
    double *__restrict__ inputs, double *__restrict__ targets) {
  const int64_t n_y = points * C;
  const int64_t width = n_y + (tcol ? 1 : 0) + x_dim;
  for (int64_t r = blockIdx.x; r < n_rows; r += gridDim.x) {
    const int64_t b = r / points, p = r - b * points;
    const double *st = states + (b / steps_per_state) * n_y;
    double *row = inputs + r * width;
    if (C == 1) {
      for (int64_t col = threadIdx.x; col < n_y; col += kRowBlock) row[col] = st[col];
    } else {
      for (int64_t col = threadIdx.x; col < n_y; col += kRowBlock) {
        const int64_t q = col / C, c = col - q * C;
        row[col] = st[c * points + q];
      }
    }
    if (threadIdx.x < C) targets[r * C + threadIdx.x] = nexts[(b * C + threadIdx.x) * points + p];
    if (tcol && threadIdx.x == 0) row[n_y] = tcol[b % n_steps];
    if ((int)threadIdx.x < x_dim)
      row[n_y + (tcol ? 1 : 0) + threadIdx.x] = coords[p * x_dim + threadIdx.x];
  }
}

}  // namespace
}  // namespace pml

extern "C" int pml_supervised_rows(const double *states, int64_t steps_per_state,
                                   const double *nexts, int64_t n_blocks, int64_t n_steps,
                                   int64_t points, int n_fields, const double *tcol,
                                   const double *coords, int x_dim, double *inputs,
                                   double *targets, void *stream) {
  if (n_blocks < 0 || n_steps < 1 || points < 1 || n_fields < 1 || n_fields > kRowBlock ||
      x_dim < 0 || x_dim > kRowBlock || steps_per_state < 1 || n_blocks % n_steps != 0)
    return fail(PML_E_INVALID, "bad supervised-row arguments");
  if (n_blocks == 0) return PML_OK;
  if (!states || !nexts || !inputs || !targets || (x_dim > 0 && !coords))
    return fail(PML_E_INVALID, "bad supervised-row arguments");
  const int64_t rows = n_blocks * points;
  const unsigned grid = (unsigned)(rows < (1LL << 20) ? rows : (1LL << 20));
  hipLaunchKernelGGL(k_supervised_rows, dim3(grid), dim3(kRowBlock), 0, as_stream(stream),
                     states, steps_per_state, nexts, rows, n_steps, points, n_fields, tcol,
                     coords, x_dim, inputs, targets);
  PML_LAUNCH_CHECK();
  return PML_OK;
}
