// Fused finite-difference time stepping kernels for gfx950.
//
// Two kernel families advance a (possibly ragged) batch of 2-D grids:
//
// * fdm_stream -- single-stage steps (forward Euler, and any integrator for
//   stage-invariant systems such as Cahn-Hilliard).  Barrier-free and
//   LDS-free: each wavefront owns a 64-column strip of RW rows, each lane one
//   column; the lane walks down its column through a register ring with
//   rows prefetched ahead, takes its left/right neighbours from the adjacent
//   lanes (DPP wave shifts) and reads and writes the state through per-row
//   buffer resources.  HBM sees one read and one write of the state per
//   step; the two halo rows per strip are mostly served by L2.
//
// * fdm_fused -- any stage count, with STEPS steps fused per launch.  One
//   workgroup loads an output tile plus a halo of H = stages * STEPS
//   vertices into LDS once, evaluates every Runge-Kutta stage on a region
//   that shrinks by one vertex per stage, and writes the output tile once.
//
// Boundary semantics restate pararealml/operators/fdm/numerical_differentiator.py:
//   * first derivatives use zero halos and are then overwritten on the
//     boundary by the Neumann values (1012-1057);
//   * second derivatives use ghosts y[1] - 2dx g_lo / y[n-2] + 2dx g_hi where
//     a Neumann value exists and 0 otherwise (1059-1095, 1188-1242);
// and the Dirichlet constraints are re-imposed on every stage input and on
// the step result (numerical_integrator.py:42-132).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "pml_common.h"
#include "pml_equations.h"

namespace pml {

struct GridDev {
  int n0, n1;
  int64_t off;  // state element offset of field 0
  int64_t bc;   // offset of the grid's boundary value block
  double inv_dx0sq, inv_dx1sq;
  double inv_2dx0, inv_2dx1;
  double m2dx0, m2dx1;  // -2 dx (lower ghost multiplier)
  double p2dx0, p2dx1;  // +2 dx (upper ghost multiplier)
};

// A work unit: an output tile (fdm_fused) or a strip of rows (fdm_stream)
// with origin (i0, j0) in grid g.
struct TileDev {
  int g, i0, j0, pad;
};

struct FdmArgs {
  const TileDev *tiles;
  const GridDev *grids;
  const double *bc;
  const void *in_hi;
  const void *in_lo;
  void *out_hi;
  void *out_lo;
  void *snap_hi;
  void *snap_lo;
  double dt;
  double p[6];
  int n_tiles;
  int ndim;
};

// Bijective block -> unit map that gives every XCD (blocks are dealt to the
// 8 XCDs round robin) a contiguous range of units, so that units sharing
// halo rows run behind the same L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n >> 3, rem = n & 7;
  const int x = b & 7, local = b >> 3;
  return x * q + min(x, rem) + local;
}

template <class P>
struct Store;
template <>
struct Store<PrecF64> {
  __device__ static void load(const void *hi_p, const void *, int64_t i, double &hi, double &lo) {
    hi = static_cast<const double *>(hi_p)[i];
    lo = 0.0;
  }
  __device__ static void store(void *p_hi, void *, int64_t i, double hi, double) {
    static_cast<double *>(p_hi)[i] = hi;
  }
};
template <>
struct Store<PrecF32> {
  __device__ static void load(const void *hi_p, const void *, int64_t i, float &hi, float &lo) {
    hi = static_cast<const float *>(hi_p)[i];
    lo = 0.0f;
  }
  __device__ static void store(void *p_hi, void *, int64_t i, float hi, float) {
    static_cast<float *>(p_hi)[i] = hi;
  }
};
template <>
struct Store<PrecF32C> {
  __device__ static void load(const void *hi_p, const void *lo_p, int64_t i, float &hi,
                              float &lo) {
    hi = static_cast<const float *>(hi_p)[i];
    lo = static_cast<const float *>(lo_p)[i];
  }
  __device__ static void store(void *p_hi, void *p_lo, int64_t i, float hi, float lo) {
    static_cast<float *>(p_hi)[i] = hi;
    static_cast<float *>(p_lo)[i] = lo;
  }
};

// y <- y + d for the state precision.
template <class P>
__device__ __forceinline__ void state_add(typename P::C &hi, typename P::C &lo,
                                          typename P::C d) {
  if constexpr (P::kComp) {
    comp_add(hi, lo, d);
  } else {
    hi = hi + d;
  }
}

// Working (single-word) value of y + d without changing y.
template <class P>
__device__ __forceinline__ typename P::C state_plus(typename P::C hi, typename P::C lo,
                                                    typename P::C d) {
  if constexpr (P::kComp) {
    return hi + (lo + d);
  } else {
    return hi + d;
  }
}

// Working value of y + d / 2 (the midpoint and RK4 stage inputs), with the
// halving fused into the add (exact: d / 2 needs no rounding).
template <class P>
__device__ __forceinline__ typename P::C state_plus_half(typename P::C hi, typename P::C lo,
                                                         typename P::C d) {
  using C = typename P::C;
  if constexpr (P::kComp) {
    return hi + fma_pow2(C(0.5), d, lo);
  } else {
    return fma_pow2(C(0.5), d, hi);
  }
}

template <class P>
__device__ __forceinline__ void state_set(typename P::C &hi, typename P::C &lo, double v) {
  if constexpr (P::kComp) {
    split_f64(v, hi, lo);
  } else {
    hi = static_cast<typename P::C>(v);
    lo = 0;
  }
}

// Dirichlet value of field f at boundary vertex (gi, gj); NaN when free.
__device__ __forceinline__ double dirichlet_at(const double *fb, int n0, int n1, int gi,
                                               int gj, bool ax0) {
  double v = NAN;
  if (ax0) {
    if (gi == 0) v = fb[gj];
    else if (gi == n0 - 1) v = fb[n1 + gj];
  }
  if (v != v) {
    if (gj == 0) v = fb[2 * n1 + gi];
    else if (gj == n1 - 1) v = fb[2 * n1 + n0 + gi];
  }
  return v;
}

template <typename C>
struct StencilK {
  C inv_dx0sq, inv_dx1sq, inv_2dx0, inv_2dx1, m2dx0, m2dx1, p2dx0, p2dx1;
  __device__ explicit StencilK(const GridDev &G)
      : inv_dx0sq(static_cast<C>(G.inv_dx0sq)),
        inv_dx1sq(static_cast<C>(G.inv_dx1sq)),
        inv_2dx0(static_cast<C>(G.inv_2dx0)),
        inv_2dx1(static_cast<C>(G.inv_2dx1)),
        m2dx0(static_cast<C>(G.m2dx0)),
        m2dx1(static_cast<C>(G.m2dx1)),
        p2dx0(static_cast<C>(G.p2dx0)),
        p2dx1(static_cast<C>(G.p2dx1)) {}
};

// Stencil quantities of one field at vertex (gi, gj) from its raw
// neighbours yw = (gi, gj-1), ye = (gi, gj+1), yn = (gi-1, gj),
// ys = (gi+1, gj); neighbours outside the grid are never used.  nb points at
// the field's Neumann edge arrays.
template <typename C, bool WL, bool WD>
__device__ __forceinline__ void field_quant(Quant<C> &q, C yc, C yw, C ye, C yn, C ys,
                                            bool lo1, bool hi1, bool lo0, bool hi0, bool ax0,
                                            const double *nb, int gi, int gj, int n0, int n1,
                                            const StencilK<C> &K) {
  q.v = yc;
  q.lap = 0;
  q.d0 = 0;
  q.d1 = 0;
  if (!WL && !WD) return;
  C lap1 = 0, d1 = 0, lap0 = 0, d0 = 0;
  if (!(lo1 || hi1 || lo0 || hi0)) {
    if (WL) lap1 = (fma_pow2(C(-2), yc, ye) + yw) * K.inv_dx1sq;
    if (WD) d1 = (ye - yw) * K.inv_2dx1;
    if (ax0) {
      if (WL) lap0 = (fma_pow2(C(-2), yc, ys) + yn) * K.inv_dx0sq;
      if (WD) d0 = (ys - yn) * K.inv_2dx0;
    }
  } else {
    const C zw = lo1 ? C(0) : yw, ze = hi1 ? C(0) : ye;
    const double g_lo1 = lo1 ? nb[2 * n1 + gi] : NAN;
    const double g_hi1 = hi1 ? nb[2 * n1 + n0 + gi] : NAN;
    if (WL) {
      C gw = zw, ge = ze;
      if (lo1) gw = (g_lo1 == g_lo1) ? ye + K.m2dx1 * static_cast<C>(g_lo1) : C(0);
      if (hi1) ge = (g_hi1 == g_hi1) ? yw + K.p2dx1 * static_cast<C>(g_hi1) : C(0);
      lap1 = (fma_pow2(C(-2), yc, ge) + gw) * K.inv_dx1sq;
    }
    if (WD) {
      d1 = (ze - zw) * K.inv_2dx1;
      if (lo1 && g_lo1 == g_lo1) d1 = static_cast<C>(g_lo1);
      if (hi1 && g_hi1 == g_hi1) d1 = static_cast<C>(g_hi1);
    }
    if (ax0) {
      const C zn = lo0 ? C(0) : yn, zs = hi0 ? C(0) : ys;
      const double g_lo0 = lo0 ? nb[gj] : NAN;
      const double g_hi0 = hi0 ? nb[n1 + gj] : NAN;
      if (WL) {
        C gn = zn, gs = zs;
        if (lo0) gn = (g_lo0 == g_lo0) ? ys + K.m2dx0 * static_cast<C>(g_lo0) : C(0);
        if (hi0) gs = (g_hi0 == g_hi0) ? yn + K.p2dx0 * static_cast<C>(g_hi0) : C(0);
        lap0 = (fma_pow2(C(-2), yc, gs) + gn) * K.inv_dx0sq;
      }
      if (WD) {
        d0 = (zs - zn) * K.inv_2dx0;
        if (lo0 && g_lo0 == g_lo0) d0 = static_cast<C>(g_lo0);
        if (hi0 && g_hi0 == g_hi0) d0 = static_cast<C>(g_hi0);
      }
    }
  }
  if (ax0) {
    q.lap = lap0 + lap1;
    q.d0 = d0;
    q.d1 = d1;
  } else {
    q.lap = lap1;
    q.d0 = d1;  // the single physical axis lies along axis 1
    q.d1 = 0;
  }
}

template <class Eq, int f>
struct Need {
  static constexpr bool kL = (Eq::kLap >> f) & 1u;
  static constexpr bool kD = (Eq::kD >> f) & 1u;
};

// All fields' quantities; Y(f, dir) yields the raw neighbour of field f.
template <class Eq, typename C, typename Getter>
__device__ __forceinline__ void all_quants(Quant<C> *q, const Getter &Y, bool lo1, bool hi1,
                                           bool lo0, bool hi0, bool ax0, const double *bcg,
                                           int64_t fstride, int gi, int gj, int n0, int n1,
                                           const StencilK<C> &K) {
  static_assert(Eq::F <= 4, "the stencil quantities are unrolled for up to 4 fields");
  const int64_t noff = 2 * static_cast<int64_t>(n0 + n1);
#define PML_FQ(F_)                                                                        \
  if constexpr (F_ < Eq::F) {                                                             \
    field_quant<C, Need<Eq, F_>::kL, Need<Eq, F_>::kD>(                                   \
        q[F_], Y(F_, 0), Y(F_, 1), Y(F_, 2), Y(F_, 3), Y(F_, 4), lo1, hi1, lo0, hi0, ax0, \
        bcg + F_ * fstride + noff, gi, gj, n0, n1, K);                                    \
  }
  PML_FQ(0)
  PML_FQ(1)
  PML_FQ(2)
  PML_FQ(3)
#undef PML_FQ
}

// ---------------------------------------------------------------------------
// fdm_stream: single-stage steps, one wavefront per 64-column strip of RW
// rows, one lane per column.  The lane walks down its column through a ring
// of PD + 3 row slots (rows i-1 .. i+1 in use, rows up to i+1+PD in flight);
// the row loop is unrolled by the ring length so that no slot is ever
// copied (a copy would wait for the prefetch).  Left and right neighbours
// come from the adjacent lanes by DPP wave shifts; lanes 0 and 63 load the
// strip's halo columns and feed them in as the shifted-in value.  Rows are
// handled by one of three wave-uniform bodies:
//   * interior strip (no grid column edge in the strip): plain stencils;
//   * edge strip: the row's axis-1 edge values are the same for every lane
//     (scalar loads) and are selected per lane, branch-free;
//   * first / last grid row: the generic per-vertex boundary code.
// Out-of-grid rows are never used (their loads are clamped into the grid).

// Per-row, per-field boundary values of the axis-1 edges (NaN = none).
template <typename C>
struct RowEdges {
  C n_lo, n_hi;       // Neumann values at (i, 0) and (i, n1-1)
  double d_lo, d_hi;  // Dirichlet values at (i, 0) and (i, n1-1)
};

// Stencil quantities of one field away from the first and last rows:
// identical arithmetic to field_quant.  kCols selects the column-edge
// handling (lo1 / hi1 with the row's edge values E).
template <typename C, bool WL, bool WD, bool kCols>
__device__ __forceinline__ void field_quant_row(Quant<C> &q, C yc, C yw, C ye, C yn, C ys,
                                                bool lo1, bool hi1, bool ax0,
                                                const RowEdges<C> &E, const StencilK<C> &K) {
  q.v = yc;
  q.lap = 0;
  q.d0 = 0;
  q.d1 = 0;
  if (!WL && !WD) return;
  C lap1 = 0, d1 = 0, lap0 = 0, d0 = 0;
  if (WL) {
    C gw = yw, ge = ye;
    if (kCols) {
      gw = lo1 ? (E.n_lo == E.n_lo ? ye + K.m2dx1 * E.n_lo : C(0)) : yw;
      ge = hi1 ? (E.n_hi == E.n_hi ? yw + K.p2dx1 * E.n_hi : C(0)) : ye;
    }
    lap1 = (fma_pow2(C(-2), yc, ge) + gw) * K.inv_dx1sq;
    if (ax0) lap0 = (fma_pow2(C(-2), yc, ys) + yn) * K.inv_dx0sq;
  }
  if (WD) {
    if (kCols) {
      const C zw = lo1 ? C(0) : yw, ze = hi1 ? C(0) : ye;
      d1 = (ze - zw) * K.inv_2dx1;
      d1 = (lo1 && E.n_lo == E.n_lo) ? E.n_lo : d1;
      d1 = (hi1 && E.n_hi == E.n_hi) ? E.n_hi : d1;
    } else {
      d1 = (ye - yw) * K.inv_2dx1;
    }
    if (ax0) d0 = (ys - yn) * K.inv_2dx0;
  }
  if (ax0) {
    q.lap = lap0 + lap1;
    q.d0 = d0;
    q.d1 = d1;
  } else {
    q.lap = lap1;
    q.d0 = d1;  // the single physical axis lies along axis 1
    q.d1 = 0;
  }
}

// Lane l receives lane l-1's x (lane 0: fill), or lane l+1's (lane 63: fill).
template <typename C>
__device__ __forceinline__ C wave_shr1(C x, C fill);
template <typename C>
__device__ __forceinline__ C wave_shl1(C x, C fill);
// DPP wave_shr:1 (0x138) / wave_shl:1 (0x130), bound_ctrl off: the lane
// whose source is outside the wavefront keeps `fill`.
template <>
__device__ __forceinline__ float wave_shr1<float>(float x, float fill) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill),
                                                               __builtin_bit_cast(int, x), 0x138,
                                                               0xf, 0xf, false));
}
template <>
__device__ __forceinline__ float wave_shl1<float>(float x, float fill) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill),
                                                               __builtin_bit_cast(int, x), 0x130,
                                                               0xf, 0xf, false));
}
template <>
__device__ __forceinline__ double wave_shr1<double>(double x, double fill) {
  const long long r = __builtin_amdgcn_update_dpp(__builtin_bit_cast(long long, fill),
                                                  __builtin_bit_cast(long long, x), 0x138, 0xf,
                                                  0xf, false);
  return __builtin_bit_cast(double, r);
}
template <>
__device__ __forceinline__ double wave_shl1<double>(double x, double fill) {
  const long long r = __builtin_amdgcn_update_dpp(__builtin_bit_cast(long long, fill),
                                                  __builtin_bit_cast(long long, x), 0x130, 0xf,
                                                  0xf, false);
  return __builtin_bit_cast(double, r);
}

// The same shifts with bound_ctrl on and a zero `old`: the lane whose
// source is outside the wavefront reads 0, and the shift can fold into the
// instruction that uses it (no copy of a fill value first).
template <typename C>
__device__ __forceinline__ C wave_shr1_z(C x);
template <typename C>
__device__ __forceinline__ C wave_shl1_z(C x);
template <>
__device__ __forceinline__ float wave_shr1_z<float>(float x) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, true));
}
template <>
__device__ __forceinline__ float wave_shl1_z<float>(float x) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, true));
}
template <>
__device__ __forceinline__ double wave_shr1_z<double>(double x) {
  const long long zero = 0;
  const long long r = __builtin_amdgcn_update_dpp(zero, __builtin_bit_cast(long long, x), 0x138,
                                                  0xf, 0xf, true);
  return __builtin_bit_cast(double, r);
}
template <>
__device__ __forceinline__ double wave_shl1_z<double>(double x) {
  const long long zero = 0;
  const long long r = __builtin_amdgcn_update_dpp(zero, __builtin_bit_cast(long long, x), 0x130,
                                                  0xf, 0xf, true);
  return __builtin_bit_cast(double, r);
}

// Value of lane `l` (wave-uniform index) of x, as a scalar.
template <typename T>
__device__ __forceinline__ T lane_value(T x, int l) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
  } else {
    static_assert(sizeof(T) == 8, "4- or 8-byte values");
    struct Two {
      int lo, hi;
    };
    Two v = __builtin_bit_cast(Two, x);
    v.lo = __builtin_amdgcn_readlane(v.lo, l);
    v.hi = __builtin_amdgcn_readlane(v.hi, l);
    return __builtin_bit_cast(T, v);
  }
}

// Raw buffer access: a resource over `bytes` bytes from p (gfx9 word 3:
// 32-bit data format, no stride); loads outside the range return 0 and
// stores outside it are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, static_cast<int>(bytes),
                                           0x00020000);
}

// N consecutive values of type C (4 * N * sizeof(C) / 4 bytes, one buffer
// instruction) at byte offset `off`.
template <typename C, int N>
struct Vec {
  C v[N];
};
template <typename C, int N>
__device__ __forceinline__ Vec<C, N> buf_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
  constexpr int kBytes = static_cast<int>(sizeof(C)) * N;
  if constexpr (kBytes == 4) {
    return __builtin_bit_cast(Vec<C, N>, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  } else if constexpr (kBytes == 8) {
    return __builtin_bit_cast(Vec<C, N>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  } else {
    static_assert(kBytes == 16, "4, 8 or 16 bytes per lane");
    return __builtin_bit_cast(Vec<C, N>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  }
}
template <typename C, int N>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned off,
                                          const Vec<C, N> &x) {
  constexpr int kBytes = static_cast<int>(sizeof(C)) * N;
  using u2 = unsigned __attribute__((ext_vector_type(2)));
  using u4 = unsigned __attribute__((ext_vector_type(4)));
  if constexpr (kBytes == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, off, 0, 0);
  } else if constexpr (kBytes == 8) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x), r, off, 0, 0);
  } else {
    static_assert(kBytes == 16, "4, 8 or 16 bytes per lane");
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, x), r, off, 0, 0);
  }
}

// CPL columns per lane: the strip is 64 CPL columns wide, lane l owns
// columns j0 + CPL l .. j0 + CPL l + CPL - 1 and moves them with one
// CPL-wide buffer instruction per row and field (CPL = 2 needs every grid's
// row length and state offset even, which the host checks).  SNAP compiles
// the snapshot stores in (launches that write a snapshot) or out.
template <class Eq, class Integ, class P, int RW, int PD, bool FULL, int CPL, bool SNAP>
__global__ __launch_bounds__(256) void fdm_stream(FdmArgs a) {
  using C = typename P::C;
  using V = Vec<C, CPL>;
  constexpr int F = Eq::F;
  constexpr int NW = PD + 3;  // ring slots
  constexpr int SW = 64 * CPL;  // strip width
  static_assert(Eq::kInvariant || Integ::kStages == 1, "single-stage steps only");
  static_assert(CPL == 1 || CPL == 2, "one or two columns per lane");
  const int lane = threadIdx.x & 63;
  // The unit is the same for the whole wavefront: keep it (and everything
  // derived from it) in scalar registers.
  const int unit = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x) * 4 +
                                                  static_cast<int>(threadIdx.x >> 6));
  if (unit >= a.n_tiles) return;  // whole wavefront: no barriers follow
  const TileDev U = a.tiles[unit];
  const GridDev G = a.grids[U.g];
  const int n0 = G.n0, n1 = G.n1;
  const int jl = U.j0 + CPL * lane;  // the lane's first column
  const bool ax0 = a.ndim == 2;
  const unsigned plane = static_cast<unsigned>(n0) * n1;
  const double *bcg = a.bc + G.bc;
  const int fstride = 4 * (n0 + n1);
  const StencilK<C> K(G);
  C pc[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) pc[i] = static_cast<C>(a.p[i]);
  const C dt = static_cast<C>(a.dt);
  const int i_beg = U.i0;
  const int i_end = min(U.i0 + RW, n0);
  const bool edge_strip = U.j0 == 0 || U.j0 + SW >= n1;  // wave-uniform
  bool lo1[CPL], hi1[CPL];
  int jj[CPL];  // columns for the per-lane boundary lookups (clamped)
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    lo1[c] = jl + c == 0;
    hi1[c] = jl + c == n1 - 1;
    jj[c] = min(jl + c, n1 - 1);
  }

  const C *in_hi = static_cast<const C *>(a.in_hi) + G.off;
  const C *in_lo = P::kComp ? static_cast<const C *>(a.in_lo) + G.off : nullptr;
  C *out_hi = static_cast<C *>(a.out_hi) + G.off;
  C *out_lo = P::kComp ? static_cast<C *>(a.out_lo) + G.off : nullptr;
  C *sn_hi = SNAP ? static_cast<C *>(a.snap_hi) + G.off : nullptr;
  C *sn_lo = (SNAP && P::kComp) ? static_cast<C *>(a.snap_lo) + G.off : nullptr;

  // Ring slot s holds row r with r = s (mod NW): values, lo parts, halo.
  V w[NW][F], wl[NW][F];
  C wh[NW][F];
  // Lane byte offsets into a row: the lane's columns and its halo column
  // (lanes 0 / 63: the columns left / right of the strip).  Offsets outside
  // the row read 0 (and store nothing) through the row's buffer range.
  const unsigned jb = static_cast<unsigned>(jl) * static_cast<unsigned>(sizeof(C));
  const int hcol = lane == 0 ? U.j0 - 1 : (lane == 63 ? U.j0 + SW : jl);
  const unsigned hb = static_cast<unsigned>(hcol) * static_cast<unsigned>(sizeof(C));
  const unsigned row_bytes = static_cast<unsigned>(n1) * static_cast<unsigned>(sizeof(C));
  // Row r of the input; rows outside [0, n0) and past `last` read as 0
  // without touching memory (empty buffer range), so that the load is
  // unconditional.
  auto load = [&](int r, int last, V *v, V *vl, C *h) {
    const bool ok = r >= 0 && r < n0 && r <= last;
    const size_t ro = static_cast<size_t>(min(max(r, 0), n0 - 1)) * n1;
    const unsigned nb = ok ? row_bytes : 0u;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const auto rs = row_rsrc(in_hi + static_cast<size_t>(f) * plane + ro, nb);
      v[f] = buf_load<C, CPL>(rs, jb);
      if (P::kComp)
        vl[f] = buf_load<C, CPL>(row_rsrc(in_lo + static_cast<size_t>(f) * plane + ro, nb), jb);
      h[f] = buf_load<C, 1>(rs, hb).v[0];
    }
  };

  // Neighbours of the lane's cells along axis 1: west[c], east[c].
  auto lateral = [&](const V *x, const C *h, C (*west)[CPL], C (*east)[CPL]) {
#pragma unroll
    for (int f = 0; f < F; ++f) {
      west[f][0] = wave_shr1<C>(x[f].v[CPL - 1], h[f]);
      east[f][CPL - 1] = wave_shl1<C>(x[f].v[0], h[f]);
#pragma unroll
      for (int c = 1; c < CPL; ++c) {
        west[f][c] = x[f].v[c - 1];
        east[f][c - 1] = x[f].v[c];
      }
    }
  };

  // Writes row i: update, Dirichlet values, store.
  auto finish = [&](int i, Quant<C> (*q)[F], const V *ct, const V *ctl, double (*dv)[F]) {
    V ho[F], lo[F];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      C rr[F], ra[F];
#pragma unroll
      for (int f = 0; f < F; ++f) rr[f] = ra[f] = 0;
      Eq::template rhs<C>(pc, q[c], rr);
      if (Eq::kAlg != 0u) Eq::template alg<C>(pc, q[c], ra);
#pragma unroll
      for (int f = 0; f < F; ++f) {
        C hv = ct[f].v[c], lv = P::kComp ? ctl[f].v[c] : C(0);
        if ((Eq::kAlg >> f) & 1u) {
          hv = ra[f];
          lv = 0;
        } else {
          const C kk = dt * rr[f];
          C d = kk;
          if (Eq::kInvariant && Integ::kStages == 4)
            d = (fma_pow2(C(2), kk, fma_pow2(C(2), kk, kk)) + kk) * C(1.0 / 6.0);
          state_add<P>(hv, lv, d);
        }
        if (dv && dv[c][f] == dv[c][f]) state_set<P>(hv, lv, dv[c][f]);
        ho[f].v[c] = hv;
        lo[f].v[c] = lv;
      }
    }
    const size_t ro = static_cast<size_t>(i) * n1;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const size_t fo = static_cast<size_t>(f) * plane + ro;
      buf_store<C, CPL>(row_rsrc(out_hi + fo, row_bytes), jb, ho[f]);
      if (P::kComp) buf_store<C, CPL>(row_rsrc(out_lo + fo, row_bytes), jb, lo[f]);
      if (SNAP) {
        buf_store<C, CPL>(row_rsrc(sn_hi + fo, row_bytes), jb, ho[f]);
        if (P::kComp) buf_store<C, CPL>(row_rsrc(sn_lo + fo, row_bytes), jb, lo[f]);
      }
    }
  };

  // The ring computes every row of the strip with the column-edge rules
  // only; the first and last grid rows (2-D) are computed again below with
  // the full boundary code, and their second store (same lane, later in
  // program order) is the one that remains.  The step body has no branch
  // but the wave-uniform strip kind, so the waits for the prefetched rows
  // are counted exactly.
  // i_beg is a multiple of RW, and RW of NW: row i_beg + k sits in slot
  // k mod NW at every pass.
  static_assert(RW % NW == 0, "rows per strip must be a multiple of the ring length");
  // Rows i_beg-1 .. i_beg+PD into the ring.
#pragma unroll
  for (int k = -1; k <= PD; ++k) {
    const int s = (k + NW) % NW;
    load(i_beg + k, i_end, w[s], wl[s], wh[s]);
  }
  // FULL unrolls the passes too (no loop back-edge).
#pragma unroll(FULL ? RW / NW : 1)
  for (int p = 0; p < RW / NW; ++p) {
    const int ib = i_beg + p * NW;
    if (ib >= i_end) break;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const int i = ib + k;
      if (i >= i_end) break;
      // Prefetch row i+1+PD into the slot of row i-2 (rows past i_end read
      // as 0).
      const int sp = (k + PD + 1) % NW, su = (k + NW - 1) % NW, sc = k % NW, sd = (k + 1) % NW;
      load(i + 1 + PD, i_end, w[sp], wl[sp], wh[sp]);
      C west[F][CPL], east[F][CPL];
      lateral(w[sc], wh[sc], west, east);
      Quant<C> q[CPL][F];
      if (edge_strip) {
        RowEdges<C> E[F];
        double dv[CPL][F];
#pragma unroll
        for (int f = 0; f < F; ++f) {
          // The boundary block is read-only during the launch: read the
          // row's (uniform) edge values through the scalar cache.
          using cdbl = const __attribute__((address_space(4))) double;
          cdbl *fb = (cdbl *)(bcg + f * fstride);
          cdbl *nb = fb + 2 * (n0 + n1);
          E[f].d_lo = fb[2 * n1 + i];
          E[f].d_hi = fb[2 * n1 + n0 + i];
          E[f].n_lo = static_cast<C>(nb[2 * n1 + i]);
          E[f].n_hi = static_cast<C>(nb[2 * n1 + n0 + i]);
#pragma unroll
          for (int c = 0; c < CPL; ++c)
            dv[c][f] = lo1[c] ? E[f].d_lo : (hi1[c] ? E[f].d_hi : static_cast<double>(NAN));
        }
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
#define PML_FQR(F_)                                                                         \
  if constexpr (F_ < F) {                                                                   \
    field_quant_row<C, Need<Eq, F_>::kL, Need<Eq, F_>::kD, true>(                           \
        q[c][F_], w[sc][F_].v[c], west[F_][c], east[F_][c], w[su][F_].v[c], w[sd][F_].v[c], \
        lo1[c], hi1[c], ax0, E[F_], K);                                                     \
  }
          PML_FQR(0)
          PML_FQR(1)
          PML_FQR(2)
          PML_FQR(3)
#undef PML_FQR
        }
        finish(i, q, w[sc], wl[sc], dv);
      } else {
        const RowEdges<C> E{};
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
#define PML_FQR(F_)                                                                         \
  if constexpr (F_ < F) {                                                                   \
    field_quant_row<C, Need<Eq, F_>::kL, Need<Eq, F_>::kD, false>(                          \
        q[c][F_], w[sc][F_].v[c], west[F_][c], east[F_][c], w[su][F_].v[c], w[sd][F_].v[c], \
        false, false, ax0, E, K);                                                           \
  }
          PML_FQR(0)
          PML_FQR(1)
          PML_FQR(2)
          PML_FQR(3)
#undef PML_FQR
        }
        finish(i, q, w[sc], wl[sc], nullptr);
      }
    }
  }

  // First and last grid rows (2-D): the generic per-vertex boundary code,
  // with their rows loaded again (cache hits).
  if (!ax0) return;
#pragma unroll 1
  for (int e = 0; e < 2; ++e) {
    const int i = e == 0 ? 0 : n0 - 1;
    if (i < i_beg || i >= i_end) continue;
    V u[F], cr[F], cl[F], d[F], dummy[F];
    C ch[F], hdummy[F];
    load(i - 1, n0, u, dummy, hdummy);
    load(i, n0, cr, cl, ch);
    load(i + 1, n0, d, dummy, hdummy);
    C west[F][CPL], east[F][CPL];
    lateral(cr, ch, west, east);
    Quant<C> q[CPL][F];
    double dv[CPL][F];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      auto Y = [&](int f, int dir) -> C {
        switch (dir) {
          case 0: return cr[f].v[c];
          case 1: return west[f][c];
          case 2: return east[f][c];
          case 3: return u[f].v[c];
          default: return d[f].v[c];
        }
      };
      all_quants<Eq, C>(q[c], Y, lo1[c], hi1[c], i == 0, i == n0 - 1, ax0, bcg, fstride, i,
                        jj[c], n0, n1, K);
#pragma unroll
      for (int f = 0; f < F; ++f)
        dv[c][f] = dirichlet_at(bcg + f * fstride, n0, n1, i, jj[c], ax0);
    }
    finish(i, q, cr, cl, dv);
  }
}

// ---------------------------------------------------------------------------
// fdm_stream2: NSTEP (1 or 2) single-stage steps per launch.  Lane l of a
// wavefront owns column c = j0 - 1 + l; the first step is computed on all
// 64 columns j0-1 .. j0+62 and the second on lanes 1 .. 62 (the output
// strip j0 .. j0+61; strips are 62 columns apart), so the second step's
// neighbours are all inside the wavefront.  Walking down the rows, the
// first step of row r (input rows r-1 .. r+1, from a ring of prefetched
// rows) and the second step of row r-1 (first-step rows r-2 .. r, from a
// second ring) are computed together, so the intermediate state never
// leaves the registers.  All boundary rules are selects: the axis-1 edge
// values of a row are scalar loads, the axis-0 edge values of the lane's
// column are loaded into registers once; rows and columns outside the grid
// only ever feed selects that replace them.

// Stencil quantities of one field with every edge rule as a select
// (identical arithmetic to field_quant): e_* are the row's axis-1 Neumann
// values, r_* the column's axis-0 Neumann values (NaN = none).
template <typename C, bool WL, bool WD>
__device__ __forceinline__ void field_quant_sel(Quant<C> &q, C yc, C yw, C ye, C yn, C ys,
                                                bool lo1, bool hi1, bool lo0, bool hi0,
                                                bool ax0, C e_lo, C e_hi, C r_lo, C r_hi,
                                                const StencilK<C> &K) {
  q.v = yc;
  q.lap = 0;
  q.d0 = 0;
  q.d1 = 0;
  if (!WL && !WD) return;
  C lap1 = 0, d1 = 0, lap0 = 0, d0 = 0;
  if (WL) {
    const C gw = lo1 ? (e_lo == e_lo ? ye + K.m2dx1 * e_lo : C(0)) : yw;
    const C ge = hi1 ? (e_hi == e_hi ? yw + K.p2dx1 * e_hi : C(0)) : ye;
    lap1 = (fma_pow2(C(-2), yc, ge) + gw) * K.inv_dx1sq;
    if (ax0) {
      const C gn = lo0 ? (r_lo == r_lo ? ys + K.m2dx0 * r_lo : C(0)) : yn;
      const C gs = hi0 ? (r_hi == r_hi ? yn + K.p2dx0 * r_hi : C(0)) : ys;
      lap0 = (fma_pow2(C(-2), yc, gs) + gn) * K.inv_dx0sq;
    }
  }
  if (WD) {
    const C zw = lo1 ? C(0) : yw, ze = hi1 ? C(0) : ye;
    d1 = (ze - zw) * K.inv_2dx1;
    d1 = (lo1 && e_lo == e_lo) ? e_lo : d1;
    d1 = (hi1 && e_hi == e_hi) ? e_hi : d1;
    if (ax0) {
      const C zn = lo0 ? C(0) : yn, zs = hi0 ? C(0) : ys;
      d0 = (zs - zn) * K.inv_2dx0;
      d0 = (lo0 && r_lo == r_lo) ? r_lo : d0;
      d0 = (hi0 && r_hi == r_hi) ? r_hi : d0;
    }
  }
  if (ax0) {
    q.lap = lap0 + lap1;
    q.d0 = d0;
    q.d1 = d1;
  } else {
    q.lap = lap1;
    q.d0 = d1;  // the single physical axis lies along axis 1
    q.d1 = 0;
  }
}

template <class Eq, class Integ, class P, int RW, int PD, int NSTEP, bool SNAP>
__global__ __launch_bounds__(256) void fdm_stream2(FdmArgs a) {
  using C = typename P::C;
  // Dirichlet values: exact FP64 where the state can hold them.
  using D = typename std::conditional<P::kComp || sizeof(C) == 8, double, C>::type;
  constexpr int F = Eq::F;
  constexpr int NW = PD + 3;  // slots of each ring
  constexpr int OW = 62;      // output columns per strip
  static_assert(Eq::kInvariant || Integ::kStages == 1, "single-stage steps only");
  static_assert(NSTEP == 1 || NSTEP == 2, "one or two steps per launch");
  static_assert(PD >= 1, "the ring must prefetch at least one row");
  static_assert(RW + 2 <= 64, "one lane per first-step row of the unit");
  const int lane = threadIdx.x & 63;
  const int unit = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x) * 4 +
                                                  static_cast<int>(threadIdx.x >> 6));
  if (unit >= a.n_tiles) return;  // whole wavefront: no barriers follow
  const TileDev U = a.tiles[unit];
  const GridDev G = a.grids[U.g];
  const int n0 = G.n0, n1 = G.n1;
  const bool ax0 = a.ndim == 2;
  const unsigned plane = static_cast<unsigned>(n0) * n1;
  const double *bcg = a.bc + G.bc;
  const int fstride = 4 * (n0 + n1);
  const StencilK<C> K(G);
  C pc[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) pc[i] = static_cast<C>(a.p[i]);
  const C dt = static_cast<C>(a.dt);
  const int i_beg = U.i0;
  const int i_end = min(U.i0 + RW, n0);
  const int col = U.j0 - 1 + lane;
  const bool lo1 = col == 0, hi1 = col == n1 - 1;
  const int cc = min(max(col, 0), n1 - 1);
  // Wave-uniform unit kinds: does any computed column / row lie on (or
  // beyond) a grid edge?
  const bool lo_edge = U.j0 - 1 <= 0, hi_edge = U.j0 + 62 >= n1 - 1;
  const bool edge_rows = !ax0 || i_beg - 1 <= 0 || i_end >= n0 - 1;
  // column kind: 1 / 2 the lower / upper axis-1 edge only, 3 both; units
  // that also reach a row edge take 3 (fewer compiled paths)
  const int cols = edge_rows ? (lo_edge || hi_edge ? 3 : 0)
                             : (lo_edge ? 1 : 0) | (hi_edge ? 2 : 0);

  const C *in_hi = static_cast<const C *>(a.in_hi) + G.off;
  const C *in_lo = P::kComp ? static_cast<const C *>(a.in_lo) + G.off : nullptr;
  C *out_hi = static_cast<C *>(a.out_hi) + G.off;
  C *out_lo = P::kComp ? static_cast<C *>(a.out_lo) + G.off : nullptr;
  C *sn_hi = SNAP ? static_cast<C *>(a.snap_hi) + G.off : nullptr;
  C *sn_lo = (SNAP && P::kComp) ? static_cast<C *>(a.snap_lo) + G.off : nullptr;
  const unsigned esz = static_cast<unsigned>(sizeof(C));
  const unsigned jb = static_cast<unsigned>(col) * esz;  // col = -1 reads 0
  const int hcol = lane == 0 ? U.j0 - 2 : (lane == 63 ? U.j0 + 63 : col);
  const unsigned hb = static_cast<unsigned>(hcol) * esz;
  // Only lanes 1 .. 62 store (lanes 0 and 63 carry the neighbour columns).
  const unsigned sb = (lane >= 1 && lane <= OW) ? jb : 0x80000000u;
  const unsigned row_bytes = static_cast<unsigned>(n1) * esz;

  auto load = [&](int r, C *v, C *vl, C *h) {
    const bool ok = r >= 0 && r < n0 && r <= i_end + 1;
    const size_t ro = static_cast<size_t>(min(max(r, 0), n0 - 1)) * n1;
    const unsigned nb = ok ? row_bytes : 0u;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const auto rs = row_rsrc(in_hi + static_cast<size_t>(f) * plane + ro, nb);
      v[f] = buf_load<C, 1>(rs, jb).v[0];
      if (P::kComp)
        vl[f] = buf_load<C, 1>(row_rsrc(in_lo + static_cast<size_t>(f) * plane + ro, nb), jb).v[0];
      h[f] = buf_load<C, 1>(rs, hb).v[0];
    }
  };
  auto store_row = [&](int r, const C *vh, const C *vl) {
    const size_t ro = static_cast<size_t>(r) * n1;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const size_t fo = static_cast<size_t>(f) * plane + ro;
      buf_store<C, 1>(row_rsrc(out_hi + fo, row_bytes), sb, Vec<C, 1>{{vh[f]}});
      if (P::kComp) buf_store<C, 1>(row_rsrc(out_lo + fo, row_bytes), sb, Vec<C, 1>{{vl[f]}});
      if (SNAP) {
        buf_store<C, 1>(row_rsrc(sn_hi + fo, row_bytes), sb, Vec<C, 1>{{vh[f]}});
        if (P::kComp) buf_store<C, 1>(row_rsrc(sn_lo + fo, row_bytes), sb, Vec<C, 1>{{vl[f]}});
      }
    }
  };

  // The row walk for one unit kind: kLoC / kHiC / kRows select the edge
  // rules compiled in (interior units run plain stencils; a unit beside one
  // axis-1 edge carries that edge's rules only).
  auto run = [&](auto cols_tag, auto rows_tag) {
    constexpr int kColKind = decltype(cols_tag)::value;
    constexpr bool kLoC = (kColKind & 1) != 0;
    constexpr bool kHiC = (kColKind & 2) != 0;
    constexpr bool kCols = kColKind != 0;
    constexpr bool kRows = decltype(rows_tag)::value;
    // Axis-0 edge values of the lane's column (rows 0 and n0-1).
    C rn_lo[F], rn_hi[F];
    D rd_lo[F], rd_hi[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      rn_lo[f] = rn_hi[f] = static_cast<C>(NAN);
      rd_lo[f] = rd_hi[f] = static_cast<D>(NAN);
      if (kRows && ax0) {
        const double *fb = bcg + f * fstride;
        const double *nb = fb + 2 * (n0 + n1);
        rd_lo[f] = static_cast<D>(fb[cc]);
        rd_hi[f] = static_cast<D>(fb[n1 + cc]);
        rn_lo[f] = static_cast<C>(nb[cc]);
        rn_hi[f] = static_cast<C>(nb[n1 + cc]);
      }
    }
    // Axis-1 edge values of the unit's first-step rows, lane k holding row
    // i_beg - 1 + k (loaded once; each step reads its row's values from
    // that lane, so the row walk issues no scalar memory traffic).
    C pe_lo[F], pe_hi[F];
    D pd_lo[F], pd_hi[F];
    if (kCols) {
      const int rk = min(max(i_beg - 1 + lane, 0), n0 - 1);
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const double *fb = bcg + f * fstride;
        const double *nb = fb + 2 * (n0 + n1);
        if (kLoC) {
          pd_lo[f] = static_cast<D>(fb[2 * n1 + rk]);
          pe_lo[f] = static_cast<C>(nb[2 * n1 + rk]);
        }
        if (kHiC) {
          pd_hi[f] = static_cast<D>(fb[2 * n1 + n0 + rk]);
          pe_hi[f] = static_cast<C>(nb[2 * n1 + n0 + rk]);
        }
      }
    }
    // One step of row r = i_beg - 1 + k from rows (yu, yc, yd) of the
    // previous state (hi parts; yc_lo the lo parts of row r), west / east
    // from the lanes.
    auto step_row = [&](int k, const C *yu, const C *yc, const C *yc_lo, const C *yd,
                        const C *fill, C *nh, C *nl) {
      const int r = i_beg - 1 + k;
      const bool lo0 = kRows && ax0 && r == 0, hi0 = kRows && ax0 && r == n0 - 1;
      C e_lo[F], e_hi[F];
      D d_lo[F], d_hi[F];
#pragma unroll
      for (int f = 0; f < F; ++f) {
        e_lo[f] = e_hi[f] = static_cast<C>(NAN);
        d_lo[f] = d_hi[f] = static_cast<D>(NAN);
        if (kLoC) {
          d_lo[f] = lane_value(pd_lo[f], k);
          e_lo[f] = lane_value(pe_lo[f], k);
        }
        if (kHiC) {
          d_hi[f] = lane_value(pd_hi[f], k);
          e_hi[f] = lane_value(pe_hi[f], k);
        }
      }
      const bool l1 = kLoC && lo1, h1 = kHiC && hi1;
      Quant<C> q[F];
#define PML_FQS(F_)                                                                          \
  if constexpr (F_ < F) {                                                                    \
    const C yw = fill ? wave_shr1<C>(yc[F_], fill[F_]) : wave_shr1_z<C>(yc[F_]);            \
    const C ye = fill ? wave_shl1<C>(yc[F_], fill[F_]) : wave_shl1_z<C>(yc[F_]);            \
    field_quant_sel<C, Need<Eq, F_>::kL, Need<Eq, F_>::kD>(q[F_], yc[F_], yw, ye, yu[F_],    \
                                                           yd[F_], l1, h1, lo0, hi0, ax0,    \
                                                           e_lo[F_], e_hi[F_], rn_lo[F_],    \
                                                           rn_hi[F_], K);                    \
  }
      PML_FQS(0)
      PML_FQS(1)
      PML_FQS(2)
      PML_FQS(3)
#undef PML_FQS
      C rr[F], ra[F];
#pragma unroll
      for (int f = 0; f < F; ++f) rr[f] = ra[f] = 0;
      Eq::template rhs<C>(pc, q, rr);
      if (Eq::kAlg != 0u) Eq::template alg<C>(pc, q, ra);
#pragma unroll
      for (int f = 0; f < F; ++f) {
        C hv = yc[f], lv = P::kComp ? yc_lo[f] : C(0);
        if ((Eq::kAlg >> f) & 1u) {
          hv = ra[f];
          lv = 0;
        } else {
          const C kk = dt * rr[f];
          C d = kk;
          if (Eq::kInvariant && Integ::kStages == 4)
            d = (fma_pow2(C(2), kk, fma_pow2(C(2), kk, kk)) + kk) * C(1.0 / 6.0);
          state_add<P>(hv, lv, d);
        }
        if (kCols || kRows) {
          // Dirichlet: the axis-0 edge value first, then the axis-1 one
          // (selects, no divergent branch).
          D dv = lo0 ? rd_lo[f] : (hi0 ? rd_hi[f] : static_cast<D>(NAN));
          dv = (dv == dv) ? dv : (l1 ? d_lo[f] : (h1 ? d_hi[f] : static_cast<D>(NAN)));
          C sh, sl;
          state_set<P>(sh, sl, static_cast<double>(dv));
          const bool fix = dv == dv;
          hv = fix ? sh : hv;
          lv = fix ? sl : lv;
        }
        nh[f] = hv;
        nl[f] = lv;
      }
    };

    // Input ring: row i_beg - 2 + m in slot m mod NW; first-step ring: row
    // i_beg - 1 + t in slot t mod NW.
    C w[NW][F], wl[NW][F], wh[NW][F];
    C ah[NW][F], al[NW][F];
#pragma unroll
    for (int m = 0; m <= PD + 1; ++m) load(i_beg - 2 + m, w[m % NW], wl[m % NW], wh[m % NW]);
    // Iteration t: first step of row r = i_beg - 1 + t, second step of row
    // r - 1 (NSTEP = 2).  Fully unrolled: every slot index is a constant.
#pragma unroll
    for (int t = 0; t <= RW + 1; ++t) {
      const int r = i_beg - 1 + t;
      if (NSTEP == 2 ? r - 1 >= i_end : r >= i_end) break;
      {
        const int m = t + 2 + PD;  // prefetch into the slot of input row m - NW
        load(i_beg - 2 + m, w[m % NW], wl[m % NW], wh[m % NW]);
      }
      const int su = t % NW, sc = (t + 1) % NW, sd = (t + 2) % NW;
      if (NSTEP == 1) {
        if (t == 0) continue;  // row i_beg - 1 is not written
        C nh[F], nl[F];
        step_row(t, w[su], w[sc], wl[sc], w[sd], wh[sc], nh, nl);
        store_row(r, nh, nl);
      } else {
        step_row(t, w[su], w[sc], wl[sc], w[sd], wh[sc], ah[t % NW], al[t % NW]);
        if (t >= 2) {
          C nh[F], nl[F];
          const int au = (t - 2) % NW, ac = (t - 1) % NW, ad = t % NW;
          // The fill (lanes 0 / 63) is never stored.
          // lanes 0 / 63 of the second step are not stored: no fill
          step_row(t - 1, ah[au], ah[ac], al[ac], ah[ad], nullptr, nh, nl);
          store_row(r - 1, nh, nl);
        }
      }
    }
  };
  using T = std::true_type;
  using Fa = std::false_type;
  using C0 = std::integral_constant<int, 0>;
  using C3 = std::integral_constant<int, 3>;
  if (edge_rows) {
    if (cols) run(C3{}, T{});
    else run(C0{}, T{});
  } else if (cols == 0) {
    run(C0{}, Fa{});
  } else if (cols == 1) {
    run(std::integral_constant<int, 1>{}, Fa{});
  } else if (cols == 2) {
    run(std::integral_constant<int, 2>{}, Fa{});
  } else {
    run(C3{}, Fa{});
  }
}

// ---------------------------------------------------------------------------
// fdm_stream_rk: one multi-stage step (explicit midpoint, RK4) per launch
// for systems whose right-hand side is re-evaluated at every stage.  A row
// pipeline: while walking down, stage s works on the row s-1 above stage 1,
// so every stage input Y_s stays in a three-row register ring and each row
// of the state crosses HBM once per step.  Stage s is valid on one column
// fewer per side than stage s-1.  With one or two fields, lane l owns column
// j0 - (S-1) + l and lanes 0 / 63 also load the halo columns beside the
// wavefront, so the result is written by lanes S-1 .. 64-S (strips of
// 66 - 2S columns).  With three or more fields registers limit occupancy:
// lane l owns column j0 - S + l and loads no halo, the result is written by
// lanes S .. 63-S (strips of 64 - 2S columns), which keeps the kernel at
// three waves per SIMD.  The axis-1 edge values of a row are held in one
// register per field (lower edge in lane k, upper in lane 32 + k).  The
// per-stage arithmetic, stage weights and Dirichlet rules are those of
// fdm_fused (bitwise-identical results); boundary rules are selects as in
// fdm_stream2, units are specialised on touching an edge.
template <class Eq, class Integ>
constexpr bool stream_rk_halo() {
  return Eq::F < 3;
}
// Lanes left of a stream_rk strip's first output column; strips are
// 64 - 2 * lead columns wide.
template <class Eq, class Integ>
constexpr int stream_rk_lead() {
  return stream_rk_halo<Eq, Integ>() ? Integ::kStages - 1 : Integ::kStages;
}
template <class Eq, class Integ, class P, int RW, int PD, bool SNAP>
__global__ __launch_bounds__(256) void fdm_stream_rk(FdmArgs a) {
  using C = typename P::C;
  using D = typename std::conditional<P::kComp || sizeof(C) == 8, double, C>::type;
  constexpr int F = Eq::F;
  constexpr int S = Integ::kStages;
  static_assert(!Eq::kInvariant && (S == 2 || S == 4), "multi-stage steps");
  static_assert(Eq::kAlg == 0u, "no algebraic fields");
  constexpr int NW = PD + S + 1;        // input ring slots
  constexpr int NK = S;                 // k-accumulator ring slots (RK4)
  constexpr int NT = RW + 2 * (S - 1);  // pipeline iterations
  constexpr bool kHalo = stream_rk_halo<Eq, Integ>();
  constexpr int kLead = stream_rk_lead<Eq, Integ>();
  static_assert(PD >= 1, "the ring must prefetch at least one row");
  static_assert(NT <= 32, "one lane per stage-1 row of the unit and edge");
  const int lane = threadIdx.x & 63;
  const int unit = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x) * 4 +
                                                  static_cast<int>(threadIdx.x >> 6));
  if (unit >= a.n_tiles) return;  // whole wavefront: no barriers follow
  const TileDev U = a.tiles[unit];
  const GridDev G = a.grids[U.g];
  const int n0 = G.n0, n1 = G.n1;
  const bool ax0 = a.ndim == 2;
  const unsigned plane = static_cast<unsigned>(n0) * n1;
  const double *bcg = a.bc + G.bc;
  const int fstride = 4 * (n0 + n1);
  const StencilK<C> K(G);
  C pc[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) pc[i] = static_cast<C>(a.p[i]);
  const C dt = static_cast<C>(a.dt);
  const C hdt = static_cast<C>(a.dt * 0.5);
  const int i_beg = U.i0;
  const int i_end = min(U.i0 + RW, n0);
  const int c0 = U.j0 - kLead;  // column of lane 0
  const int col = c0 + lane;
  const bool lo1 = col == 0, hi1 = col == n1 - 1;
  const int cc = min(max(col, 0), n1 - 1);
  const int r_first = i_beg - (S - 1);  // first stage-1 row
  // Which grid edges the unit's rows and lanes reach (wave-uniform): units
  // beside one axis-1 edge only run a path without the other edges' rules,
  // and when every row of the unit holds a Dirichlet value on that edge,
  // the edge column's own stencil is never used (its value is replaced at
  // every stage), so that path runs the interior stencils.
  const bool rows_edge = !ax0 || r_first <= 0 || i_end - 1 + (S - 1) >= n0 - 1;
  const bool lo_edge = c0 <= 0, hi_edge = c0 + 63 >= n1 - 1;
  int kind = rows_edge || (lo_edge && hi_edge) ? 3 : (lo_edge ? 1 : (hi_edge ? 2 : 0));
  if (kind == 1 || kind == 2) {
    // lane k (k < NT) holds row r_first + k of the lower edge, lane 32 + k
    // the same row of the upper edge
    const int rk = min(max(r_first + (lane & 31), 0), n0 - 1) + (lane < 32 ? 0 : n0);
    const unsigned long long rows = (1ull << NT) - 1ull;
    const unsigned long long want = kind == 1 ? rows : rows << 32;
    bool all = true;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const double v = bcg[f * fstride + 2 * n1 + rk];
      all = all && (__ballot(v == v) & want) == want;
    }
    if (all) kind += 3;  // 4: Dirichlet lower edge, 5: Dirichlet upper edge
  } else if (kind == 3 && ax0 && !lo_edge && !hi_edge) {
    // 6: the unit reaches row 0 and / or row n0 - 1 only, and every one of
    // its columns holds a Dirichlet value there: those rows are replaced at
    // every stage, the rows beyond them feed only them
    const bool top = r_first <= 0, bot = i_end - 1 + (S - 1) >= n0 - 1;
    bool all = true;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const double *fb = bcg + f * fstride;
      const double vt = fb[cc], vb = fb[n1 + cc];
      if (top) all = all && __ballot(vt == vt) == ~0ull;
      if (bot) all = all && __ballot(vb == vb) == ~0ull;
    }
    if (all) kind = 6;
  }

  const C *in_hi = static_cast<const C *>(a.in_hi) + G.off;
  const C *in_lo = P::kComp ? static_cast<const C *>(a.in_lo) + G.off : nullptr;
  C *out_hi = static_cast<C *>(a.out_hi) + G.off;
  C *out_lo = P::kComp ? static_cast<C *>(a.out_lo) + G.off : nullptr;
  C *sn_hi = SNAP ? static_cast<C *>(a.snap_hi) + G.off : nullptr;
  C *sn_lo = (SNAP && P::kComp) ? static_cast<C *>(a.snap_lo) + G.off : nullptr;
  const unsigned esz = static_cast<unsigned>(sizeof(C));
  const unsigned jb = static_cast<unsigned>(col) * esz;  // columns < 0 read 0
  const int hcol = lane == 0 ? c0 - 1 : (lane == 63 ? c0 + 64 : col);
  const unsigned hb = static_cast<unsigned>(hcol) * esz;
  const unsigned sb = (lane >= kLead && lane <= 63 - kLead) ? jb : 0x80000000u;
  const unsigned row_bytes = static_cast<unsigned>(n1) * esz;

  auto load = [&](int r, C *v, C *vl, C *h) {
    const bool ok = r >= 0 && r < n0 && r <= i_end - 1 + S;
    const size_t ro = static_cast<size_t>(min(max(r, 0), n0 - 1)) * n1;
    const unsigned nb = ok ? row_bytes : 0u;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const auto rs = row_rsrc(in_hi + static_cast<size_t>(f) * plane + ro, nb);
      v[f] = buf_load<C, 1>(rs, jb).v[0];
      if (P::kComp)
        vl[f] = buf_load<C, 1>(row_rsrc(in_lo + static_cast<size_t>(f) * plane + ro, nb), jb).v[0];
      if (kHalo) h[f] = buf_load<C, 1>(rs, hb).v[0];
    }
  };
  auto store_row = [&](int r, const C *vh, const C *vl) {
    const size_t ro = static_cast<size_t>(r) * n1;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const size_t fo = static_cast<size_t>(f) * plane + ro;
      buf_store<C, 1>(row_rsrc(out_hi + fo, row_bytes), sb, Vec<C, 1>{{vh[f]}});
      if (P::kComp) buf_store<C, 1>(row_rsrc(out_lo + fo, row_bytes), sb, Vec<C, 1>{{vl[f]}});
      if (SNAP) {
        buf_store<C, 1>(row_rsrc(sn_hi + fo, row_bytes), sb, Vec<C, 1>{{vh[f]}});
        if (P::kComp) buf_store<C, 1>(row_rsrc(sn_lo + fo, row_bytes), sb, Vec<C, 1>{{vl[f]}});
      }
    }
  };

  auto run = [&](auto kind_tag) {
    // 0: interior; 1 / 2: beside the lower / upper axis-1 edge only, rows
    // interior; 3: any edge (1-D grids included); 4 / 5: as 1 / 2 with a
    // Dirichlet value on every row of that edge; 6: beside the axis-0 edges
    // only, Dirichlet on all the unit's columns there
    constexpr int kKind = decltype(kind_tag)::value;
    constexpr bool kE = kKind != 0;
    constexpr bool kRows = kKind == 3;    // axis-0 edge stencil rules
    constexpr bool kRowDir = kKind == 6;  // axis-0 Dirichlet rows only
    constexpr bool kLo1 = kKind == 1 || kKind == 3;  // edge stencil rules
    constexpr bool kHi1 = kKind == 2 || kKind == 3;
    constexpr bool kNeedLo = kLo1 || kKind == 4;     // edge values needed
    constexpr bool kNeedHi = kHi1 || kKind == 5;
    // Edge values: the lane's column on rows 0 / n0-1, and the axis-1 edge
    // values of stage-1 row r_first + k, lower edge in lane k and upper edge
    // in lane 32 + k (one register for both).
    C rn_lo[F], rn_hi[F], pe[F];
    D rd_lo[F], rd_hi[F], pd[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      rn_lo[f] = rn_hi[f] = pe[f] = static_cast<C>(NAN);
      rd_lo[f] = rd_hi[f] = pd[f] = static_cast<D>(NAN);
      if (kE) {
        const double *fb = bcg + f * fstride;
        const double *nb = fb + 2 * (n0 + n1);
        if ((kRows || kRowDir) && ax0) {
          rd_lo[f] = static_cast<D>(fb[cc]);
          rd_hi[f] = static_cast<D>(fb[n1 + cc]);
        }
        if (kRows && ax0) {
          rn_lo[f] = static_cast<C>(nb[cc]);
          rn_hi[f] = static_cast<C>(nb[n1 + cc]);
        }
        const int rk = min(max(r_first + (lane & 31), 0), n0 - 1) + (lane < 32 ? 0 : n0);
        pd[f] = static_cast<D>(fb[2 * n1 + rk]);
        pe[f] = static_cast<C>(nb[2 * n1 + rk]);
      }
    }
    // dt-weighted right-hand side of row r = r_first + k from stage-input
    // rows (yu, yc, yd); also returns the row's Dirichlet values.  `fill`
    // feeds lanes 0 / 63 their outer neighbours (the halo columns); where
    // those lanes' results are not valid anyway it is null and the shifts
    // read 0 there.
    auto stage_k = [&](int k, const C *yu, const C *yc, const C *yd, const C *fill, C w, C *kk,
                       D *dv) {
      const int r = r_first + k;
      const bool lo0 = kRows && ax0 && r == 0, hi0 = kRows && ax0 && r == n0 - 1;
      const bool dlo0 = kRowDir && r == 0, dhi0 = kRowDir && r == n0 - 1;
      const bool l1 = kLo1 && lo1, h1 = kHi1 && hi1;
      C e_lo[F], e_hi[F];
#pragma unroll
      for (int f = 0; f < F; ++f) {
        e_lo[f] = e_hi[f] = static_cast<C>(NAN);
        dv[f] = static_cast<D>(NAN);
        if (kE) {
          // The broadcasts a path needs are read unconditionally: behind a
          // lane-varying condition the compiler would branch on EXEC
          // around each one.
          D d_l = static_cast<D>(NAN), d_h = static_cast<D>(NAN);
          if (kNeedLo) {
            if (kLo1) e_lo[f] = lane_value(pe[f], k);
            d_l = lane_value(pd[f], k);
          }
          if (kNeedHi) {
            if (kHi1) e_hi[f] = lane_value(pe[f], 32 + k);
            d_h = lane_value(pd[f], 32 + k);
          }
          D v = (lo0 || dlo0) ? rd_lo[f] : ((hi0 || dhi0) ? rd_hi[f] : static_cast<D>(NAN));
          v = (v == v) ? v
                       : ((kNeedLo && lo1) ? d_l
                                           : ((kNeedHi && hi1) ? d_h : static_cast<D>(NAN)));
          dv[f] = v;
        }
      }
      Quant<C> q[F];
#define PML_FQK(F_)                                                                          \
  if constexpr (F_ < F) {                                                                    \
    const C yw = fill ? wave_shr1<C>(yc[F_], fill[F_]) : wave_shr1_z<C>(yc[F_]);            \
    const C ye = fill ? wave_shl1<C>(yc[F_], fill[F_]) : wave_shl1_z<C>(yc[F_]);            \
    field_quant_sel<C, Need<Eq, F_>::kL, Need<Eq, F_>::kD>(q[F_], yc[F_], yw, ye, yu[F_],    \
                                                           yd[F_], l1, h1, lo0, hi0, ax0,    \
                                                           e_lo[F_], e_hi[F_], rn_lo[F_],    \
                                                           rn_hi[F_], K);                    \
  }
      PML_FQK(0)
      PML_FQK(1)
      PML_FQK(2)
      PML_FQK(3)
#undef PML_FQK
      C rr[F];
#pragma unroll
      for (int f = 0; f < F; ++f) rr[f] = 0;
      Eq::template rhs<C>(pc, q, rr);
#pragma unroll
      for (int f = 0; f < F; ++f) kk[f] = w * rr[f];
    };

    // Input ring: row i_beg - S + m in slot m mod NW.  Stage rings: Y_s
    // (s >= 2) of the row produced at iteration t in slot t mod 3.  k ring:
    // the accumulator of the row whose stage 1 ran at iteration t, slot
    // t mod NK.
    C w[NW][F], wl[NW][F], wh[kHalo ? NW : 1][F];
    C y[S > 1 ? S - 1 : 1][3][F];
    C kacc[NK][F];
#pragma unroll
    for (int m = 0; m <= PD + 1; ++m)
      load(i_beg - S + m, w[m % NW], wl[m % NW], wh[kHalo ? m % NW : 0]);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (i_beg - 2 * (S - 1) + t >= i_end) break;  // stage S is past the strip
      {
        const int m = t + 2 + PD;
        load(i_beg - S + m, w[m % NW], wl[m % NW], wh[kHalo ? m % NW : 0]);
      }
#pragma unroll
      for (int s = 1; s <= S; ++s) {
        if (t < 2 * (s - 1)) continue;  // stage s has no needed row yet
        const int k = t - (s - 1);      // stage-1 iteration of this row
        const int my = (t + 2 - s) % NW;  // the row's input slot
        C kk[F];
        D dv[F];
        if (s == 1) {
          // Stage 1 is valid on every lane when the halo columns are loaded.
          stage_k(k, w[t % NW], w[(t + 1) % NW], w[(t + 2) % NW],
                  kHalo ? wh[(t + 1) % NW] : nullptr, S == 2 ? hdt : dt, kk, dv);
        } else {
          // Y_s rows r-1, r, r+1 were produced at iterations t-2, t-1, t.
          const C *yu = y[s - 2][(t + 1) % 3], *yc = y[s - 2][(t + 2) % 3],
                  *yd = y[s - 2][t % 3];
          stage_k(k, yu, yc, yd, nullptr, dt, kk, dv);
        }
        C *acc = kacc[k % NK];
#pragma unroll
        for (int f = 0; f < F; ++f) {
          const C yh = w[my][f], yl = P::kComp ? wl[my][f] : C(0);
          // the Dirichlet-edge paths know their fixed lane (every row of
          // the edge holds a value): no FP64 compare
          const int rr_row = r_first + k;
          const bool fix = kKind == 4   ? lo1
                           : kKind == 5 ? hi1
                           : kKind == 6 ? (rr_row == 0 || rr_row == n0 - 1)
                                        : dv[f] == dv[f];
          if (s < S) {
            C nv;
            if (S == 2) {
              nv = state_plus<P>(yh, yl, kk[f]);
            } else if (s == 1) {
              acc[f] = kk[f];
              nv = state_plus_half<P>(yh, yl, kk[f]);
            } else if (s == 2) {
              acc[f] = fma_pow2(C(2), kk[f], acc[f]);
              nv = state_plus_half<P>(yh, yl, kk[f]);
            } else {
              acc[f] = fma_pow2(C(2), kk[f], acc[f]);
              nv = state_plus<P>(yh, yl, kk[f]);
            }
            y[s - 1][t % 3][f] = fix ? static_cast<C>(dv[f]) : nv;
          } else {
            C hv = yh, lv = yl;
            if (S == 2) {
              state_add<P>(hv, lv, kk[f]);
            } else {
              acc[f] = acc[f] + kk[f];
              state_add<P>(hv, lv, acc[f] * C(1.0 / 6.0));
            }
            C sh, sl;
            state_set<P>(sh, sl, static_cast<double>(dv[f]));
            kk[f] = fix ? sh : hv;  // the step's result (hi, lo)
            dv[f] = static_cast<D>(0);
            acc[f] = fix ? sl : lv;
          }
        }
        if (s == S) store_row(r_first + k, kk, acc);
      }
    }
  };
  if (kind == 0) run(std::integral_constant<int, 0>{});
  else if (kind == 6) run(std::integral_constant<int, 6>{});
  else if (kind == 4) run(std::integral_constant<int, 4>{});
  else if (kind == 5) run(std::integral_constant<int, 5>{});
  else if (kind == 1) run(std::integral_constant<int, 1>{});
  else if (kind == 2) run(std::integral_constant<int, 2>{});
  else run(std::integral_constant<int, 3>{});
}

// ---------------------------------------------------------------------------
// fdm_stream_n: NSTEP single-stage steps per launch as a row pipeline (the
// generalisation of fdm_stream2 to deeper temporal blocking): step s works
// on the row s-1 above step 1 and reads step s-1's rows from a three-row
// register ring (hi and lo parts), so the state crosses HBM once per NSTEP
// steps.  Lane l owns column j0 - (NSTEP-1) + l; step s is valid on lanes
// s-1 .. 64-s and the result is written by lanes NSTEP-1 .. 64-NSTEP
// (strips of 66 - 2 NSTEP columns).  Per-step arithmetic as fdm_stream2.
template <class Eq, class Integ, class P, int RW, int PD, int NSTEP, bool SNAP>
__global__ __launch_bounds__(256) void fdm_stream_n(FdmArgs a) {
  using C = typename P::C;
  using D = typename std::conditional<P::kComp || sizeof(C) == 8, double, C>::type;
  constexpr int F = Eq::F;
  constexpr int NW = PD + 3;             // input ring slots
  constexpr int NT = RW + 2 * (NSTEP - 1);  // pipeline iterations
  static_assert(Eq::kInvariant || Integ::kStages == 1, "single-stage steps only");
  static_assert(NSTEP >= 1 && NSTEP <= 4, "one to four steps per launch");
  static_assert(PD >= 1, "the ring must prefetch at least one row");
  static_assert(NT <= 64, "one lane per step-1 row of the unit");
  const int lane = threadIdx.x & 63;
  const int unit = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x) * 4 +
                                                  static_cast<int>(threadIdx.x >> 6));
  if (unit >= a.n_tiles) return;  // whole wavefront: no barriers follow
  const TileDev U = a.tiles[unit];
  const GridDev G = a.grids[U.g];
  const int n0 = G.n0, n1 = G.n1;
  const bool ax0 = a.ndim == 2;
  const unsigned plane = static_cast<unsigned>(n0) * n1;
  const double *bcg = a.bc + G.bc;
  const int fstride = 4 * (n0 + n1);
  const StencilK<C> K(G);
  C pc[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) pc[i] = static_cast<C>(a.p[i]);
  const C dt = static_cast<C>(a.dt);
  const int i_beg = U.i0;
  const int i_end = min(U.i0 + RW, n0);
  const int c0 = U.j0 - (NSTEP - 1);  // column of lane 0
  const int col = c0 + lane;
  const bool lo1 = col == 0, hi1 = col == n1 - 1;
  const int cc = min(max(col, 0), n1 - 1);
  const int r_first = i_beg - (NSTEP - 1);  // first step-1 row
  // Which grid edges the unit reaches (wave-uniform), as in fdm_stream_rk:
  // 0 interior; 1 / 2 one axis-1 edge only; 4 / 5 the same with a Dirichlet
  // value on all the unit's rows of it (interior stencils, the edge lane
  // replaced at every step); 6 row edges only, Dirichlet on all the unit's
  // columns there; 3 anything else (1-D grids included).
  const bool rows_edge = !ax0 || r_first <= 0 || i_end - 1 + (NSTEP - 1) >= n0 - 1;
  const bool lo_edge = c0 <= 0, hi_edge = c0 + 63 >= n1 - 1;
  int kind = rows_edge || (lo_edge && hi_edge) ? 3 : (lo_edge ? 1 : (hi_edge ? 2 : 0));
  if (kind == 1 || kind == 2) {
    const int rk = min(max(r_first + lane, 0), n0 - 1) + (kind == 1 ? 0 : n0);
    const unsigned long long want = NT >= 64 ? ~0ull : (1ull << NT) - 1ull;
    bool all = true;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const double v = bcg[f * fstride + 2 * n1 + rk];
      all = all && (__ballot(v == v) & want) == want;
    }
    if (all) kind += 3;
  } else if (kind == 3 && ax0 && !lo_edge && !hi_edge) {
    const bool top = r_first <= 0, bot = i_end - 1 + (NSTEP - 1) >= n0 - 1;
    bool all = true;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const double *fb = bcg + f * fstride;
      const double vt = fb[cc], vb = fb[n1 + cc];
      if (top) all = all && __ballot(vt == vt) == ~0ull;
      if (bot) all = all && __ballot(vb == vb) == ~0ull;
    }
    if (all) kind = 6;
  }

  const C *in_hi = static_cast<const C *>(a.in_hi) + G.off;
  const C *in_lo = P::kComp ? static_cast<const C *>(a.in_lo) + G.off : nullptr;
  C *out_hi = static_cast<C *>(a.out_hi) + G.off;
  C *out_lo = P::kComp ? static_cast<C *>(a.out_lo) + G.off : nullptr;
  C *sn_hi = SNAP ? static_cast<C *>(a.snap_hi) + G.off : nullptr;
  C *sn_lo = (SNAP && P::kComp) ? static_cast<C *>(a.snap_lo) + G.off : nullptr;
  const unsigned esz = static_cast<unsigned>(sizeof(C));
  const unsigned jb = static_cast<unsigned>(col) * esz;  // columns < 0 read 0
  const int hcol = lane == 0 ? c0 - 1 : (lane == 63 ? c0 + 64 : col);
  const unsigned hb = static_cast<unsigned>(hcol) * esz;
  const unsigned sb = (lane >= NSTEP - 1 && lane <= 64 - NSTEP) ? jb : 0x80000000u;
  const unsigned row_bytes = static_cast<unsigned>(n1) * esz;

  auto load = [&](int r, C *v, C *vl, C *h) {
    const bool ok = r >= 0 && r < n0 && r <= i_end + NSTEP - 1;
    const size_t ro = static_cast<size_t>(min(max(r, 0), n0 - 1)) * n1;
    const unsigned nb = ok ? row_bytes : 0u;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const auto rs = row_rsrc(in_hi + static_cast<size_t>(f) * plane + ro, nb);
      v[f] = buf_load<C, 1>(rs, jb).v[0];
      if (P::kComp)
        vl[f] = buf_load<C, 1>(row_rsrc(in_lo + static_cast<size_t>(f) * plane + ro, nb), jb).v[0];
      h[f] = buf_load<C, 1>(rs, hb).v[0];
    }
  };
  auto store_row = [&](int r, const C *vh, const C *vl) {
    const size_t ro = static_cast<size_t>(r) * n1;
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const size_t fo = static_cast<size_t>(f) * plane + ro;
      buf_store<C, 1>(row_rsrc(out_hi + fo, row_bytes), sb, Vec<C, 1>{{vh[f]}});
      if (P::kComp) buf_store<C, 1>(row_rsrc(out_lo + fo, row_bytes), sb, Vec<C, 1>{{vl[f]}});
      if (SNAP) {
        buf_store<C, 1>(row_rsrc(sn_hi + fo, row_bytes), sb, Vec<C, 1>{{vh[f]}});
        if (P::kComp) buf_store<C, 1>(row_rsrc(sn_lo + fo, row_bytes), sb, Vec<C, 1>{{vl[f]}});
      }
    }
  };

  auto run = [&](auto kind_tag) {
    constexpr int kKind = decltype(kind_tag)::value;
    constexpr bool kE = kKind != 0;
    constexpr bool kRows = kKind == 3;                 // axis-0 edge rules
    constexpr bool kRowDir = kKind == 6;               // axis-0 Dirichlet rows
    constexpr bool kLo1 = kKind == 1 || kKind == 3;    // axis-1 edge rules
    constexpr bool kHi1 = kKind == 2 || kKind == 3;
    constexpr bool kNeedLo = kLo1 || kKind == 4;       // axis-1 edge values
    constexpr bool kNeedHi = kHi1 || kKind == 5;
    C rn_lo[F], rn_hi[F], pe_lo[F], pe_hi[F];
    D rd_lo[F], rd_hi[F], pd_lo[F], pd_hi[F];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      rn_lo[f] = rn_hi[f] = pe_lo[f] = pe_hi[f] = static_cast<C>(NAN);
      rd_lo[f] = rd_hi[f] = pd_lo[f] = pd_hi[f] = static_cast<D>(NAN);
      if (kE) {
        const double *fb = bcg + f * fstride;
        const double *nb = fb + 2 * (n0 + n1);
        if ((kRows || kRowDir) && ax0) {
          rd_lo[f] = static_cast<D>(fb[cc]);
          rd_hi[f] = static_cast<D>(fb[n1 + cc]);
        }
        if (kRows && ax0) {
          rn_lo[f] = static_cast<C>(nb[cc]);
          rn_hi[f] = static_cast<C>(nb[n1 + cc]);
        }
        const int rk = min(max(r_first + lane, 0), n0 - 1);
        if (kNeedLo) pd_lo[f] = static_cast<D>(fb[2 * n1 + rk]);
        if (kNeedHi) pd_hi[f] = static_cast<D>(fb[2 * n1 + n0 + rk]);
        if (kLo1) pe_lo[f] = static_cast<C>(nb[2 * n1 + rk]);
        if (kHi1) pe_hi[f] = static_cast<C>(nb[2 * n1 + n0 + rk]);
      }
    }
    // One step of row r = r_first + k from rows (yu, yc, yd) of the
    // previous state (hi parts; yc_lo the lo parts of row r).
    auto step_row = [&](int k, const C *yu, const C *yc, const C *yc_lo, const C *yd,
                        const C *fill, C *nh, C *nl) {
      const int r = r_first + k;
      const bool lo0 = kRows && ax0 && r == 0, hi0 = kRows && ax0 && r == n0 - 1;
      const bool dlo0 = kRowDir && r == 0, dhi0 = kRowDir && r == n0 - 1;
      const bool l1 = kLo1 && lo1, h1 = kHi1 && hi1;
      C e_lo[F], e_hi[F];
      D d_lo[F], d_hi[F];
#pragma unroll
      for (int f = 0; f < F; ++f) {
        e_lo[f] = e_hi[f] = static_cast<C>(NAN);
        d_lo[f] = d_hi[f] = static_cast<D>(NAN);
        if (kNeedLo) d_lo[f] = lane_value(pd_lo[f], k);
        if (kNeedHi) d_hi[f] = lane_value(pd_hi[f], k);
        if (kLo1) e_lo[f] = lane_value(pe_lo[f], k);
        if (kHi1) e_hi[f] = lane_value(pe_hi[f], k);
      }
      Quant<C> q[F];
#define PML_FQN(F_)                                                                          \
  if constexpr (F_ < F) {                                                                    \
    const C yw = fill ? wave_shr1<C>(yc[F_], fill[F_]) : wave_shr1_z<C>(yc[F_]);            \
    const C ye = fill ? wave_shl1<C>(yc[F_], fill[F_]) : wave_shl1_z<C>(yc[F_]);            \
    field_quant_sel<C, Need<Eq, F_>::kL, Need<Eq, F_>::kD>(q[F_], yc[F_], yw, ye, yu[F_],    \
                                                           yd[F_], l1, h1, lo0, hi0, ax0,    \
                                                           e_lo[F_], e_hi[F_], rn_lo[F_],    \
                                                           rn_hi[F_], K);                    \
  }
      PML_FQN(0)
      PML_FQN(1)
      PML_FQN(2)
      PML_FQN(3)
#undef PML_FQN
      C rr[F], ra[F];
#pragma unroll
      for (int f = 0; f < F; ++f) rr[f] = ra[f] = 0;
      Eq::template rhs<C>(pc, q, rr);
      if (Eq::kAlg != 0u) Eq::template alg<C>(pc, q, ra);
#pragma unroll
      for (int f = 0; f < F; ++f) {
        C hv = yc[f], lv = P::kComp ? yc_lo[f] : C(0);
        if ((Eq::kAlg >> f) & 1u) {
          hv = ra[f];
          lv = 0;
        } else {
          const C kk = dt * rr[f];
          C d = kk;
          if (Eq::kInvariant && Integ::kStages == 4)
            d = (fma_pow2(C(2), kk, fma_pow2(C(2), kk, kk)) + kk) * C(1.0 / 6.0);
          state_add<P>(hv, lv, d);
        }
        if (kE) {
          D dv = (lo0 || dlo0) ? rd_lo[f] : ((hi0 || dhi0) ? rd_hi[f] : static_cast<D>(NAN));
          dv = (dv == dv) ? dv
                          : ((kNeedLo && lo1) ? d_lo[f]
                                              : ((kNeedHi && hi1) ? d_hi[f]
                                                                  : static_cast<D>(NAN)));
          C sh, sl;
          state_set<P>(sh, sl, static_cast<double>(dv));
          // the Dirichlet paths know their fixed lane / rows: no FP64 compare
          const bool fix = kKind == 4   ? lo1
                           : kKind == 5 ? hi1
                           : kKind == 6 ? (r == 0 || r == n0 - 1)
                                        : dv == dv;
          hv = fix ? sh : hv;
          lv = fix ? sl : lv;
        }
        nh[f] = hv;
        nl[f] = lv;
      }
    };

    // Input ring: row i_beg - NSTEP + m in slot m mod NW.  Level rings:
    // the output of step s (s < NSTEP) for the row produced at iteration t
    // in slot t mod 3.
    C w[NW][F], wl[NW][F], wh[NW][F];
    C yh[NSTEP > 1 ? NSTEP - 1 : 1][3][F], yl[NSTEP > 1 ? NSTEP - 1 : 1][3][F];
#pragma unroll
    for (int m = 0; m <= PD + 1; ++m)
      load(i_beg - NSTEP + m, w[m % NW], wl[m % NW], wh[m % NW]);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (i_beg - 2 * (NSTEP - 1) + t >= i_end) break;  // the last step is past the strip
      {
        const int m = t + 2 + PD;
        load(i_beg - NSTEP + m, w[m % NW], wl[m % NW], wh[m % NW]);
      }
#pragma unroll
      for (int s = 1; s <= NSTEP; ++s) {
        if (t < 2 * (s - 1)) continue;  // step s has no needed row yet
        const int k = t - (s - 1);
        C nh[F], nl[F];
        if (s == 1) {
          step_row(k, w[t % NW], w[(t + 1) % NW], wl[(t + 1) % NW], w[(t + 2) % NW],
                   wh[(t + 1) % NW], nh, nl);
        } else {
          // Step s-1 rows r-1, r, r+1 were produced at iterations t-2, t-1, t.
          const int su = (t + 1) % 3, sc = (t + 2) % 3, sd = t % 3;
          // lanes 0 / 63 are not valid after the first step: no fill
          step_row(k, yh[s - 2][su], yh[s - 2][sc], yl[s - 2][sc], yh[s - 2][sd], nullptr,
                   nh, nl);
        }
        if (s < NSTEP) {
#pragma unroll
          for (int f = 0; f < F; ++f) {
            yh[s - 1][t % 3][f] = nh[f];
            yl[s - 1][t % 3][f] = nl[f];
          }
        } else {
          store_row(r_first + k, nh, nl);
        }
      }
    }
  };
  if (kind == 0) run(std::integral_constant<int, 0>{});
  else if (kind == 6) run(std::integral_constant<int, 6>{});
  else if (kind == 4) run(std::integral_constant<int, 4>{});
  else if (kind == 5) run(std::integral_constant<int, 5>{});
  else if (kind == 1) run(std::integral_constant<int, 1>{});
  else if (kind == 2) run(std::integral_constant<int, 2>{});
  else run(std::integral_constant<int, 3>{});
}

// Imposes the Dirichlet values on a state in place: one workgroup per
// (grid, field) walks the grid's edge vertices.
template <class P>
__global__ __launch_bounds__(256) void fdm_constrain(const GridDev *grids, const double *bc,
                                                     void *st_hi, void *st_lo, int ndim) {
  using C = typename P::C;
  const GridDev G = grids[blockIdx.x];
  const int f = blockIdx.y;
  const int n0 = G.n0, n1 = G.n1;
  const bool ax0 = ndim == 2;
  const double *fb = bc + G.bc + static_cast<int64_t>(f) * 4 * (n0 + n1);
  const int64_t base = G.off + static_cast<int64_t>(f) * n0 * n1;
  const int n_edge = 2 * (n0 + n1);
  for (int t = threadIdx.x; t < n_edge; t += blockDim.x) {
    int i, jc;
    if (t < n1) {
      i = 0;
      jc = t;
    } else if (t < 2 * n1) {
      i = n0 - 1;
      jc = t - n1;
    } else if (t < 2 * n1 + n0) {
      i = t - 2 * n1;
      jc = 0;
    } else {
      i = t - 2 * n1 - n0;
      jc = n1 - 1;
    }
    const double dv = dirichlet_at(fb, n0, n1, i, jc, ax0);
    if (dv == dv) {
      C hi, lo;
      state_set<P>(hi, lo, dv);
      Store<P>::store(st_hi, st_lo, base + static_cast<int64_t>(i) * n1 + jc, hi, lo);
    }
  }
}

// ---------------------------------------------------------------------------
// fdm_fused: LDS tiles with shrinking stage regions, STEPS steps per launch.
template <class Eq, class Integ, class P, int OX, int OY, int STEPS, int NT>
__global__ __launch_bounds__(NT) void fdm_fused(FdmArgs a) {
  using C = typename P::C;
  constexpr int F = Eq::F;
  constexpr int S = Eq::kInvariant ? 1 : Integ::kStages;
  constexpr int H = S * STEPS;
  constexpr int EX = OX + 2 * H;
  constexpr int EY = OY + 2 * H;
  constexpr int CELLS = EX * EY;
  constexpr int PER = (CELLS + NT - 1) / NT;

  __shared__ C L[F * CELLS];

  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  if (tile >= a.n_tiles) return;
  const TileDev T = a.tiles[tile];
  const GridDev G = a.grids[T.g];
  const int n0 = G.n0, n1 = G.n1;
  const int64_t plane = static_cast<int64_t>(n0) * n1;
  const int bi = T.i0 - H, bj = T.j0 - H;
  const bool ax0 = a.ndim == 2;
  const double *bcg = a.bc + G.bc;
  const int64_t fstride = 4 * static_cast<int64_t>(n0 + n1);
  const StencilK<C> K(G);

  C pc[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) pc[i] = static_cast<C>(a.p[i]);
  const C dt = static_cast<C>(a.dt);
  const C hdt = static_cast<C>(a.dt * 0.5);

  C yh[PER][F], yl[PER][F], kacc[PER][F], kn[PER][F], an[PER][F];

  // ---- load the extended tile -------------------------------------------
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int idx = threadIdx.x + k * NT;
    if (idx < CELLS) {
      const int r = idx / EX, c = idx - (idx / EX) * EX;
      const int gi = bi + r, gj = bj + c;
      const bool in = gi >= 0 && gi < n0 && gj >= 0 && gj < n1;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        C hi = 0, lo = 0;
        if (in) {
          Store<P>::load(a.in_hi, a.in_lo, G.off + f * plane + static_cast<int64_t>(gi) * n1 + gj,
                         hi, lo);
        }
        yh[k][f] = hi;
        yl[k][f] = lo;
        L[f * CELLS + idx] = hi;
      }
    }
  }
  __syncthreads();

  for (int step = 0; step < STEPS; ++step) {
#pragma unroll
    for (int stage = 0; stage < S; ++stage) {
      const int hh = H - (step * S + stage + 1);
      const int rlo = H - hh, rhi = H + OY + hh, clo = H - hh, chi = H + OX + hh;
      // ---- evaluate the right-hand side on the stage region ------------
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = threadIdx.x + k * NT;
        const int r = idx / EX, c = idx - (idx / EX) * EX;
        const int gi = bi + r, gj = bj + c;
        if (idx < CELLS && r >= rlo && r < rhi && c >= clo && c < chi && gi >= 0 &&
            gi < n0 && gj >= 0 && gj < n1) {
          const bool lo1 = gj == 0, hi1 = gj == n1 - 1;
          const bool lo0 = ax0 && gi == 0, hi0 = ax0 && gi == n0 - 1;
          Quant<C> q[F];
          // Out-of-grid neighbours are never read (edge flags select ghosts).
          auto Y = [&](int f, int d) -> C {
            const C *Lf = L + f * CELLS;
            switch (d) {
              case 0: return Lf[idx];
              case 1: return lo1 ? C(0) : Lf[idx - 1];
              case 2: return hi1 ? C(0) : Lf[idx + 1];
              case 3: return (ax0 && !lo0) ? Lf[idx - EX] : C(0);
              default: return (ax0 && !hi0) ? Lf[idx + EX] : C(0);
            }
          };
          all_quants<Eq, C>(q, Y, lo1, hi1, lo0, hi0, ax0, bcg, fstride, gi, gj, n0, n1, K);
          C rr[F];
#pragma unroll
          for (int f = 0; f < F; ++f) rr[f] = 0;
          Eq::template rhs<C>(pc, q, rr);
          const C w = (!Eq::kInvariant && Integ::kStages == 2 && stage == 0) ? hdt : dt;
#pragma unroll
          for (int f = 0; f < F; ++f) kn[k][f] = w * rr[f];
          if (Eq::kAlg != 0u && stage == 0) {
            C ra[F];
#pragma unroll
            for (int f = 0; f < F; ++f) ra[f] = 0;
            Eq::template alg<C>(pc, q, ra);
#pragma unroll
            for (int f = 0; f < F; ++f) an[k][f] = ra[f];
          }
        }
      }
      __syncthreads();
      // ---- stage update ------------------------------------------------
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = threadIdx.x + k * NT;
        const int r = idx / EX, c = idx - (idx / EX) * EX;
        const int gi = bi + r, gj = bj + c;
        if (idx < CELLS && r >= rlo && r < rhi && c >= clo && c < chi && gi >= 0 &&
            gi < n0 && gj >= 0 && gj < n1) {
          const bool edge = gj == 0 || gj == n1 - 1 || (ax0 && (gi == 0 || gi == n0 - 1));
          const bool last = stage == S - 1;
#pragma unroll
          for (int f = 0; f < F; ++f) {
            const bool alg = (Eq::kAlg >> f) & 1u;
            C nv;  // new LDS value (stage input, or the step result)
            if (alg) {
              if (last) {
                yh[k][f] = an[k][f];
                yl[k][f] = 0;
              }
              nv = yh[k][f];
            } else if (Eq::kInvariant) {
              const C kk = kn[k][f];
              C d = kk;
              if (Integ::kStages == 4) d = (fma_pow2(C(2), kk, fma_pow2(C(2), kk, kk)) + kk) * C(1.0 / 6.0);
              state_add<P>(yh[k][f], yl[k][f], d);
              nv = yh[k][f];
            } else if (Integ::kStages == 1) {
              state_add<P>(yh[k][f], yl[k][f], kn[k][f]);
              nv = yh[k][f];
            } else if (Integ::kStages == 2) {
              if (stage == 0) {
                nv = state_plus<P>(yh[k][f], yl[k][f], kn[k][f]);
              } else {
                state_add<P>(yh[k][f], yl[k][f], kn[k][f]);
                nv = yh[k][f];
              }
            } else {  // RK4
              const C kk = kn[k][f];
              if (stage == 0) {
                kacc[k][f] = kk;
                nv = state_plus_half<P>(yh[k][f], yl[k][f], kk);
              } else if (stage == 1) {
                kacc[k][f] = fma_pow2(C(2), kk, kacc[k][f]);
                nv = state_plus_half<P>(yh[k][f], yl[k][f], kk);
              } else if (stage == 2) {
                kacc[k][f] = fma_pow2(C(2), kk, kacc[k][f]);
                nv = state_plus<P>(yh[k][f], yl[k][f], kk);
              } else {
                kacc[k][f] = kacc[k][f] + kk;
                state_add<P>(yh[k][f], yl[k][f], kacc[k][f] * C(1.0 / 6.0));
                nv = yh[k][f];
              }
            }
            if (edge) {
              const double dv = dirichlet_at(bcg + f * fstride, n0, n1, gi, gj, ax0);
              if (dv == dv) {
                if (last || alg) {
                  state_set<P>(yh[k][f], yl[k][f], dv);
                  nv = yh[k][f];
                } else {
                  nv = static_cast<C>(dv);
                }
              }
            }
            L[f * CELLS + idx] = nv;
          }
        }
      }
      __syncthreads();
    }
  }

  // ---- write the output tile ----------------------------------------------
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int idx = threadIdx.x + k * NT;
    const int r = idx / EX, c = idx - (idx / EX) * EX;
    const int gi = bi + r, gj = bj + c;
    if (idx < CELLS && r >= H && r < H + OY && c >= H && c < H + OX && gi < n0 && gj < n1) {
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const int64_t e = G.off + f * plane + static_cast<int64_t>(gi) * n1 + gj;
        Store<P>::store(a.out_hi, a.out_lo, e, yh[k][f], yl[k][f]);
        if (a.snap_hi) Store<P>::store(a.snap_hi, a.snap_lo, e, yh[k][f], yl[k][f]);
      }
    }
  }
}

// ===========================================================================
// Kernel variant tables.

using KernelFn = void (*)(FdmArgs);

enum VariantKind { kTileKind = 0, kStreamKind = 1 };

struct Variant {
  KernelFn fn;              // `steps` steps per launch
  KernelFn fn_single;       // one step per launch, same unit shape
  KernelFn fn_snap;         // fn for launches that write a snapshot
  KernelFn fn_single_snap;  // fn_single for launches that write a snapshot
  int kind;                 // kTileKind or kStreamKind
  int ox, oy;               // unit shape: output tile / strip width x rows
  int steps, nt;
  int cpl;                  // columns per lane (stream kind)
  const char *name;
};

#define PML_TILE(EQ, IN, PR, OX, OY, ST, NT, NAME)                                       \
  Variant {                                                                               \
    &fdm_fused<EQ, IN, PR, OX, OY, ST, NT>, &fdm_fused<EQ, IN, PR, OX, OY, 1, NT>,        \
        &fdm_fused<EQ, IN, PR, OX, OY, ST, NT>, &fdm_fused<EQ, IN, PR, OX, OY, 1, NT>,    \
        kTileKind, OX, OY, ST, NT, 1, NAME                                                \
  }
#define PML_STREAM(EQ, IN, PR, RW, PD, FULL, CPL, NAME)                                   \
  Variant {                                                                               \
    &fdm_stream<EQ, IN, PR, RW, PD, FULL, CPL, false>,                                    \
        &fdm_stream<EQ, IN, PR, RW, PD, FULL, CPL, false>,                                \
        &fdm_stream<EQ, IN, PR, RW, PD, FULL, CPL, true>,                                 \
        &fdm_stream<EQ, IN, PR, RW, PD, FULL, CPL, true>, kStreamKind, 64 * CPL, RW, 1,   \
        256, CPL, NAME                                                                    \
  }
#define PML_STREAMRK(EQ, IN, PR, RW, PD, NAME)                                            \
  Variant {                                                                               \
    &fdm_stream_rk<EQ, IN, PR, RW, PD, false>, &fdm_stream_rk<EQ, IN, PR, RW, PD, false>, \
        &fdm_stream_rk<EQ, IN, PR, RW, PD, true>, &fdm_stream_rk<EQ, IN, PR, RW, PD, true>, \
        kStreamKind, 64 - 2 * stream_rk_lead<EQ, IN>(), RW, 1, 256, 1, NAME               \
  }
#define PML_STREAMN(EQ, IN, PR, RW, PD, NS, NAME)                                         \
  Variant {                                                                               \
    &fdm_stream_n<EQ, IN, PR, RW, PD, NS, false>,                                         \
        &fdm_stream_n<EQ, IN, PR, RW, PD, 1, false>,                                      \
        &fdm_stream_n<EQ, IN, PR, RW, PD, NS, true>,                                      \
        &fdm_stream_n<EQ, IN, PR, RW, PD, 1, true>, kStreamKind, 66 - 2 * NS, RW, NS, 256, 1, \
        NAME                                                                              \
  }
#define PML_STREAM2(EQ, IN, PR, RW, PD, NAME)                                             \
  Variant {                                                                               \
    &fdm_stream2<EQ, IN, PR, RW, PD, 2, false>, &fdm_stream2<EQ, IN, PR, RW, PD, 1, false>, \
        &fdm_stream2<EQ, IN, PR, RW, PD, 2, true>,                                        \
        &fdm_stream2<EQ, IN, PR, RW, PD, 1, true>, kStreamKind, 62, RW, 2, 256, 1, NAME   \
  }

// Variants per (equation, integrator, precision); the first is the default.
template <class EQ, class IN, class PR>
const std::vector<Variant> &variants_for() {
  static const std::vector<Variant> v = [] {
    std::vector<Variant> out;
    constexpr int S = EQ::kInvariant ? 1 : IN::kStages;
    constexpr bool wide = sizeof(typename PR::C) == 8;
    if constexpr (S == 1) {
      out.push_back(PML_STREAM(EQ, IN, PR, 8, 1, true, 1, "stream-rw8-pd1-full"));
      out.push_back(PML_STREAM(EQ, IN, PR, 6, 3, true, 1, "stream-rw6-pd3-full"));
      out.push_back(PML_STREAM(EQ, IN, PR, 4, 1, true, 2, "stream-c2-rw4-pd1-full"));
      out.push_back(PML_STREAM2(EQ, IN, PR, 8, 1, "stream2-rw8-pd1-s2"));
      out.push_back(PML_STREAM2(EQ, IN, PR, 8, 2, "stream2-rw8-pd2-s2"));
      out.push_back(PML_STREAMN(EQ, IN, PR, 8, 2, 3, "streamn-rw8-pd2-s3"));
      out.push_back(PML_STREAMN(EQ, IN, PR, 12, 2, 3, "streamn-rw12-pd2-s3"));
      // short strips: more wavefronts and shorter row pipelines for single
      // grids too small to fill the chip (Parareal coarse solves)
      out.push_back(PML_STREAMN(EQ, IN, PR, 4, 1, 3, "streamn-rw4-pd1-s3"));
      out.push_back(PML_STREAMN(EQ, IN, PR, 2, 1, 3, "streamn-rw2-pd1-s3"));
      if constexpr (wide) {
        out.push_back(PML_TILE(EQ, IN, PR, 32, 8, 1, 256, "tile-ox32-oy8-s1"));
      } else {
        out.push_back(PML_TILE(EQ, IN, PR, 64, 16, 1, 256, "tile-ox64-oy16-s1"));
        if constexpr (EQ::F <= 2) {
          out.push_back(PML_TILE(EQ, IN, PR, 64, 16, 2, 256, "tile-ox64-oy16-s2"));
          out.push_back(PML_TILE(EQ, IN, PR, 64, 32, 4, 512, "tile-ox64-oy32-s4"));
        }
      }
    } else if constexpr (EQ::kAlg != 0u) {
      // multi-stage steps with algebraic fields: the LDS tile kernel only
      if constexpr (wide) {
        out.push_back(PML_TILE(EQ, IN, PR, 32, 8, 1, 256, "tile-ox32-oy8-s1"));
      } else {
        out.push_back(PML_TILE(EQ, IN, PR, 64 - 2 * S, 16, 1, 256, "tile-ox64m-oy16-s1"));
      }
    } else {
      out.push_back(PML_STREAMRK(EQ, IN, PR, 8, 1, "streamrk-rw8-pd1"));
      out.push_back(PML_STREAMRK(EQ, IN, PR, 16, 1, "streamrk-rw16-pd1"));
      out.push_back(PML_STREAMRK(EQ, IN, PR, 8, 2, "streamrk-rw8-pd2"));
      out.push_back(PML_STREAMRK(EQ, IN, PR, 4, 1, "streamrk-rw4-pd1"));
      out.push_back(PML_STREAMRK(EQ, IN, PR, 2, 1, "streamrk-rw2-pd1"));
      if constexpr (EQ::F <= 2) {
        // longer strips with a deeper prefetch: the Parareal fine batch of
        // compensated FP32 diffusion is latency-bound at pd1 (7 % slower)
        out.push_back(PML_STREAMRK(EQ, IN, PR, 16, 4, "streamrk-rw16-pd4"));
        out.push_back(PML_STREAMRK(EQ, IN, PR, 24, 4, "streamrk-rw24-pd4"));
      }
      if constexpr (wide) {
        out.push_back(PML_TILE(EQ, IN, PR, 32, 8, 1, 256, "tile-ox32-oy8-s1"));
      } else {
        out.push_back(PML_TILE(EQ, IN, PR, 64 - 2 * S, 16, 1, 256, "tile-ox64m-oy16-s1"));
        out.push_back(PML_TILE(EQ, IN, PR, 64 - 2 * S, 32, 1, 256, "tile-ox64m-oy32-s1"));
      }
    }
    return out;
  }();
  return v;
}

template <class EQ, class PR>
const std::vector<Variant> *variants_by_int(int integ) {
  switch (integ) {
    case PML_INT_FE: return &variants_for<EQ, IntFE, PR>();
    case PML_INT_MIDPOINT: return &variants_for<EQ, IntMidpoint, PR>();
    case PML_INT_RK4: return &variants_for<EQ, IntRK4, PR>();
    default: return nullptr;
  }
}

// Defined in the per-precision translation units (pml_fdm_<prec>_<group>.hip).
// Each returns the variant table for (eq, integrator, ndim) and the field
// count, or nullptr when that unit does not cover the equation.
#define PML_DECLARE_TABLE(NAME) \
  const std::vector<Variant> *NAME(int eq, int integ, int ndim, int *n_fields);
PML_DECLARE_TABLE(fdm_table_f32_a)
PML_DECLARE_TABLE(fdm_table_f32_b)
PML_DECLARE_TABLE(fdm_table_f64_a)
PML_DECLARE_TABLE(fdm_table_f64_b)
PML_DECLARE_TABLE(fdm_table_f32c_a)
PML_DECLARE_TABLE(fdm_table_f32c_b)
PML_DECLARE_TABLE(fdm_table_f32_c)
PML_DECLARE_TABLE(fdm_table_f64_c)
PML_DECLARE_TABLE(fdm_table_f32c_c)
PML_DECLARE_TABLE(fdm_table_f32_d)
PML_DECLARE_TABLE(fdm_table_f64_d)
PML_DECLARE_TABLE(fdm_table_f32c_d)
PML_DECLARE_TABLE(fdm_table_f32_e)
PML_DECLARE_TABLE(fdm_table_f64_e)
PML_DECLARE_TABLE(fdm_table_f32c_e)
#undef PML_DECLARE_TABLE

// Group a: diffusion.  Group d: Cahn-Hilliard, convection-diffusion.
// Group b: shallow water, wave.  Group e: Burgers.  Group c: Navier-Stokes
// (2-D).  Five groups per precision keep the translation units small enough
// to compile in parallel.
#define PML_DEFINE_TABLE_A(NAME, PR)                                                   \
  const std::vector<Variant> *NAME(int eq, int integ, int ndim, int *n_fields) {       \
    (void)ndim;                                                                        \
    if (eq != PML_EQ_DIFFUSION) return nullptr;                                        \
    *n_fields = 1;                                                                     \
    return variants_by_int<EqDiffusion, PR>(integ);                                    \
  }
#define PML_DEFINE_TABLE_D(NAME, PR)                                                   \
  const std::vector<Variant> *NAME(int eq, int integ, int ndim, int *n_fields) {       \
    (void)ndim;                                                                        \
    switch (eq) {                                                                      \
      case PML_EQ_CAHN_HILLIARD:                                                       \
        *n_fields = 2;                                                                 \
        return variants_by_int<EqCahnHilliard, PR>(integ);                             \
      case PML_EQ_CONVECTION_DIFFUSION:                                                \
        *n_fields = 1;                                                                 \
        return variants_by_int<EqConvectionDiffusion, PR>(integ);                      \
      default: return nullptr;                                                         \
    }                                                                                  \
  }
#define PML_DEFINE_TABLE_B(NAME, PR)                                                   \
  const std::vector<Variant> *NAME(int eq, int integ, int ndim, int *n_fields) {       \
    switch (eq) {                                                                      \
      case PML_EQ_SHALLOW_WATER:                                                       \
        *n_fields = 3;                                                                 \
        return ndim == 2 ? variants_by_int<EqShallowWater, PR>(integ) : nullptr;       \
      case PML_EQ_WAVE: *n_fields = 2; return variants_by_int<EqWave, PR>(integ);      \
      default: return nullptr;                                                         \
    }                                                                                  \
  }
#define PML_DEFINE_TABLE_E(NAME, PR)                                                   \
  const std::vector<Variant> *NAME(int eq, int integ, int ndim, int *n_fields) {       \
    if (eq != PML_EQ_BURGERS) return nullptr;                                          \
    if (ndim == 1) {                                                                   \
      *n_fields = 1;                                                                   \
      return variants_by_int<EqBurgers1, PR>(integ);                                   \
    }                                                                                  \
    *n_fields = 2;                                                                     \
    return variants_by_int<EqBurgers2, PR>(integ);                                     \
  }
#define PML_DEFINE_TABLE_C(NAME, PR)                                                   \
  const std::vector<Variant> *NAME(int eq, int integ, int ndim, int *n_fields) {       \
    if (eq != PML_EQ_NAVIER_STOKES || ndim != 2) return nullptr;                       \
    *n_fields = 4;                                                                     \
    return variants_by_int<EqNavierStokes, PR>(integ);                                 \
  }

}  // namespace pml
