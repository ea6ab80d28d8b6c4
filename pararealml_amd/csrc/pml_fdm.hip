// Fused finite-difference time stepping for gfx950.
//
// One workgroup advances one output tile of one grid of a (possibly ragged)
// batch by STEPS time steps.  It loads the tile plus a halo of
// H = stages_per_step * STEPS vertices into LDS once, evaluates every
// Runge-Kutta stage on a region that shrinks by one vertex per stage, and
// writes the output tile once: HBM sees one read and one write of the state
// per launch, whatever the integrator and the number of fused steps.
//
// Boundary semantics restate pararealml/operators/fdm/numerical_differentiator.py:
//   * first derivatives use zero halos and are then overwritten on the
//     boundary by the Neumann values (1012-1057);
//   * second derivatives use ghosts y[1] - 2dx g_lo / y[n-2] + 2dx g_hi where
//     a Neumann value exists and 0 otherwise (1059-1095, 1188-1242);
// and the Dirichlet constraints are re-imposed on every stage input and on
// the step result (numerical_integrator.py:42-132).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
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

// Bijective block -> tile map that gives every XCD (blocks are dealt to the
// 8 XCDs round robin) a contiguous range of tiles, so that tiles sharing
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
  __device__ static void load(const FdmArgs &a, int64_t i, double &hi, double &lo) {
    hi = static_cast<const double *>(a.in_hi)[i];
    lo = 0.0;
  }
  __device__ static void store(void *p_hi, void *, int64_t i, double hi, double) {
    static_cast<double *>(p_hi)[i] = hi;
  }
};
template <>
struct Store<PrecF32> {
  __device__ static void load(const FdmArgs &a, int64_t i, float &hi, float &lo) {
    hi = static_cast<const float *>(a.in_hi)[i];
    lo = 0.0f;
  }
  __device__ static void store(void *p_hi, void *, int64_t i, float hi, float) {
    static_cast<float *>(p_hi)[i] = hi;
  }
};
template <>
struct Store<PrecF32C> {
  __device__ static void load(const FdmArgs &a, int64_t i, float &hi, float &lo) {
    hi = static_cast<const float *>(a.in_hi)[i];
    lo = static_cast<const float *>(a.in_lo)[i];
  }
  __device__ static void store(void *p_hi, void *p_lo, int64_t i, float hi, float lo) {
    static_cast<float *>(p_hi)[i] = hi;
    static_cast<float *>(p_lo)[i] = lo;
  }
};

// y <- y + d for the state precision; returns the rounded working value.
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
  constexpr unsigned kNeed = Eq::kLap | Eq::kD;

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

  C pc[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) pc[i] = static_cast<C>(a.p[i]);
  const C dt = static_cast<C>(a.dt);
  const C hdt = static_cast<C>(a.dt * 0.5);
  const C inv_dx0sq = static_cast<C>(G.inv_dx0sq), inv_dx1sq = static_cast<C>(G.inv_dx1sq);
  const C inv_2dx0 = static_cast<C>(G.inv_2dx0), inv_2dx1 = static_cast<C>(G.inv_2dx1);
  const C m2dx0 = static_cast<C>(G.m2dx0), m2dx1 = static_cast<C>(G.m2dx1);
  const C p2dx0 = static_cast<C>(G.p2dx0), p2dx1 = static_cast<C>(G.p2dx1);

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
        if (in) Store<P>::load(a, G.off + f * plane + static_cast<int64_t>(gi) * n1 + gj, hi, lo);
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
          Quant<C> q[F];
          const bool lo1 = gj == 0, hi1 = gj == n1 - 1;
          const bool lo0 = ax0 && gi == 0, hi0 = ax0 && gi == n0 - 1;
          const bool edge = lo1 || hi1 || lo0 || hi0;
#pragma unroll
          for (int f = 0; f < F; ++f) {
            const C *Lf = L + f * CELLS;
            const C yc = Lf[idx];
            q[f].v = yc;
            q[f].lap = 0;
            q[f].d0 = 0;
            q[f].d1 = 0;
            if (!((kNeed >> f) & 1u)) continue;
            const bool wl = (Eq::kLap >> f) & 1u, wd = (Eq::kD >> f) & 1u;
            // axis 1 (contiguous)
            C yw = lo1 ? C(0) : Lf[idx - 1];
            C ye = hi1 ? C(0) : Lf[idx + 1];
            C lap1 = 0, d1 = 0, lap0 = 0, d0 = 0;
            if (!edge) {
              if (wl) lap1 = ((ye - C(2) * yc) + yw) * inv_dx1sq;
              if (wd) d1 = (ye - yw) * inv_2dx1;
              if (ax0) {
                const C yn = Lf[idx - EX], ys = Lf[idx + EX];
                if (wl) lap0 = ((ys - C(2) * yc) + yn) * inv_dx0sq;
                if (wd) d0 = (ys - yn) * inv_2dx0;
              }
            } else {
              const double *nb = bcg + f * fstride + 2 * (n0 + n1);
              double g_lo1 = lo1 ? nb[2 * n1 + gi] : NAN;
              double g_hi1 = hi1 ? nb[2 * n1 + n0 + gi] : NAN;
              if (wl) {
                C gw = yw, ge = ye;
                if (lo1) gw = (g_lo1 == g_lo1) ? Lf[idx + 1] + m2dx1 * static_cast<C>(g_lo1) : C(0);
                if (hi1) ge = (g_hi1 == g_hi1) ? Lf[idx - 1] + p2dx1 * static_cast<C>(g_hi1) : C(0);
                lap1 = ((ge - C(2) * yc) + gw) * inv_dx1sq;
              }
              if (wd) {
                d1 = (ye - yw) * inv_2dx1;
                if (lo1 && g_lo1 == g_lo1) d1 = static_cast<C>(g_lo1);
                if (hi1 && g_hi1 == g_hi1) d1 = static_cast<C>(g_hi1);
              }
              if (ax0) {
                const C yn = lo0 ? C(0) : Lf[idx - EX];
                const C ys = hi0 ? C(0) : Lf[idx + EX];
                double g_lo0 = lo0 ? nb[gj] : NAN;
                double g_hi0 = hi0 ? nb[n1 + gj] : NAN;
                if (wl) {
                  C gn = yn, gs = ys;
                  if (lo0) gn = (g_lo0 == g_lo0) ? Lf[idx + EX] + m2dx0 * static_cast<C>(g_lo0) : C(0);
                  if (hi0) gs = (g_hi0 == g_hi0) ? Lf[idx - EX] + p2dx0 * static_cast<C>(g_hi0) : C(0);
                  lap0 = ((gs - C(2) * yc) + gn) * inv_dx0sq;
                }
                if (wd) {
                  d0 = (ys - yn) * inv_2dx0;
                  if (lo0 && g_lo0 == g_lo0) d0 = static_cast<C>(g_lo0);
                  if (hi0 && g_hi0 == g_hi0) d0 = static_cast<C>(g_hi0);
                }
              }
            }
            if (ax0) {
              q[f].lap = lap0 + lap1;
              q[f].d0 = d0;
              q[f].d1 = d1;
            } else {
              q[f].lap = lap1;
              q[f].d0 = d1;  // the single physical axis lies along axis 1
              q[f].d1 = 0;
            }
          }
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
              if (Integ::kStages == 4) d = (((kk + C(2) * kk) + C(2) * kk) + kk) / C(6);
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
                nv = state_plus<P>(yh[k][f], yl[k][f], kk / C(2));
              } else if (stage == 1) {
                kacc[k][f] = kacc[k][f] + C(2) * kk;
                nv = state_plus<P>(yh[k][f], yl[k][f], kk / C(2));
              } else if (stage == 2) {
                kacc[k][f] = kacc[k][f] + C(2) * kk;
                nv = state_plus<P>(yh[k][f], yl[k][f], kk);
              } else {
                kacc[k][f] = kacc[k][f] + kk;
                state_add<P>(yh[k][f], yl[k][f], kacc[k][f] / C(6));
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
// Host side: variant tables, tile tables and the step loop.

using KernelFn = void (*)(FdmArgs);

struct Variant {
  KernelFn fn;
  KernelFn fn_single;  // STEPS == 1 variant with the same tile shape
  int ox, oy, steps, nt;
  const char *name;
};

#define PML_VARIANT(EQ, IN, PR, OX, OY, ST, NT, NAME)                                 \
  Variant {                                                                            \
    &fdm_fused<EQ, IN, PR, OX, OY, ST, NT>, &fdm_fused<EQ, IN, PR, OX, OY, 1, NT>, OX, \
        OY, ST, NT, NAME                                                               \
  }

template <class EQ, class IN, class PR>
const std::vector<Variant> &variants_for() {
  // Tile shapes: 64-wide rows (one wavefront per row), heights and fused
  // step counts chosen per stage count.  The first entry is the default.
  static const std::vector<Variant> v = [] {
    std::vector<Variant> out;
    constexpr int S = EQ::kInvariant ? 1 : IN::kStages;
    if constexpr (sizeof(typename PR::C) == 8) {
      out.push_back(PML_VARIANT(EQ, IN, PR, 32, 8, 1, 256, "ox32-oy8-s1"));
    } else if constexpr (S == 1) {
      out.push_back(PML_VARIANT(EQ, IN, PR, 64, 16, 1, 256, "ox64-oy16-s1"));
      out.push_back(PML_VARIANT(EQ, IN, PR, 64, 32, 1, 256, "ox64-oy32-s1"));
      out.push_back(PML_VARIANT(EQ, IN, PR, 64, 16, 2, 256, "ox64-oy16-s2"));
      out.push_back(PML_VARIANT(EQ, IN, PR, 64, 32, 4, 512, "ox64-oy32-s4"));
    } else {
      out.push_back(PML_VARIANT(EQ, IN, PR, 64 - 2 * S, 16, 1, 256, "ox64m-oy16-s1"));
      out.push_back(PML_VARIANT(EQ, IN, PR, 64 - 2 * S, 32, 1, 256, "ox64m-oy32-s1"));
    }
    return out;
  }();
  return v;
}

template <class EQ, class IN>
const std::vector<Variant> *variants_by_prec(int prec) {
  switch (prec) {
    case PML_F32: return &variants_for<EQ, IN, PrecF32>();
    case PML_F64: return &variants_for<EQ, IN, PrecF64>();
    case PML_F32C: return &variants_for<EQ, IN, PrecF32C>();
    default: return nullptr;
  }
}

template <class EQ>
const std::vector<Variant> *variants_by_int(int integ, int prec) {
  switch (integ) {
    case PML_INT_FE: return variants_by_prec<EQ, IntFE>(prec);
    case PML_INT_MIDPOINT: return variants_by_prec<EQ, IntMidpoint>(prec);
    case PML_INT_RK4: return variants_by_prec<EQ, IntRK4>(prec);
    default: return nullptr;
  }
}

static const std::vector<Variant> *find_variants(int eq, int integ, int prec, int ndim,
                                                 int *n_fields) {
  switch (eq) {
    case PML_EQ_DIFFUSION: *n_fields = 1; return variants_by_int<EqDiffusion>(integ, prec);
    case PML_EQ_CAHN_HILLIARD: *n_fields = 2; return variants_by_int<EqCahnHilliard>(integ, prec);
    case PML_EQ_SHALLOW_WATER:
      *n_fields = 3;
      return ndim == 2 ? variants_by_int<EqShallowWater>(integ, prec) : nullptr;
    case PML_EQ_CONVECTION_DIFFUSION:
      *n_fields = 1;
      return variants_by_int<EqConvectionDiffusion>(integ, prec);
    case PML_EQ_WAVE: *n_fields = 2; return variants_by_int<EqWave>(integ, prec);
    case PML_EQ_BURGERS:
      if (ndim == 1) {
        *n_fields = 1;
        return variants_by_int<EqBurgers1>(integ, prec);
      }
      *n_fields = 2;
      return variants_by_int<EqBurgers2>(integ, prec);
    default: return nullptr;
  }
}

// Device-side copies of the tile and grid tables, reused across calls.
struct DeviceTables {
  std::mutex mu;
  void *tiles = nullptr;
  size_t tiles_cap = 0;
  void *grids = nullptr;
  size_t grids_cap = 0;
};
static DeviceTables g_tables;

static int ensure_capacity(void **p, size_t *cap, size_t bytes) {
  if (*cap >= bytes) return PML_OK;
  if (*p) PML_HIP_CHECK(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  size_t want = std::max<size_t>(bytes, 4096);
  PML_HIP_CHECK(hipMalloc(p, want));
  *cap = want;
  return PML_OK;
}

}  // namespace pml

using namespace pml;

extern "C" int pml_fdm_configs(int eq_id, int integrator_id, int precision) {
  int nf = 0;
  const auto *v = find_variants(eq_id, integrator_id, precision, 2, &nf);
  return v ? static_cast<int>(v->size()) : 0;
}

extern "C" const char *pml_fdm_config_name(int eq_id, int integrator_id, int precision,
                                           int config) {
  int nf = 0;
  const auto *v = find_variants(eq_id, integrator_id, precision, 2, &nf);
  if (!v || config < 0 || config >= static_cast<int>(v->size())) return "";
  return (*v)[config].name;
}

extern "C" int pml_fdm_solve(int eq_id, const double *eq_params, int n_params,
                             int integrator_id, int precision, int ndim, const pml_grid *grids,
                             int n_grids, const double *bc_values, void *state_hi,
                             void *state_lo, void *work_hi, void *work_lo, int64_t total_elems,
                             double dt, int64_t n_steps, void *snapshots_hi,
                             void *snapshots_lo, int64_t snapshot_every, int tile_config,
                             int use_graph, void *stream) {
  int n_fields = 0;
  if (ndim != 1 && ndim != 2) return fail(PML_E_INVALID, "ndim must be 1 or 2");
  const auto *vars = find_variants(eq_id, integrator_id, precision, ndim, &n_fields);
  if (!vars) return fail(PML_E_UNSUPPORTED, "no device kernel for this equation/integrator/precision");
  if (tile_config < 0 || tile_config >= static_cast<int>(vars->size()))
    return fail(PML_E_INVALID, "tile_config out of range");
  if (n_params < 0 || n_params > 6) return fail(PML_E_INVALID, "at most 6 equation parameters");
  if (n_grids <= 0 || !grids) return fail(PML_E_INVALID, "no grids");
  if (!state_hi || !work_hi || !bc_values) return fail(PML_E_INVALID, "null buffer");
  if (precision == PML_F32C && (!state_lo || !work_lo))
    return fail(PML_E_INVALID, "compensated precision needs lo planes");
  if (n_steps < 0) return fail(PML_E_INVALID, "n_steps must be non-negative");
  if (snapshots_hi && snapshot_every <= 0)
    return fail(PML_E_INVALID, "snapshot_every must be positive");
  if (snapshots_hi && precision == PML_F32C && !snapshots_lo)
    return fail(PML_E_INVALID, "compensated snapshots need a lo plane");

  const Variant &V = (*vars)[tile_config];
  // Fused steps per launch must divide the snapshot interval.
  const bool can_fuse = V.steps > 1 && (!snapshots_hi || snapshot_every % V.steps == 0);
  const int steps_per_launch = can_fuse ? V.steps : 1;

  // Grid and tile tables.
  std::vector<GridDev> gd(n_grids);
  std::vector<TileDev> td;
  for (int g = 0; g < n_grids; ++g) {
    const pml_grid &s = grids[g];
    if (s.n1 < 3 || (ndim == 2 && s.n0 < 3) || (ndim == 1 && s.n0 != 1))
      return fail(PML_E_INVALID, "each active axis needs at least 3 vertices");
    if (!(s.dx0 > 0) || !(s.dx1 > 0)) return fail(PML_E_INVALID, "spacing must be positive");
    const int64_t need = s.state_offset + static_cast<int64_t>(n_fields) * s.n0 * s.n1;
    if (s.state_offset < 0 || need > total_elems)
      return fail(PML_E_INVALID, "grid state exceeds total_elems");
    GridDev &d = gd[g];
    d.n0 = s.n0;
    d.n1 = s.n1;
    d.off = s.state_offset;
    d.bc = s.bc_offset;
    d.inv_dx0sq = 1.0 / (s.dx0 * s.dx0);
    d.inv_dx1sq = 1.0 / (s.dx1 * s.dx1);
    d.inv_2dx0 = 1.0 / (2.0 * s.dx0);
    d.inv_2dx1 = 1.0 / (2.0 * s.dx1);
    d.m2dx0 = -2.0 * s.dx0;
    d.m2dx1 = -2.0 * s.dx1;
    d.p2dx0 = 2.0 * s.dx0;
    d.p2dx1 = 2.0 * s.dx1;
    for (int i0 = 0; i0 < s.n0; i0 += V.oy)
      for (int j0 = 0; j0 < s.n1; j0 += V.ox) td.push_back(TileDev{g, i0, j0, 0});
  }
  hipStream_t st = as_stream(stream);
  std::lock_guard<std::mutex> lock(g_tables.mu);
  int rc = ensure_capacity(&g_tables.tiles, &g_tables.tiles_cap, td.size() * sizeof(TileDev));
  if (rc) return rc;
  rc = ensure_capacity(&g_tables.grids, &g_tables.grids_cap, gd.size() * sizeof(GridDev));
  if (rc) return rc;
  PML_HIP_CHECK(hipMemcpyAsync(g_tables.tiles, td.data(), td.size() * sizeof(TileDev),
                               hipMemcpyHostToDevice, st));
  PML_HIP_CHECK(hipMemcpyAsync(g_tables.grids, gd.data(), gd.size() * sizeof(GridDev),
                               hipMemcpyHostToDevice, st));

  FdmArgs a{};
  a.tiles = static_cast<const TileDev *>(g_tables.tiles);
  a.grids = static_cast<const GridDev *>(g_tables.grids);
  a.bc = bc_values;
  a.dt = dt;
  for (int i = 0; i < 6; ++i) a.p[i] = i < n_params ? eq_params[i] : 0.0;
  a.n_tiles = static_cast<int>(td.size());
  a.ndim = ndim;
  const dim3 grid_dim(a.n_tiles), block_dim(V.nt);
  const size_t esz = precision == PML_F64 ? 8 : 4;

  void *cur_hi = state_hi, *cur_lo = state_lo, *nxt_hi = work_hi, *nxt_lo = work_lo;
  int64_t done = 0, snap_idx = 0;

  auto launch = [&](KernelFn fn, int steps) -> int {
    a.in_hi = cur_hi;
    a.in_lo = cur_lo;
    a.out_hi = nxt_hi;
    a.out_lo = nxt_lo;
    a.snap_hi = nullptr;
    a.snap_lo = nullptr;
    if (snapshots_hi && (done + steps) % snapshot_every == 0) {
      a.snap_hi = static_cast<char *>(snapshots_hi) + snap_idx * total_elems * esz;
      a.snap_lo = snapshots_lo ? static_cast<char *>(snapshots_lo) + snap_idx * total_elems * esz
                               : nullptr;
      ++snap_idx;
    }
    hipLaunchKernelGGL(fn, grid_dim, block_dim, 0, st, a);
    PML_LAUNCH_CHECK();
    std::swap(cur_hi, nxt_hi);
    std::swap(cur_lo, nxt_lo);
    done += steps;
    return PML_OK;
  };

  // Launch loop.  With graphs, a fixed block of launches is captured once
  // and replayed; the block length is even so the ping-pong parity is the
  // same at every replay, and no snapshot may fall inside a replayed block.
  const int64_t fused_launches = n_steps / steps_per_launch;
  int64_t li = 0;
  // Stream capture needs a created stream (not the null stream).
  if (use_graph && st != nullptr && !snapshots_hi && fused_launches >= 64) {
    const int64_t block = 32;  // launches per graph (even)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    PML_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int64_t i = 0; i < block; ++i) {
      rc = launch(can_fuse ? V.fn : V.fn_single, steps_per_launch);
      if (rc) {
        hipStreamEndCapture(st, &graph);
        if (graph) hipGraphDestroy(graph);
        return rc;
      }
    }
    PML_HIP_CHECK(hipStreamEndCapture(st, &graph));
    PML_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    // The capture itself recorded the first block without running it.
    const int64_t reps = fused_launches / block;
    for (int64_t r = 0; r < reps; ++r) {
      hipError_t e = hipGraphLaunch(exec, st);
      if (e != hipSuccess) {
        hipGraphExecDestroy(exec);
        hipGraphDestroy(graph);
        return fail(PML_E_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
      }
    }
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
    // After an even number of launches per replay the buffers are back in
    // place; account for the captured steps.
    done = reps * block * steps_per_launch;
    li = reps * block;
  }
  for (; li < fused_launches; ++li) {
    rc = launch(can_fuse ? V.fn : V.fn_single, steps_per_launch);
    if (rc) return rc;
  }
  while (done < n_steps) {
    rc = launch(V.fn_single, 1);
    if (rc) return rc;
  }
  if (cur_hi != state_hi) {
    PML_HIP_CHECK(hipMemcpyAsync(state_hi, cur_hi, total_elems * esz, hipMemcpyDeviceToDevice, st));
    if (precision == PML_F32C)
      PML_HIP_CHECK(
          hipMemcpyAsync(state_lo, cur_lo, total_elems * esz, hipMemcpyDeviceToDevice, st));
  }
  return PML_OK;
}
