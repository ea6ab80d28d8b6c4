// Fused right-hand sides of the equations that have device kernels.
//
// Each equation describes, per field, which stencil quantities it needs
// (kLap: Laplacian, kD: first derivatives, as bit masks over fields), which
// fields are algebraic (LHS.Y: recomputed from the pre-step state after the
// integrator step, pararealml/operators/fdm/fdm_operator.py:127-139) and
// whether its time-derivative right-hand sides are independent of the
// time-derivative fields themselves (kInvariant).  For an invariant system
// every Runge-Kutta stage sees the same right-hand side -- the stage inputs
// differ only in fields the right-hand side does not read -- so the stages
// are evaluated once and combined with the integrator's own weights.
//
// The expressions restate pararealml/differential_equation.py:608-819.
#pragma once

namespace pml {

// x * s + y for a power of two s (so x * s is exact): one fused
// multiply-add with the same result as the separate multiply and add, in
// either precision.
template <typename C>
__device__ __forceinline__ C fma_pow2(C s, C x, C y) {
  if constexpr (sizeof(C) == 4) return __builtin_fmaf(s, x, y);
  else return __builtin_fma(s, x, y);
}

// acc + x * y: one fused multiply-add in FP32, where the multi-field
// kernels are bound by VALU issue; in FP64 the separate multiply and add of
// the reference's operation order (bitwise the same as acc - (-x) * y).
template <typename C>
__device__ __forceinline__ C mad(C x, C y, C acc) {
  if constexpr (sizeof(C) == 4) return __builtin_fmaf(x, y, acc);
  else return acc + x * y;
}

template <typename C>
struct Quant {
  C v;    // value
  C lap;  // Laplacian (sum over active axes, axis order)
  C d0;   // derivative along physical axis 0
  C d1;   // derivative along physical axis 1 (0 for 1-D problems)
};

// u_t = d lap(u)    (differential_equation.py:608-630)
struct EqDiffusion {
  static constexpr int F = 1;
  static constexpr unsigned kLap = 0x1u, kD = 0x0u, kAlg = 0x0u;
  static constexpr bool kInvariant = false;
  template <typename C>
  __device__ __forceinline__ static void rhs(const C *p, const Quant<C> *q, C *r) {
    r[0] = p[0] * q[0].lap;
  }
  template <typename C>
  __device__ __forceinline__ static void alg(const C *, const Quant<C> *, C *) {}
};

// u_t = d lap(u) - (v0 u_x0 + v1 u_x1)    (differential_equation.py:633-670)
struct EqConvectionDiffusion {
  static constexpr int F = 1;
  static constexpr unsigned kLap = 0x1u, kD = 0x1u, kAlg = 0x0u;
  static constexpr bool kInvariant = false;
  template <typename C>
  __device__ __forceinline__ static void rhs(const C *p, const Quant<C> *q, C *r) {
    r[0] = mad(-p[2], q[0].d1, mad(-p[1], q[0].d0, p[0] * q[0].lap));
  }
  template <typename C>
  __device__ __forceinline__ static void alg(const C *, const Quant<C> *, C *) {}
};

// (u, w)_t = (w, c^2 lap(u))    (differential_equation.py:673-697)
struct EqWave {
  static constexpr int F = 2;
  static constexpr unsigned kLap = 0x1u, kD = 0x0u, kAlg = 0x0u;
  static constexpr bool kInvariant = false;
  template <typename C>
  __device__ __forceinline__ static void rhs(const C *p, const Quant<C> *q, C *r) {
    r[0] = q[1].v;
    r[1] = p[0] * q[0].lap;
  }
  template <typename C>
  __device__ __forceinline__ static void alg(const C *, const Quant<C> *, C *) {}
};

// c_t = d lap(mu);  mu = c^3 - c - gamma lap(c)  (LHS.Y)
// (differential_equation.py:700-731).  c_t reads only mu, which is
// algebraic, so the system is stage invariant.
struct EqCahnHilliard {
  static constexpr int F = 2;
  static constexpr unsigned kLap = 0x3u, kD = 0x0u, kAlg = 0x2u;
  static constexpr bool kInvariant = true;
  template <typename C>
  __device__ __forceinline__ static void rhs(const C *p, const Quant<C> *q, C *r) {
    r[0] = p[0] * q[1].lap;
  }
  template <typename C>
  __device__ __forceinline__ static void alg(const C *p, const Quant<C> *q, C *r) {
    const C c = q[0].v;
    r[1] = mad(-p[1], q[0].lap, c * c * c - c);
  }
};

// 1-D Burgers: u_t = lap(u)/re - u u_x    (differential_equation.py:734-765)
struct EqBurgers1 {
  static constexpr int F = 1;
  static constexpr unsigned kLap = 0x1u, kD = 0x1u, kAlg = 0x0u;
  static constexpr bool kInvariant = false;
  template <typename C>
  __device__ __forceinline__ static void rhs(const C *p, const Quant<C> *q, C *r) {
    r[0] = mad(-q[0].v, q[0].d0, p[0] * q[0].lap);
  }
  template <typename C>
  __device__ __forceinline__ static void alg(const C *, const Quant<C> *, C *) {}
};

// 2-D Burgers: u_i,t = lap(u_i)/re - (u u_i,x0 + v u_i,x1)
struct EqBurgers2 {
  static constexpr int F = 2;
  static constexpr unsigned kLap = 0x3u, kD = 0x3u, kAlg = 0x0u;
  static constexpr bool kInvariant = false;
  template <typename C>
  __device__ __forceinline__ static void rhs(const C *p, const Quant<C> *q, C *r) {
    r[0] = mad(p[0], q[0].lap, -mad(q[1].v, q[0].d1, q[0].v * q[0].d0));
    r[1] = mad(p[0], q[1].lap, -mad(q[1].v, q[1].d1, q[0].v * q[1].d0));
  }
  template <typename C>
  __device__ __forceinline__ static void alg(const C *, const Quant<C> *, C *) {}
};

// Shallow water in (eta, u, v), parameters (h, b, nu, f, g)
// (differential_equation.py:768-819).
struct EqShallowWater {
  static constexpr int F = 3;
  static constexpr unsigned kLap = 0x6u, kD = 0x7u, kAlg = 0x0u;
  static constexpr bool kInvariant = false;
  template <typename C>
  __device__ __forceinline__ static void rhs(const C *p, const Quant<C> *q, C *r) {
    const C h = p[0], b = p[1], nu = p[2], f = p[3], g = p[4];
    const C eta = q[0].v, u = q[1].v, v = q[2].v;
    r[0] = mad(-v, q[0].d1,
               mad(-eta, q[2].d1, mad(-u, q[0].d0, mad(-eta, q[1].d0, -h * (q[1].d0 + q[2].d1)))));
    r[1] = mad(f, v, mad(-b, u, mad(-g, q[0].d0, mad(-v, q[1].d1, mad(-u, q[1].d0,
                                                                        nu * q[1].lap)))));
    r[2] = mad(-f, u, mad(-b, v, mad(-g, q[0].d1, mad(-v, q[2].d1, mad(-u, q[2].d0,
                                                                         nu * q[2].lap)))));
  }
  template <typename C>
  __device__ __forceinline__ static void alg(const C *, const Quant<C> *, C *) {}
};

// Navier-Stokes in vorticity / stream function (w, psi, u, v), parameter
// 1/re (differential_equation.py:822-850): w_t = lap(w)/re - (u w_x0 +
// v w_x1); u = psi_x1 and v = -psi_x0 are algebraic (from the pre-step
// state); psi is an LHS.Y_LAPLACIAN field (lap(psi) = -w), re-solved by the
// host after each step, so its right-hand side here is zero and the step
// leaves it unchanged.
struct EqNavierStokes {
  static constexpr int F = 4;
  static constexpr unsigned kLap = 0x1u, kD = 0x3u, kAlg = 0xCu;
  static constexpr bool kInvariant = false;
  template <typename C>
  __device__ __forceinline__ static void rhs(const C *p, const Quant<C> *q, C *r) {
    r[0] = mad(p[0], q[0].lap, -mad(q[3].v, q[0].d1, q[2].v * q[0].d0));
  }
  template <typename C>
  __device__ __forceinline__ static void alg(const C *, const Quant<C> *q, C *r) {
    r[2] = q[1].d1;
    r[3] = -q[1].d0;
  }
};

// Integrators (pararealml/operators/fdm/numerical_integrator.py:42-132).
struct IntFE {
  static constexpr int kStages = 1;
};
struct IntMidpoint {
  static constexpr int kStages = 2;
};
struct IntRK4 {
  static constexpr int kStages = 4;
};

}  // namespace pml
