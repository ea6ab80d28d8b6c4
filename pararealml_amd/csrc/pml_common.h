// Shared definitions of the gfx950 kernels: status handling, precision
// policies and compensated (hi + lo) FP32 arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/pml.h"

namespace pml {

// Thread-local error message returned by pml_last_error().
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define PML_HIP_CHECK(expr)                                                   \
  do {                                                                        \
    hipError_t err__ = (expr);                                                \
    if (err__ != hipSuccess) {                                                \
      return ::pml::fail(PML_E_HIP, std::string(#expr) + ": " +               \
                                        hipGetErrorString(err__));            \
    }                                                                         \
  } while (0)

#define PML_LAUNCH_CHECK()                                                    \
  do {                                                                        \
    hipError_t err__ = hipGetLastError();                                     \
    if (err__ != hipSuccess) {                                                \
      return ::pml::fail(PML_E_HIP, std::string("kernel launch: ") +          \
                                        hipGetErrorString(err__));            \
    }                                                                         \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------
// Precision policies.  C is the arithmetic type; kComp marks the compensated
// FP32 state (separate hi and lo planes, value = hi + lo).
struct PrecF64 {
  using C = double;
  static constexpr bool kComp = false;
};
struct PrecF32 {
  using C = float;
  static constexpr bool kComp = false;
};
struct PrecF32C {
  using C = float;
  static constexpr bool kComp = true;
};

// Error-free transformation a + b = s + e (Knuth's two-sum); valid for any
// magnitudes.  Compiled without fast-math so the operations stay as written.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// (hi, lo) += d with renormalisation; keeps |lo| <= ulp(hi)/2.
__device__ __forceinline__ void comp_add(float &hi, float &lo, float d) {
  float s, e;
  two_sum(hi, d, s, e);
  e += lo;
  float h = s + e;
  lo = e - (h - s);
  hi = h;
}

// Exact split of a double into a float pair (hi = nearest float).
__device__ __forceinline__ void split_f64(double v, float &hi, float &lo) {
  hi = static_cast<float>(v);
  lo = static_cast<float>(v - static_cast<double>(hi));
}

}  // namespace pml
