// Device evaluation of symbolic right-hand sides compiled to a small
// register-machine program (the device form of what the reference's
// SymbolMapper builds with sympy.lambdify, pararealml/operators/
// symbol_mapper.py:160-253), and the batched ODE time loop that runs such a
// program (FDMOperator.solve on a system without a mesh,
// pararealml/operators/fdm/fdm_operator.py:48-77 with the explicit
// integrators of numerical_integrator.py:42-132).
//
// Program: n_instr instructions of four int32 {op, dst, a, b}; registers
// live in LDS, one column per thread (register r of thread i at
// regs[r * blockDim.x + i]), so an instruction's operands are dynamic
// indices without private-memory spills.  Every thread runs the same
// instruction stream, so the code and constants are read through uniform
// (scalar) loads.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "pml_common.h"

namespace pml {

enum ExprOp : int {
  OP_CONST = 0,  // r[dst] = consts[a]
  OP_LOAD = 1,   // r[dst] = input a at this element
  OP_TIME = 2,   // r[dst] = t
  OP_ADD = 3,
  OP_SUB = 4,
  OP_MUL = 5,
  OP_DIV = 6,
  OP_POW = 7,    // pow(r[a], r[b])
  OP_POWI = 8,   // r[a] ** b, b an integer (negative: 1 / r[a] ** -b)
  OP_NEG = 9,
  OP_STORE = 10,  // output dst at this element = r[a]
  OP_SQRT = 11,
  OP_EXP = 12,
  OP_LOG = 13,
  OP_SIN = 14,
  OP_COS = 15,
  OP_TAN = 16,
  OP_TANH = 17,
  OP_ABS = 18,
  OP_ASIN = 19,
  OP_ACOS = 20,
  OP_ATAN = 21,
  OP_SINH = 22,
  OP_COSH = 23,
  OP_ATAN2 = 24,
  OP_MIN = 25,
  OP_MAX = 26,
  OP_SIGN = 27,
  OP_COUNT = 28
};

template <typename T>
__device__ __forceinline__ T powi(T x, int k) {
  const bool neg = k < 0;
  unsigned int e = neg ? static_cast<unsigned int>(-k) : static_cast<unsigned int>(k);
  // left-to-right binary powering: x^2 = x*x exactly as NumPy's square
  T r = T(1);
  bool started = false;
  for (int bit = 31; bit >= 0; --bit) {
    if (started) r = r * r;
    if ((e >> bit) & 1u) {
      r = started ? r * x : x;
      started = true;
    }
  }
  return neg ? T(1) / r : r;
}

// Runs the program for one element: load(k) gives input k, store(k, v)
// receives output k.
template <typename T, class Load, class Store>
__device__ __forceinline__ void run_program(const int4 *__restrict__ code, int n_instr,
                                            const double *__restrict__ consts, T *regs, int rs,
                                            T t, Load load, Store store) {
  for (int pc = 0; pc < n_instr; ++pc) {
    const int4 in = code[pc];
    T *d = regs + in.y * rs;
    // every operand index was range-checked when the program was created
    const T x = in.x >= OP_ADD ? regs[in.z * rs] : T(0);
    switch (in.x) {
      case OP_CONST: *d = static_cast<T>(consts[in.z]); break;
      case OP_LOAD: *d = load(in.z); break;
      case OP_TIME: *d = t; break;
      case OP_ADD: *d = x + regs[in.w * rs]; break;
      case OP_SUB: *d = x - regs[in.w * rs]; break;
      case OP_MUL: *d = x * regs[in.w * rs]; break;
      case OP_DIV: *d = x / regs[in.w * rs]; break;
      case OP_POW: *d = pow(x, regs[in.w * rs]); break;
      case OP_POWI: *d = powi(x, in.w); break;
      case OP_NEG: *d = -x; break;
      case OP_STORE: store(in.y, x); break;
      case OP_SQRT: *d = sqrt(x); break;
      case OP_EXP: *d = exp(x); break;
      case OP_LOG: *d = log(x); break;
      case OP_SIN: *d = sin(x); break;
      case OP_COS: *d = cos(x); break;
      case OP_TAN: *d = tan(x); break;
      case OP_TANH: *d = tanh(x); break;
      case OP_ABS: *d = fabs(x); break;
      case OP_ASIN: *d = asin(x); break;
      case OP_ACOS: *d = acos(x); break;
      case OP_ATAN: *d = atan(x); break;
      case OP_SINH: *d = sinh(x); break;
      case OP_COSH: *d = cosh(x); break;
      case OP_ATAN2: *d = atan2(x, regs[in.w * rs]); break;
      case OP_MIN: {
        const T y = regs[in.w * rs];
        *d = (x != x || y != y) ? x + y : (y < x ? y : x);
        break;
      }
      case OP_MAX: {
        const T y = regs[in.w * rs];
        *d = (x != x || y != y) ? x + y : (y > x ? y : x);
        break;
      }
      case OP_SIGN: *d = x > T(0) ? T(1) : (x < T(0) ? T(-1) : x); break;
      default: break;
    }
  }
}

struct ExprIO {
  const void *in[PML_EXPR_MAX_IO];
  int64_t in_stride[PML_EXPR_MAX_IO];
  void *out[PML_EXPR_MAX_IO];
  int64_t out_stride[PML_EXPR_MAX_IO];
};

template <typename T>
__global__ void __launch_bounds__(128) k_expr_eval(const int4 *code, int n_instr,
                                                   const double *consts, int n_regs, ExprIO io,
                                                   double t, int64_t count) {
  extern __shared__ unsigned char smem[];
  T *regs = reinterpret_cast<T *>(smem) + threadIdx.x;
  const int rs = blockDim.x;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * blockDim.x) {
    auto load = [&](int k) -> T {
      return static_cast<const T *>(io.in[k])[e * io.in_stride[k]];
    };
    auto store = [&](int k, T v) { static_cast<T *>(io.out[k])[e * io.out_stride[k]] = v; };
    run_program<T>(code, n_instr, consts, regs, rs, static_cast<T>(t), load, store);
  }
  (void)n_regs;
}

// One thread per batch element; the whole time loop in registers + LDS.
// LDS per thread: the program's registers, then y, the stage input and the
// stage accumulator (n rows each), then the stage derivative (n rows).
template <typename T>
__global__ void __launch_bounds__(64) k_ode_solve(const int4 *code, int n_instr,
                                                  const double *consts, int n_regs, int n,
                                                  int integrator, int64_t batch,
                                                  const double *t0, const double *t_step,
                                                  double dt, int64_t n_steps, T *y, T *snaps,
                                                  int64_t snap_every) {
  extern __shared__ unsigned char smem[];
  const int rs = blockDim.x;
  T *regs = reinterpret_cast<T *>(smem) + threadIdx.x;
  T *ys = regs + static_cast<int64_t>(n_regs) * rs;  // y
  T *yt = ys + static_cast<int64_t>(n) * rs;         // stage input
  T *acc = yt + static_cast<int64_t>(n) * rs;        // k1 + 2 k2 + 2 k3 (RK4)
  T *kk = acc + static_cast<int64_t>(n) * rs;        // f(t, .) of the stage
  const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (b >= batch) return;
  for (int i = 0; i < n; ++i) ys[i * rs] = y[b * n + i];
  const T h = static_cast<T>(dt);
  const T half = h / T(2);
  const double tb = t0[b], step = t_step[b];
  int64_t snap = 0;
  auto f = [&](T t, const T *src) {
    auto load = [&](int k) -> T { return src[k * rs]; };
    auto store = [&](int k, T v) { kk[k * rs] = v; };
    run_program<T>(code, n_instr, consts, regs, rs, t, load, store);
  };
  for (int64_t s = 0; s < n_steps; ++s) {
    // the reference's time grid: linspace(t0, t0 + n dt, n + 1)
    const T t = static_cast<T>(static_cast<double>(s) * step + tb);
    if (integrator == PML_INT_FE) {
      f(t, ys);
      for (int i = 0; i < n; ++i) ys[i * rs] = ys[i * rs] + h * kk[i * rs];
    } else if (integrator == PML_INT_MIDPOINT) {
      f(t, ys);
      for (int i = 0; i < n; ++i) yt[i * rs] = ys[i * rs] + half * kk[i * rs];
      f(t + half, yt);
      for (int i = 0; i < n; ++i) ys[i * rs] = ys[i * rs] + h * kk[i * rs];
    } else {
      // k1 = h f(t, y); k2 = h f(t + h/2, y + k1/2); k3 = h f(t + h/2,
      // y + k2/2); k4 = h f(t + h, y + k3); y + (k1 + 2 k2 + 2 k3 + k4) / 6
      f(t, ys);
      for (int i = 0; i < n; ++i) {
        const T k1 = h * kk[i * rs];
        acc[i * rs] = k1;
        yt[i * rs] = ys[i * rs] + k1 / T(2);
      }
      f(t + half, yt);
      for (int i = 0; i < n; ++i) {
        const T k2 = h * kk[i * rs];
        acc[i * rs] = acc[i * rs] + T(2) * k2;
        yt[i * rs] = ys[i * rs] + k2 / T(2);
      }
      f(t + half, yt);
      for (int i = 0; i < n; ++i) {
        const T k3 = h * kk[i * rs];
        acc[i * rs] = acc[i * rs] + T(2) * k3;
        yt[i * rs] = ys[i * rs] + k3;
      }
      f(t + h, yt);
      for (int i = 0; i < n; ++i) {
        const T k4 = h * kk[i * rs];
        ys[i * rs] = ys[i * rs] + (acc[i * rs] + k4) / T(6);
      }
    }
    if (snaps && (s + 1) % snap_every == 0) {
      T *dst = snaps + (snap * batch + b) * n;
      for (int i = 0; i < n; ++i) dst[i] = ys[i * rs];
      ++snap;
    }
  }
  for (int i = 0; i < n; ++i) y[b * n + i] = ys[i * rs];
}

}  // namespace pml

using namespace pml;

// A validated program resident in device memory.
struct pml_program_s {
  int4 *code = nullptr;
  double *consts = nullptr;
  int n_instr = 0, n_consts = 0, n_regs = 0, n_in = 0, n_out = 0;
  bool all_outputs_stored = false;
  int device = 0;
};

extern "C" int pml_program_create(const int32_t *code, int n_instr, const double *consts,
                                  int n_consts, int n_regs, int n_inputs, int n_outputs,
                                  pml_program *out) {
  if (!out) return fail(PML_E_INVALID, "null program handle");
  *out = nullptr;
  if (!code || n_instr <= 0 || n_instr > (1 << 20)) return fail(PML_E_INVALID, "empty program");
  if (n_regs <= 0 || n_regs > PML_EXPR_MAX_REGS)
    return fail(PML_E_INVALID, "program register count out of range");
  if (n_consts < 0 || (n_consts && !consts)) return fail(PML_E_INVALID, "bad constants");
  if (n_inputs < 0 || n_inputs > PML_EXPR_MAX_IO || n_outputs <= 0 ||
      n_outputs > PML_EXPR_MAX_IO)
    return fail(PML_E_INVALID, "program input/output count out of range");
  std::vector<char> stored(n_outputs, 0);
  for (int pc = 0; pc < n_instr; ++pc) {
    const int op = code[4 * pc], dst = code[4 * pc + 1], a = code[4 * pc + 2],
              b = code[4 * pc + 3];
    auto reg = [&](int r) { return r >= 0 && r < n_regs; };
    bool ok = op >= 0 && op < OP_COUNT;
    if (ok) {
      switch (op) {
        case OP_CONST: ok = reg(dst) && a >= 0 && a < n_consts; break;
        case OP_LOAD: ok = reg(dst) && a >= 0 && a < n_inputs; break;
        case OP_TIME: ok = reg(dst); break;
        case OP_STORE:
          ok = dst >= 0 && dst < n_outputs && reg(a);
          if (ok) stored[dst] = 1;
          break;
        case OP_POWI: ok = reg(dst) && reg(a); break;
        case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_POW: case OP_ATAN2:
        case OP_MIN: case OP_MAX:
          ok = reg(dst) && reg(a) && reg(b);
          break;
        default: ok = reg(dst) && reg(a); break;
      }
    }
    if (!ok) return fail(PML_E_INVALID, "invalid instruction " + std::to_string(pc));
  }
  auto *p = new pml_program_s();
  p->n_instr = n_instr;
  p->n_consts = n_consts;
  p->n_regs = n_regs;
  p->n_in = n_inputs;
  p->n_out = n_outputs;
  p->all_outputs_stored = true;
  for (char c : stored) p->all_outputs_stored = p->all_outputs_stored && c;
  hipError_t e = hipGetDevice(&p->device);
  if (e == hipSuccess) e = hipMalloc(&p->code, sizeof(int4) * n_instr);
  if (e == hipSuccess) e = hipMalloc(&p->consts, sizeof(double) * (n_consts > 0 ? n_consts : 1));
  if (e == hipSuccess)
    e = hipMemcpy(p->code, code, sizeof(int4) * n_instr, hipMemcpyHostToDevice);
  if (e == hipSuccess && n_consts)
    e = hipMemcpy(p->consts, consts, sizeof(double) * n_consts, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (p->code) (void)hipFree(p->code);
    if (p->consts) (void)hipFree(p->consts);
    delete p;
    return fail(PML_E_HIP, std::string("program upload: ") + hipGetErrorString(e));
  }
  *out = p;
  return PML_OK;
}

extern "C" int pml_program_destroy(pml_program p) {
  if (!p) return PML_OK;
  // a launch reading the program may still be in flight
  hipError_t e = hipDeviceSynchronize();
  (void)hipFree(p->code);
  (void)hipFree(p->consts);
  delete p;
  if (e != hipSuccess) return fail(PML_E_HIP, std::string("hipDeviceSynchronize: ") +
                                                 hipGetErrorString(e));
  return PML_OK;
}

static int check_device(const pml_program_s *p) {
  int dev = -1;
  PML_HIP_CHECK(hipGetDevice(&dev));
  if (dev != p->device) return fail(PML_E_INVALID, "program created on another device");
  return PML_OK;
}

extern "C" int pml_expr_eval(int precision, pml_program p, const void *const *inputs,
                             const int64_t *in_strides, void *const *outputs,
                             const int64_t *out_strides, double t, int64_t count,
                             void *stream) {
  if (!p) return fail(PML_E_INVALID, "null program");
  if (int rc = check_device(p); rc != PML_OK) return rc;
  if (count <= 0) return fail(PML_E_INVALID, "count must be positive");
  if ((p->n_in && (!inputs || !in_strides)) || !outputs || !out_strides)
    return fail(PML_E_INVALID, "null input/output tables");
  ExprIO io{};
  for (int k = 0; k < p->n_in; ++k) {
    if (!inputs[k] || in_strides[k] < 0) return fail(PML_E_INVALID, "bad program input");
    io.in[k] = inputs[k];
    io.in_stride[k] = in_strides[k];
  }
  for (int k = 0; k < p->n_out; ++k) {
    if (!outputs[k] || out_strides[k] < 0) return fail(PML_E_INVALID, "bad program output");
    io.out[k] = outputs[k];
    io.out_stride[k] = out_strides[k];
  }
  const int nt = 128;
  int64_t nb = (count + nt - 1) / nt;
  if (nb > 65536) nb = 65536;
  hipStream_t st = as_stream(stream);
  if (precision == PML_F64) {
    const size_t lds = static_cast<size_t>(p->n_regs) * nt * sizeof(double);
    hipLaunchKernelGGL(k_expr_eval<double>, dim3(nb), dim3(nt), lds, st, p->code, p->n_instr,
                       p->consts, p->n_regs, io, t, count);
  } else if (precision == PML_F32) {
    const size_t lds = static_cast<size_t>(p->n_regs) * nt * sizeof(float);
    hipLaunchKernelGGL(k_expr_eval<float>, dim3(nb), dim3(nt), lds, st, p->code, p->n_instr,
                       p->consts, p->n_regs, io, t, count);
  } else {
    return fail(PML_E_INVALID, "programs run in FP32 or FP64");
  }
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_ode_solve(int precision, pml_program p, int integrator_id, int64_t batch,
                             const double *t0, const double *t_step, double dt, int64_t n_steps,
                             void *y, void *snapshots, int64_t snapshot_every, void *stream) {
  if (!p) return fail(PML_E_INVALID, "null program");
  if (int rc = check_device(p); rc != PML_OK) return rc;
  const int n = p->n_out;
  if (p->n_in > n || !p->all_outputs_stored)
    return fail(PML_E_INVALID, "an ODE program reads y[0..n) and writes all n derivatives");
  if (batch <= 0 || !t0 || !t_step || !y || n_steps < 0)
    return fail(PML_E_INVALID, "bad ODE arguments");
  if (integrator_id != PML_INT_FE && integrator_id != PML_INT_MIDPOINT &&
      integrator_id != PML_INT_RK4)
    return fail(PML_E_UNSUPPORTED, "the ODE kernel runs the explicit integrators");
  if (snapshots && snapshot_every <= 0)
    return fail(PML_E_INVALID, "snapshot_every must be positive");
  if (n_steps == 0) return PML_OK;
  const int nt = 64;
  const size_t esz = precision == PML_F64 ? 8 : 4;
  const size_t lds = (static_cast<size_t>(p->n_regs) + 4 * static_cast<size_t>(n)) * nt * esz;
  if (lds > 160 * 1024)
    return fail(PML_E_UNSUPPORTED, "ODE system too large for the LDS-resident kernel");
  const dim3 grid((batch + nt - 1) / nt);
  hipStream_t st = as_stream(stream);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_ode_solve<double>, grid, dim3(nt), lds, st, p->code, p->n_instr,
                       p->consts, p->n_regs, n, integrator_id, batch, t0, t_step, dt, n_steps,
                       static_cast<double *>(y), static_cast<double *>(snapshots),
                       snapshot_every);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_ode_solve<float>, grid, dim3(nt), lds, st, p->code, p->n_instr,
                       p->consts, p->n_regs, n, integrator_id, batch, t0, t_step, dt, n_steps,
                       static_cast<float *>(y), static_cast<float *>(snapshots),
                       snapshot_every);
  else
    return fail(PML_E_INVALID, "the ODE kernel runs in FP32 or FP64");
  PML_LAUNCH_CHECK();
  return PML_OK;
}
