// Convolutional surrogate kernels for gfx950: 3x3 "same" convolutions with
// zero padding on NHWC BF16 tensors, FP32 accumulation.
//
//  * conv64: 64 -> 64 channels on the matrix cores
//    (v_mfma_f32_32x32x16_bf16).  Implicit GEMM: M = pixels, N = 64 output
//    channels, K = 9 taps x 64 input channels.  A persistent workgroup of
//    8 wavefronts keeps all 9 taps of the weights in LDS and loops over
//    8 x 32-pixel output tiles; each wavefront computes one 32-pixel row
//    of the tile against all 64 output channels (two 32x32 accumulator
//    blocks).  The same main loop serves the forward pass (epilogue: bias,
//    GELU, BF16 pre-activation and activation) and the input gradient
//    (weights pre-arranged flipped and transposed; epilogue: times GELU'
//    of the previous layer's pre-activation).
//  * wgrad64: the 64 -> 64 weight gradient on the matrix cores, M = input
//    channels, N = output channels, K = pixels, one wavefront per tap,
//    accumulated over a persistent pixel-tile loop and added to an FP32
//    gradient buffer.
//  * small layers (C_in -> 64 and 64 -> C_out with C <= 4): VALU kernels.
//  * Adam on FP32 masters, writing the BF16 weight layouts the kernels read.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pml_common.h"

namespace pml {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16;

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2 / pi)

__device__ __forceinline__ float gelu_f(float z) {
  return 0.5f * z * (1.0f + tanhf(kGeluC * (z + 0.044715f * z * z * z)));
}
__device__ __forceinline__ float gelu_grad_f(float z) {
  const float u = kGeluC * (z + 0.044715f * z * z * z);
  const float t = tanhf(u);
  const float du = kGeluC * (1.0f + 3.0f * 0.044715f * z * z);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

// ---------------------------------------------------------------------------
// conv64 (MFMA).  Tile: TR = 8 rows x TC = 32 columns of output pixels.
constexpr int kTR = 8, kTC = 32;
constexpr int kPR = kTR + 2, kPC = kTC + 2;  // input patch incl. halo
constexpr int kCP = 72;                      // LDS channel pitch (bf16)
constexpr int kPatchElems = kPR * kPC * kCP;
constexpr int kWElems = 9 * 64 * kCP;
constexpr size_t kConvLds = (kPatchElems + kWElems) * sizeof(bf16);

enum ConvMode { kFwd = 0, kBwd = 1 };

struct ConvArgs {
  const bf16 *in;     // [B, H, W, 64]
  const bf16 *wk;     // [9][64 n][64 k]
  const float *bias;  // [64] (forward)
  const bf16 *zprev;  // [B, H, W, 64] (backward)
  bf16 *z_out;        // forward: pre-activation (may be null)
  bf16 *a_out;        // forward: GELU(z) / backward: gradient
  int B, H, W;
  int tiles_h, tiles_w, n_tiles;
};

template <int MODE>
__global__ __launch_bounds__(512) void k_conv64(ConvArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  bf16 *patch = reinterpret_cast<bf16 *>(smem);
  bf16 *wl = patch + kPatchElems;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;

  // Weights into LDS once: 9 x 64 rows of 64 bf16 (8 chunks of 16 B).
  for (int c = tid; c < 9 * 64 * 8; c += 512) {
    const int row = c >> 3, part = c & 7;
    const uint4 v = reinterpret_cast<const uint4 *>(a.wk)[row * 8 + part];
    *reinterpret_cast<uint4 *>(wl + row * kCP + part * 8) = v;
  }

  for (int tile = blockIdx.x; tile < a.n_tiles; tile += gridDim.x) {
    const int per_img = a.tiles_h * a.tiles_w;
    const int b = tile / per_img;
    const int rem = tile - b * per_img;
    const int h0 = (rem / a.tiles_w) * kTR;
    const int w0 = (rem % a.tiles_w) * kTC;
    __syncthreads();  // previous tile's reads of the patch are done
    for (int c = tid; c < kPR * kPC * 8; c += 512) {
      const int px = c >> 3, part = c & 7;
      const int pr = px / kPC, pc = px - (px / kPC) * kPC;
      const int h = h0 + pr - 1, w = w0 + pc - 1;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (h >= 0 && h < a.H && w >= 0 && w < a.W)
        v = reinterpret_cast<const uint4 *>(
            a.in + ((static_cast<int64_t>(b) * a.H + h) * a.W + w) * 64)[part];
      *reinterpret_cast<uint4 *>(patch + (pr * kPC + pc) * kCP + part * 8) = v;
    }
    __syncthreads();

    f32x16 acc0, acc1;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc0[r] = acc1[r] = 0.0f;
    const int i = lane & 31;        // pixel column within the tile row (A row)
    const int kh8 = (lane >> 5) * 8;  // k offset of this lane's 8 values
#pragma unroll 1
    for (int t = 0; t < 9; ++t) {
      const int th = t / 3, tw = t - (t / 3) * 3;
      const bf16 *arow = patch + ((wave + th) * kPC + (i + tw)) * kCP + kh8;
      const bf16 *b0 = wl + (t * 64 + i) * kCP + kh8;
      const bf16 *b1 = wl + (t * 64 + 32 + i) * kCP + kh8;
#pragma unroll
      for (int kc = 0; kc < 4; ++kc) {
        const bf16x8 av = *reinterpret_cast<const bf16x8 *>(arow + kc * 16);
        const bf16x8 bv0 = *reinterpret_cast<const bf16x8 *>(b0 + kc * 16);
        const bf16x8 bv1 = *reinterpret_cast<const bf16x8 *>(b1 + kc * 16);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv1, acc1, 0, 0, 0);
      }
    }

    // Epilogue.  Accumulator element r of lane l holds output channel
    // n = l % 32 (+ 32 for acc1) of the tile-row pixel 8 (r / 4) + 4 (l / 32)
    // + r % 4.
    const int h = h0 + wave;
    if (h < a.H) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int px = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int w = w0 + px;
        if (w >= a.W) continue;
        const int64_t base = ((static_cast<int64_t>(b) * a.H + h) * a.W + w) * 64;
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
          const int n = nb * 32 + (lane & 31);
          const float v = nb == 0 ? acc0[r] : acc1[r];
          if (MODE == kFwd) {
            const float z = v + a.bias[n];
            const bf16 zb = static_cast<bf16>(z);
            if (a.z_out) a.z_out[base + n] = zb;
            a.a_out[base + n] = static_cast<bf16>(gelu_f(z));
          } else {
            const float zp = static_cast<float>(a.zprev[base + n]);
            a.a_out[base + n] = static_cast<bf16>(v * gelu_grad_f(zp));
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// wgrad64 (MFMA).  Pixel tiles of 8 x 32; nine wavefronts, one per tap.
constexpr int kGP = 40;  // LDS pitch (elements) of a transposed patch row
constexpr int kGPatch = 64 * kPR * kGP;     // [ci][prow][pcol]
constexpr int kGdz = 64 * kTR * kTC;        // [co][tile pixel]
constexpr size_t kWgradLds = (kGPatch + kGdz) * sizeof(bf16);

struct WgradArgs {
  const bf16 *x;   // [B, H, W, 64] layer input
  const bf16 *dz;  // [B, H, W, 64] gradient of the pre-activation
  float *dw;       // [9][64 ci][64 co], accumulated
  int B, H, W;
  int tiles_h, tiles_w, n_tiles;
};

__global__ __launch_bounds__(576) void k_wgrad64(WgradArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  bf16 *pt = reinterpret_cast<bf16 *>(smem);
  bf16 *dzt = pt + kGPatch;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;  // wave = tap
  const int th = wave / 3, tw = wave - (wave / 3) * 3;
  f32x16 acc[2][2];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[m][n][r] = 0.0f;

  for (int tile = blockIdx.x; tile < a.n_tiles; tile += gridDim.x) {
    const int per_img = a.tiles_h * a.tiles_w;
    const int b = tile / per_img;
    const int rem = tile - b * per_img;
    const int h0 = (rem / a.tiles_w) * kTR;
    const int w0 = (rem % a.tiles_w) * kTC;
    __syncthreads();
    // Transposed staging: patch[ci][pr][pc] and dz[co][tile pixel].
    for (int c = tid; c < kPR * kPC * 8; c += 576) {
      const int px = c >> 3, part = c & 7;
      const int pr = px / kPC, pc = px - (px / kPC) * kPC;
      const int h = h0 + pr - 1, w = w0 + pc - 1;
      bf16x8 v;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = static_cast<bf16>(0.0f);
      if (h >= 0 && h < a.H && w >= 0 && w < a.W)
        v = reinterpret_cast<const bf16x8 *>(
            a.x + ((static_cast<int64_t>(b) * a.H + h) * a.W + w) * 64)[part];
#pragma unroll
      for (int e = 0; e < 8; ++e) pt[((part * 8 + e) * kPR + pr) * kGP + pc] = v[e];
    }
    for (int c = tid; c < kTR * kTC * 8; c += 576) {
      const int px = c >> 3, part = c & 7;
      const int pr = px / kTC, pc = px - (px / kTC) * kTC;
      const int h = h0 + pr, w = w0 + pc;
      bf16x8 v;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = static_cast<bf16>(0.0f);
      if (h < a.H && w < a.W)
        v = reinterpret_cast<const bf16x8 *>(
            a.dz + ((static_cast<int64_t>(b) * a.H + h) * a.W + w) * 64)[part];
#pragma unroll
      for (int e = 0; e < 8; ++e) dzt[(part * 8 + e) * (kTR * kTC) + px] = v[e];
    }
    __syncthreads();
    // K = the tile's 256 pixels in steps of 16 (half a tile row).
    const int i = lane & 31, k8 = (lane >> 5) * 8;
#pragma unroll 1
    for (int ks = 0; ks < kTR * kTC / 16; ++ks) {
      const int prow = ks >> 1, pcol0 = (ks & 1) * 16 + k8;
      // A: rows = input channels, 8 consecutive pixels of the shifted patch.
      bf16x8 av[2];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const bf16 *src = pt + ((m * 32 + i) * kPR + prow + th) * kGP + pcol0 + tw;
#pragma unroll
        for (int e = 0; e < 8; ++e) av[m][e] = src[e];
      }
      // B: output channels, 8 consecutive tile pixels.
      bf16x8 bv[2];
#pragma unroll
      for (int n = 0; n < 2; ++n)
        bv[n] = *reinterpret_cast<const bf16x8 *>(dzt + (n * 32 + i) * (kTR * kTC) +
                                                  prow * kTC + pcol0);
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int n = 0; n < 2; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[m], bv[n], acc[m][n], 0, 0, 0);
    }
  }
  // dw[tap][ci][co] += acc: element r of lane l is (ci row 8 (r/4) + 4 (l/32)
  // + r%4, co column l%32) within each 32x32 block.
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ci = m * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int co = n * 32 + (lane & 31);
        atomicAdd(&a.dw[(wave * 64 + ci) * 64 + co], acc[m][n][r]);
      }
}

// ---------------------------------------------------------------------------
// Small layers (VALU).
struct SmallArgs {
  const bf16 *in;      // [B, H, W, Ci]
  const bf16 *w;       // [9][Ci][Co]
  const float *bias;   // [Co]
  const bf16 *zprev;   // backward: [B, H, W, Ci] pre-activation of the input
  bf16 *z_out;         // forward, Co == 64: pre-activation
  bf16 *a_out;         // forward, Co == 64: activation / backward: gradient
  float *y_out;        // forward, Co < 64: output (no activation)
  int B, H, W, Ci, Co;
};

// Forward, one thread per output pixel, all Co outputs in registers.
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_small_fwd(SmallArgs a) {
  const int64_t P = static_cast<int64_t>(a.B) * a.H * a.W;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int w = p % a.W;
    const int h = (p / a.W) % a.H;
    const int64_t b = p / (static_cast<int64_t>(a.W) * a.H);
    float acc[CO];
#pragma unroll
    for (int o = 0; o < CO; ++o) acc[o] = a.bias[o];
    for (int t = 0; t < 9; ++t) {
      const int hh = h + t / 3 - 1, ww = w + t % 3 - 1;
      if (hh < 0 || hh >= a.H || ww < 0 || ww >= a.W) continue;
      const bf16 *src = a.in + ((b * a.H + hh) * a.W + ww) * CI;
      const bf16 *wt = a.w + t * CI * CO;
#pragma unroll 4
      for (int c = 0; c < CI; ++c) {
        const float xv = static_cast<float>(src[c]);
#pragma unroll
        for (int o = 0; o < CO; ++o) acc[o] += xv * static_cast<float>(wt[c * CO + o]);
      }
    }
    if (CO == 64) {
#pragma unroll
      for (int o = 0; o < CO; ++o) {
        const bf16 zb = static_cast<bf16>(acc[o]);
        if (a.z_out) a.z_out[p * CO + o] = zb;
        a.a_out[p * CO + o] = static_cast<bf16>(gelu_f(acc[o]));
      }
    } else {
#pragma unroll
      for (int o = 0; o < CO; ++o) a.y_out[p * CO + o] = acc[o];
    }
  }
}

// Input gradient of a 64 -> CO layer, times GELU' of the input's
// pre-activation; one thread per pixel, 64 outputs.
template <int CO>
__global__ __launch_bounds__(256) void k_small_bwd(SmallArgs a) {
  const int64_t P = static_cast<int64_t>(a.B) * a.H * a.W;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < P;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int w = p % a.W;
    const int h = (p / a.W) % a.H;
    const int64_t b = p / (static_cast<int64_t>(a.W) * a.H);
    float acc[64];
#pragma unroll
    for (int c = 0; c < 64; ++c) acc[c] = 0.0f;
    for (int t = 0; t < 9; ++t) {
      // input (h, w) feeds output (h - th + 1, w - tw + 1)
      const int hh = h - t / 3 + 1, ww = w - t % 3 + 1;
      if (hh < 0 || hh >= a.H || ww < 0 || ww >= a.W) continue;
      const bf16 *g = a.in + ((b * a.H + hh) * a.W + ww) * CO;
      const bf16 *wt = a.w + t * 64 * CO;
#pragma unroll
      for (int o = 0; o < CO; ++o) {
        const float gv = static_cast<float>(g[o]);
#pragma unroll
        for (int c = 0; c < 64; ++c) acc[c] += gv * static_cast<float>(wt[c * CO + o]);
      }
    }
#pragma unroll
    for (int c = 0; c < 64; ++c) {
      const float zp = static_cast<float>(a.zprev[p * 64 + c]);
      a.a_out[p * 64 + c] = static_cast<bf16>(acc[c] * gelu_grad_f(zp));
    }
  }
}

// Weight and bias gradients of a small layer: lane = the 64-wide channel
// (output channel for CI -> 64, input channel for 64 -> CO); block-level
// partial sums over a strided pixel range, then FP32 atomics.
template <int CI, int CO>
__global__ __launch_bounds__(256) void k_small_wgrad(const bf16 *x, const bf16 *dz, float *dw,
                                                     float *db, int B, int H, int W) {
  constexpr bool kIn64 = CI == 64;  // 64 -> CO layer
  constexpr int kOther = kIn64 ? CO : CI;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float acc[9][kOther];
  float bacc[kIn64 ? CO : 1];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int o = 0; o < kOther; ++o) acc[t][o] = 0.0f;
#pragma unroll
  for (int o = 0; o < (kIn64 ? CO : 1); ++o) bacc[o] = 0.0f;
  float bias64 = 0.0f;
  const int64_t P = static_cast<int64_t>(B) * H * W;
  for (int64_t p = blockIdx.x * 4 + wave; p < P; p += (int64_t)gridDim.x * 4) {
    const int w = p % W;
    const int h = (p / W) % H;
    const int64_t b = p / (static_cast<int64_t>(W) * H);
    if (!kIn64) {
      const float g = static_cast<float>(dz[p * 64 + lane]);  // co = lane
      bias64 += g;
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const int hh = h + t / 3 - 1, ww = w + t % 3 - 1;
        if (hh < 0 || hh >= H || ww < 0 || ww >= W) continue;
        const bf16 *src = x + ((b * H + hh) * W + ww) * CI;
#pragma unroll
        for (int c = 0; c < CI; ++c) acc[t][c] += static_cast<float>(src[c]) * g;
      }
    } else {
      float g[CO];
#pragma unroll
      for (int o = 0; o < CO; ++o) {
        g[o] = static_cast<float>(dz[p * CO + o]);
        bacc[o] += g[o];
      }
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const int hh = h + t / 3 - 1, ww = w + t % 3 - 1;
        if (hh < 0 || hh >= H || ww < 0 || ww >= W) continue;
        const float xv = static_cast<float>(x[((b * H + hh) * W + ww) * 64 + lane]);  // ci = lane
#pragma unroll
        for (int o = 0; o < CO; ++o) acc[t][o] += xv * g[o];
      }
    }
  }
  // Reduce the four wavefronts through LDS, then one atomic per value.
  __shared__ float red[4][64];
#pragma unroll 1
  for (int t = 0; t < 9; ++t) {
#pragma unroll 1
    for (int o = 0; o < kOther; ++o) {
      red[wave][lane] = acc[t][o];
      __syncthreads();
      if (wave == 0) {
        const float s = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
        if (!kIn64) atomicAdd(&dw[(t * CI + o) * 64 + lane], s);   // [t][ci=o][co=lane]
        else atomicAdd(&dw[(t * 64 + lane) * CO + o], s);           // [t][ci=lane][co=o]
      }
      __syncthreads();
    }
  }
  if (!kIn64) {
    red[wave][lane] = bias64;
    __syncthreads();
    if (wave == 0) atomicAdd(&db[lane], red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane]);
  } else {
    if (lane == 0) {
#pragma unroll
      for (int o = 0; o < CO; ++o) atomicAdd(&db[o], bacc[o]);
    }
  }
}

// Bias gradient of a 64-channel layer: db[c] += sum over pixels of dz[p][c].
__global__ __launch_bounds__(256) void k_bias_grad64(const bf16 *dz, float *db, int64_t P) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float s = 0.0f;
  for (int64_t p = blockIdx.x * 4 + wave; p < P; p += (int64_t)gridDim.x * 4)
    s += static_cast<float>(dz[p * 64 + lane]);
  __shared__ float red[4][64];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0) atomicAdd(&db[lane], red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane]);
}

// Adam on FP32 masters.  Writes the BF16 weights in up to three layouts:
// plain [t][ci][co], forward-MFMA [t][co][ci] and backward-MFMA
// [8 - t][ci][co] ordered as [tap][n = ci][k = co].
__global__ __launch_bounds__(256) void k_adam(float *p, float *m, float *v, const float *g,
                                              int64_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2, bf16 *plain, bf16 *wf,
                                              bf16 *wb, int ci, int co) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    const float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
    p[i] = pi;
    const bf16 pb = static_cast<bf16>(pi);
    if (plain) plain[i] = pb;
    if (wf || wb) {
      const int t = static_cast<int>(i / (static_cast<int64_t>(ci) * co));
      const int r = static_cast<int>(i - static_cast<int64_t>(t) * ci * co);
      const int c_in = r / co, c_out = r - (r / co) * co;
      if (wf) wf[(t * co + c_out) * ci + c_in] = pb;
      if (wb) wb[((8 - t) * ci + c_in) * co + c_out] = pb;
    }
  }
}

// Per-sample scale s = max |u| over all fields and the normalised BF16 NHWC
// input x = u / s.  u is field-major [B][C][H*W] in FP64 or FP32.
template <typename T>
__global__ __launch_bounds__(256) void k_normalize(const T *u, bf16 *x, float *s_out, int C,
                                                   int64_t HW) {
  const int b = blockIdx.x;
  const T *ub = u + static_cast<int64_t>(b) * C * HW;
  float m = 0.0f;
  for (int64_t i = threadIdx.x; i < C * HW; i += blockDim.x)
    m = fmaxf(m, fabsf(static_cast<float>(ub[i])));
  __shared__ float red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const float sc = fmaxf(red[0], 1e-30f);
  if (threadIdx.x == 0) s_out[b] = sc;
  for (int64_t i = threadIdx.x; i < C * HW; i += blockDim.x) {
    const int c = static_cast<int>(i / HW);
    const int64_t px = i - c * HW;
    x[(static_cast<int64_t>(b) * HW + px) * C + c] =
        static_cast<bf16>(static_cast<float>(ub[i]) / sc);
  }
}

// Prediction u_next = u + s y (field-major FP64 / FP32): the identity part
// stays in the state's precision, only the learned increment passes through
// the BF16 network.  out may alias u.
template <typename T>
__global__ __launch_bounds__(256) void k_residual(const T *u, const float *y, const float *s,
                                                  T *out, int C, int64_t HW, int B) {
  const int64_t total = static_cast<int64_t>(B) * C * HW;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / (C * HW);
    const int64_t r = i - b * C * HW;
    const int c = static_cast<int>(r / HW);
    const int64_t px = r - c * HW;
    const int64_t e = (b * HW + px) * C + c;
    out[i] = u[i] + static_cast<T>(s[b]) * static_cast<T>(y[e]);
  }
}

// MSE loss of the network output y against the normalised increment
// (target - u) / s (u, target field-major FP32) and the BF16 gradient
// 2 (y - (target - u) / s) / n of the network output.
__global__ __launch_bounds__(256) void k_loss(const float *u, const float *y, const float *s,
                                              const float *target, bf16 *dy, double *loss, int C,
                                              int64_t HW, int B) {
  const int64_t total = static_cast<int64_t>(B) * C * HW;
  const float inv_n = 2.0f / static_cast<float>(total);
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = i / (C * HW);
    const int64_t r = i - b * C * HW;
    const int c = static_cast<int>(r / HW);
    const int64_t px = r - c * HW;
    const int64_t e = (b * HW + px) * C + c;
    const float d = y[e] - (target[i] - u[i]) / s[b];
    acc += static_cast<double>(d) * d;
    dy[e] = static_cast<bf16>(d * inv_n);
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

static int n_blocks(int64_t n, int cap = 4096) {
  int64_t b = (n + 255) / 256;
  if (b > cap) b = cap;
  return static_cast<int>(b < 1 ? 1 : b);
}

static int cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) cus = 256;
  }
  return cus;
}

}  // namespace pml

using namespace pml;

#define PML_SMALL_DISPATCH(KERNEL, ARGS, ...)                                          \
  do {                                                                                 \
    if (a.Ci == 1 && a.Co == 64) hipLaunchKernelGGL((KERNEL<1, 64>), __VA_ARGS__, ARGS); \
    else if (a.Ci == 2 && a.Co == 64) hipLaunchKernelGGL((KERNEL<2, 64>), __VA_ARGS__, ARGS); \
    else if (a.Ci == 3 && a.Co == 64) hipLaunchKernelGGL((KERNEL<3, 64>), __VA_ARGS__, ARGS); \
    else if (a.Ci == 64 && a.Co == 1) hipLaunchKernelGGL((KERNEL<64, 1>), __VA_ARGS__, ARGS); \
    else if (a.Ci == 64 && a.Co == 2) hipLaunchKernelGGL((KERNEL<64, 2>), __VA_ARGS__, ARGS); \
    else if (a.Ci == 64 && a.Co == 3) hipLaunchKernelGGL((KERNEL<64, 3>), __VA_ARGS__, ARGS); \
    else return fail(PML_E_INVALID, "small layers are C (1..3) -> 64 or 64 -> C");      \
  } while (0)

extern "C" int pml_conv64(int mode, const void *in, const void *wk, const float *bias,
                          const void *zprev, void *z_out, void *a_out, int B, int H, int W,
                          void *stream) {
  if (B <= 0 || H <= 0 || W <= 0 || !in || !wk || !a_out)
    return fail(PML_E_INVALID, "bad conv64 arguments");
  if (mode == kFwd && !bias) return fail(PML_E_INVALID, "forward needs a bias");
  if (mode == kBwd && !zprev) return fail(PML_E_INVALID, "backward needs the pre-activation");
  ConvArgs a{};
  a.in = static_cast<const bf16 *>(in);
  a.wk = static_cast<const bf16 *>(wk);
  a.bias = bias;
  a.zprev = static_cast<const bf16 *>(zprev);
  a.z_out = static_cast<bf16 *>(z_out);
  a.a_out = static_cast<bf16 *>(a_out);
  a.B = B;
  a.H = H;
  a.W = W;
  a.tiles_h = (H + kTR - 1) / kTR;
  a.tiles_w = (W + kTC - 1) / kTC;
  const int64_t nt = static_cast<int64_t>(B) * a.tiles_h * a.tiles_w;
  if (nt >= (int64_t(1) << 31)) return fail(PML_E_INVALID, "too many tiles");
  a.n_tiles = static_cast<int>(nt);
  const int grid = static_cast<int>(std::min<int64_t>(nt, cu_count()));
  hipStream_t st = as_stream(stream);
  if (mode == kFwd) {
    PML_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_conv64<kFwd>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, kConvLds));
    hipLaunchKernelGGL(k_conv64<kFwd>, dim3(grid), dim3(512), kConvLds, st, a);
  } else if (mode == kBwd) {
    PML_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_conv64<kBwd>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, kConvLds));
    hipLaunchKernelGGL(k_conv64<kBwd>, dim3(grid), dim3(512), kConvLds, st, a);
  } else {
    return fail(PML_E_INVALID, "mode must be 0 (forward) or 1 (input gradient)");
  }
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_wgrad64(const void *x, const void *dz, float *dw, int B, int H, int W,
                           void *stream) {
  if (B <= 0 || H <= 0 || W <= 0 || !x || !dz || !dw)
    return fail(PML_E_INVALID, "bad wgrad64 arguments");
  WgradArgs a{};
  a.x = static_cast<const bf16 *>(x);
  a.dz = static_cast<const bf16 *>(dz);
  a.dw = dw;
  a.B = B;
  a.H = H;
  a.W = W;
  a.tiles_h = (H + kTR - 1) / kTR;
  a.tiles_w = (W + kTC - 1) / kTC;
  const int64_t nt = static_cast<int64_t>(B) * a.tiles_h * a.tiles_w;
  if (nt >= (int64_t(1) << 31)) return fail(PML_E_INVALID, "too many tiles");
  a.n_tiles = static_cast<int>(nt);
  const int grid = static_cast<int>(std::min<int64_t>(nt, cu_count()));
  PML_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_wgrad64),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kWgradLds));
  hipLaunchKernelGGL(k_wgrad64, dim3(grid), dim3(576), kWgradLds, as_stream(stream), a);
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_conv_small(int mode, const void *in, const void *w, const float *bias,
                              const void *zprev, void *z_out, void *a_out, float *y_out, int Ci,
                              int Co, int B, int H, int W, void *stream) {
  if (B <= 0 || H <= 0 || W <= 0 || !in || !w) return fail(PML_E_INVALID, "bad conv arguments");
  SmallArgs a{};
  a.in = static_cast<const bf16 *>(in);
  a.w = static_cast<const bf16 *>(w);
  a.bias = bias;
  a.zprev = static_cast<const bf16 *>(zprev);
  a.z_out = static_cast<bf16 *>(z_out);
  a.a_out = static_cast<bf16 *>(a_out);
  a.y_out = y_out;
  a.B = B;
  a.H = H;
  a.W = W;
  a.Ci = Ci;
  a.Co = Co;
  hipStream_t st = as_stream(stream);
  const int nb = n_blocks(static_cast<int64_t>(B) * H * W, 1 << 16);
  if (mode == kFwd) {
    if (!bias) return fail(PML_E_INVALID, "forward needs a bias");
    if (Co == 64 && !a_out) return fail(PML_E_INVALID, "missing activation output");
    if (Co < 64 && !y_out) return fail(PML_E_INVALID, "missing output");
    PML_SMALL_DISPATCH(k_small_fwd, a, dim3(nb), dim3(256), 0, st);
  } else if (mode == kBwd) {
    // input gradient of a 64 -> Co layer (the caller passes Ci = 64)
    if (Ci != 64 || !zprev || !a_out) return fail(PML_E_INVALID, "bad input-gradient arguments");
    if (Co == 1) hipLaunchKernelGGL(k_small_bwd<1>, dim3(nb), dim3(256), 0, st, a);
    else if (Co == 2) hipLaunchKernelGGL(k_small_bwd<2>, dim3(nb), dim3(256), 0, st, a);
    else if (Co == 3) hipLaunchKernelGGL(k_small_bwd<3>, dim3(nb), dim3(256), 0, st, a);
    else return fail(PML_E_INVALID, "input gradient covers 64 -> 1..3 layers");
  } else {
    return fail(PML_E_INVALID, "mode must be 0 or 1");
  }
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_conv_small_wgrad(const void *x, const void *dz, float *dw, float *db, int Ci,
                                    int Co, int B, int H, int W, void *stream) {
  if (B <= 0 || H <= 0 || W <= 0 || !x || !dz || !dw || !db)
    return fail(PML_E_INVALID, "bad weight-gradient arguments");
  hipStream_t st = as_stream(stream);
  const int grid = 4 * cu_count();
  const bf16 *xb = static_cast<const bf16 *>(x);
  const bf16 *dzb = static_cast<const bf16 *>(dz);
#define PML_WG(CI, CO) \
  hipLaunchKernelGGL((k_small_wgrad<CI, CO>), dim3(grid), dim3(256), 0, st, xb, dzb, dw, db, B, H, W)
  if (Ci == 1 && Co == 64) PML_WG(1, 64);
  else if (Ci == 2 && Co == 64) PML_WG(2, 64);
  else if (Ci == 3 && Co == 64) PML_WG(3, 64);
  else if (Ci == 64 && Co == 1) PML_WG(64, 1);
  else if (Ci == 64 && Co == 2) PML_WG(64, 2);
  else if (Ci == 64 && Co == 3) PML_WG(64, 3);
  else return fail(PML_E_INVALID, "small layers are C (1..3) -> 64 or 64 -> C");
#undef PML_WG
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_bias_grad64(const void *dz, float *db, int64_t pixels, void *stream) {
  if (pixels <= 0 || !dz || !db) return fail(PML_E_INVALID, "bad bias-gradient arguments");
  hipLaunchKernelGGL(k_bias_grad64, dim3(4 * cu_count()), dim3(256), 0, as_stream(stream),
                     static_cast<const bf16 *>(dz), db, pixels);
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_adam(float *p, float *m, float *v, const float *g, int64_t n, float lr,
                        float beta1, float beta2, float eps, int step, void *plain, void *wf,
                        void *wb, int ci, int co, void *stream) {
  if (n <= 0 || !p || !m || !v || !g || step < 1) return fail(PML_E_INVALID, "bad Adam arguments");
  if ((wf || wb) && (ci <= 0 || co <= 0 || n != 9LL * ci * co))
    return fail(PML_E_INVALID, "weight layouts need n = 9 ci co");
  const float c1 = 1.0f - powf(beta1, static_cast<float>(step));
  const float c2 = 1.0f - powf(beta2, static_cast<float>(step));
  hipLaunchKernelGGL(k_adam, dim3(n_blocks(n)), dim3(256), 0, as_stream(stream), p, m, v, g, n,
                     lr, beta1, beta2, eps, c1, c2, static_cast<bf16 *>(plain),
                     static_cast<bf16 *>(wf), static_cast<bf16 *>(wb), ci, co);
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_surrogate_normalize(int precision, const void *u, void *x, float *s, int B,
                                       int C, int64_t HW, void *stream) {
  if (B <= 0 || C <= 0 || HW <= 0 || !u || !x || !s)
    return fail(PML_E_INVALID, "bad normalisation arguments");
  hipStream_t st = as_stream(stream);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_normalize<double>, dim3(B), dim3(256), 0, st,
                       static_cast<const double *>(u), static_cast<bf16 *>(x), s, C, HW);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_normalize<float>, dim3(B), dim3(256), 0, st,
                       static_cast<const float *>(u), static_cast<bf16 *>(x), s, C, HW);
  else
    return fail(PML_E_INVALID, "normalisation reads FP32 or FP64");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_surrogate_residual(int precision, const void *u, const float *y,
                                      const float *s, void *out, int B, int C, int64_t HW,
                                      void *stream) {
  if (B <= 0 || C <= 0 || HW <= 0 || !u || !y || !s || !out)
    return fail(PML_E_INVALID, "bad residual arguments");
  hipStream_t st = as_stream(stream);
  const int nb = n_blocks(static_cast<int64_t>(B) * C * HW);
  if (precision == PML_F64)
    hipLaunchKernelGGL(k_residual<double>, dim3(nb), dim3(256), 0, st,
                       static_cast<const double *>(u), y, s, static_cast<double *>(out), C, HW, B);
  else if (precision == PML_F32)
    hipLaunchKernelGGL(k_residual<float>, dim3(nb), dim3(256), 0, st,
                       static_cast<const float *>(u), y, s, static_cast<float *>(out), C, HW, B);
  else
    return fail(PML_E_INVALID, "residual works in FP32 or FP64");
  PML_LAUNCH_CHECK();
  return PML_OK;
}

extern "C" int pml_surrogate_loss(const float *u, const float *y, const float *s,
                                  const float *target, void *dy, double *loss, int B, int C,
                                  int64_t HW, void *stream) {
  if (B <= 0 || C <= 0 || HW <= 0 || !u || !y || !s || !target || !dy || !loss)
    return fail(PML_E_INVALID, "bad loss arguments");
  hipLaunchKernelGGL(k_loss, dim3(n_blocks(static_cast<int64_t>(B) * C * HW, 2048)), dim3(256),
                     0, as_stream(stream), u, y, s, target, static_cast<bf16 *>(dy), loss, C, HW,
                     B);
  PML_LAUNCH_CHECK();
  return PML_OK;
}
