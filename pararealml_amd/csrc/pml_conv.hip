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
//  * wgrad64: the 64 -> 64 weight (and bias) gradient on the matrix cores, M = input
//    channels, N = output channels, K = pixels, one wavefront per tap,
//    accumulated over a persistent pixel-tile loop and added to an FP32
//    gradient buffer.
//  * small layers (C_in -> 64 and 64 -> C_out with C <= 3): VALU kernels with
//    vectorised 64-channel rows (a lane group per pixel); the first layer
//    reads the normalised input and its weights in FP32.
//  * Adam on FP32 masters, writing the BF16 weight layouts the kernels read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "pml_common.h"

namespace pml {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16;

// GELU, tanh form: 0.5 z (1 + tanh(u)), u = sqrt(2/pi) (z + 0.044715 z^3).
// Evaluated as z * sigmoid(2u) (0.5 (1 + tanh(u)) = sigmoid(2u)) with one
// hardware exp2 and one reciprocal: 2^(z (kGeluA + kGeluB z^2)) = exp(-2u).
constexpr float kGeluA = -2.3022081983f;   // -2 sqrt(2/pi) log2(e)
constexpr float kGeluB = -0.1029432396f;   // kGeluA * 0.044715
constexpr float kGeluC = 1.5957691216f;    // 2 sqrt(2/pi)
constexpr float kGeluD = 0.2140644488f;    // kGeluC * 3 * 0.044715

__device__ __forceinline__ float gelu_sig(float z) {
  const float e = __builtin_amdgcn_exp2f(z * (kGeluA + kGeluB * (z * z)));
  return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu_f(float z) { return z * gelu_sig(z); }
// d/dz [z s(z)] = s + z s (1 - s) 2 u'(z), 2 u' = kGeluC + kGeluD z^2.
__device__ __forceinline__ float gelu_grad_f(float z) {
  const float s = gelu_sig(z);
  return s + z * s * (1.0f - s) * (kGeluC + kGeluD * (z * z));
}

// ---------------------------------------------------------------------------
// conv64 (MFMA).  Output tile: kTR = 16 rows x kTC = 32 columns x 64
// channels.  The input patch (18 x 34 pixels) sits in LDS as 144-byte rows
// (64 BF16 channels + 16 bytes of padding: 16 lanes reading consecutive
// rows hit all 64 banks once, and every patch address is a per-lane base
// plus an immediate).  All nine taps of the weights sit beside it as
// unpadded 128-byte rows whose 16-byte chunks are XOR-swizzled by
// (row / 2) mod 8, which spreads any 16 consecutive rows over the 64 banks.
// The MFMA's A operand is a weight block (32 output channels x 16 input
// channels), B a pixel block (16 input channels x 32 pixels of one row).
// Wavefront w owns output rows 4 (w / 2) .. + 3 and output channels
// 32 (w % 2) .. + 31: for each tap column and 16-channel slice it reads the
// six patch rows and the three kernel-row weight blocks those four rows need
// once and issues 12 MFMAs.  The next tile's patch is loaded into registers
// while the current one is computed.
constexpr int kTR = 16, kTC = 32;
constexpr int kPR = kTR + 2, kPC = kTC + 2;  // input patch incl. halo
constexpr int kPatchPitch = 144;             // bytes per patch pixel
constexpr int kPatchChunks = kPR * kPC * 8;  // 16-byte chunks
constexpr int kPatchBytes = kPR * kPC * kPatchPitch;
constexpr int kWBytes = 9 * 64 * 128;
constexpr size_t kConvLds = kPatchBytes + kWBytes;
constexpr int kConvThreads = 512;
constexpr int kPrefetch = (kPatchChunks + kConvThreads - 1) / kConvThreads;
static_assert(kConvLds <= 160 * 1024, "conv64 LDS budget");

// Byte offset of 16-byte chunk `chunk` (0..7) of weight row `row`.
__device__ __forceinline__ int wswz(int row, int chunk) {
  return row * 128 + ((chunk ^ ((row >> 1) & 7)) << 4);
}

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
  unsigned bytes;     // bytes of each [B, H, W, 64] tensor (< 2^31)
};

// Raw buffer resources over the activation tensors: loads beyond `bytes`
// return 0 and stores beyond it are dropped, so halo pixels outside the
// image and output pixels outside it need no branches.  kDrop is an offset
// past every such buffer.
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr unsigned kDrop = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t conv_rsrc(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, static_cast<int>(bytes),
                                           0x00020000);
}

struct TilePos {
  int b, h0, w0;
};
__device__ __forceinline__ TilePos tile_pos(const ConvArgs &a, int tile) {
  const int per_img = a.tiles_h * a.tiles_w;
  const int b = tile / per_img;
  const int rem = tile - b * per_img;
  return {b, (rem / a.tiles_w) * kTR, (rem % a.tiles_w) * kTC};
}

template <int MODE>
__global__ __launch_bounds__(kConvThreads) void k_conv64(ConvArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned char *wl = smem;
  unsigned char *patch = smem + kWBytes;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int rg = wave >> 1, nh = wave & 1;
  const int i = lane & 31, kh = lane >> 5;
  const __amdgpu_buffer_rsrc_t r_in = conv_rsrc(a.in, a.bytes);
  const __amdgpu_buffer_rsrc_t r_out = conv_rsrc(a.a_out, a.bytes);
  const __amdgpu_buffer_rsrc_t r_z =
      conv_rsrc(MODE == kFwd ? static_cast<const void *>(a.z_out) : a.zprev,
                (MODE == kFwd ? a.z_out != nullptr : true) ? a.bytes : 0u);

  // Weights into LDS once: 9 x 64 rows of 8 chunks.
  for (int c = tid; c < 9 * 64 * 8; c += kConvThreads)
    *reinterpret_cast<uint4 *>(wl + wswz(c >> 3, c & 7)) = reinterpret_cast<const uint4 *>(a.wk)[c];
  // Accumulator element r = 4q + e of a lane is output channel
  // 32 nh + 8q + 4kh + e of pixel column i; the forward accumulators start
  // at the bias.
  float bias[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    bias[r] = MODE == kFwd ? a.bias[nh * 32 + 8 * (r >> 2) + 4 * kh + (r & 3)] : 0.0f;
    // Settle the load here: a load still pending at the tile loop's header
    // would make every tile wait for all its own prefetches before its
    // first MFMA.
    asm volatile("" : "+v"(bias[r]));
  }

  // Patch chunks this thread stages: pixel (pr, pc) of the patch, chunk
  // `part`, at a fixed byte offset from the tile origin.
  int rel[kPrefetch], prc[kPrefetch];
#pragma unroll
  for (int s = 0; s < kPrefetch; ++s) {
    const int c = tid + s * kConvThreads;
    const int px = c >> 3, part = c & 7;
    const int pr = px / kPC, pc = px - pr * kPC;
    rel[s] = ((pr - 1) * a.W + (pc - 1)) * 128 + part * 16;
    prc[s] = c < kPatchChunks ? (pr << 8) | pc : -1;  // -1: beyond the patch
  }
  u32x4 pre[kPrefetch];
  auto fetch = [&](int tile) {
    const TilePos tp = tile_pos(a, tile);
    const int origin = ((tp.b * a.H + tp.h0) * a.W + tp.w0) * 128;
#pragma unroll
    for (int s = 0; s < kPrefetch; ++s) {
      const unsigned h = static_cast<unsigned>(tp.h0 + (prc[s] >> 8) - 1);
      const unsigned w = static_cast<unsigned>(tp.w0 + (prc[s] & 0xff) - 1);
      const bool ok = prc[s] >= 0 && h < static_cast<unsigned>(a.H) && w < static_cast<unsigned>(a.W);
      pre[s] = __builtin_amdgcn_raw_buffer_load_b128(
          r_in, ok ? static_cast<unsigned>(origin + rel[s]) : kDrop, 0, 0);
    }
  };

  // Per-lane LDS bases: patch pixel (4 rg, i) chunk kh, and weight row
  // 32 nh + i of tap 0 for each 16-channel slice kc.
  const int pbase = ((4 * rg) * kPC + i) * kPatchPitch + kh * 16;
  int wbase[4];
#pragma unroll
  for (int kc = 0; kc < 4; ++kc) wbase[kc] = wswz(nh * 32 + i, 2 * kc + kh);

  int tile = blockIdx.x;
  if (tile < a.n_tiles) fetch(tile);
  for (; tile < a.n_tiles; tile += gridDim.x) {
    __syncthreads();  // the previous tile's patch reads are done
#pragma unroll
    for (int s = 0; s < kPrefetch; ++s) {
      const int c = tid + s * kConvThreads;
      if (c < kPatchChunks)
        *reinterpret_cast<u32x4 *>(patch + (c >> 3) * kPatchPitch + (c & 7) * 16) = pre[s];
    }
    __syncthreads();
    const TilePos tp = tile_pos(a, tile);
    if (tile + static_cast<int>(gridDim.x) < a.n_tiles) fetch(tile + gridDim.x);

    // The two wavefronts sharing a SIMD (w and w + 4) run the MFMA phase
    // at different priorities, so the first finishes its matrix work and
    // runs its VALU epilogue while the second still issues MFMAs.
    if (wave < 4) __builtin_amdgcn_s_setprio(2);
    else __builtin_amdgcn_s_setprio(1);
    f32x16 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[j][r] = bias[r];
    // 12 steps (tap column tw, 16-channel slice kc); the fragments of step
    // st + 1 are read while the MFMAs of step st run.
    bf16x8 pv[2][6], wv[2][3];
    auto load_step = [&](int st, bf16x8 (&p6)[6], bf16x8 (&w3)[3]) {
      const int tw = st >> 2, kc = st & 3;
#pragma unroll
      for (int r = 0; r < 6; ++r)
        p6[r] = *reinterpret_cast<const bf16x8 *>(patch + pbase + (r * kPC + tw) * kPatchPitch +
                                                  kc * 32);
#pragma unroll
      for (int th = 0; th < 3; ++th)
        w3[th] = *reinterpret_cast<const bf16x8 *>(wl + wbase[kc] + (th * 3 + tw) * 64 * 128);
    };
    load_step(0, pv[0], wv[0]);
#pragma unroll
    for (int st = 0; st < 12; ++st) {
      const int cur = st & 1;
      if (st + 1 < 12) load_step(st + 1, pv[cur ^ 1], wv[cur ^ 1]);
      // Keep the reads ahead of this step's MFMAs (the scheduler would
      // otherwise sink each read next to its first use).
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int th = 0; th < 3; ++th)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv[cur][th], pv[cur][j + th], acc[j],
                                                           0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_s_setprio(0);

    // Epilogue: 4 consecutive channels per 8-byte access; pixels outside
    // the image use the drop offset.  The backward pass issues all its
    // pre-activation loads before using any.
    const int w = tp.w0 + i;
    unsigned base[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int h = tp.h0 + 4 * rg + j;
      base[j] = h < a.H && w < a.W
                    ? ((static_cast<unsigned>(tp.b * a.H + h) * a.W + w) * 64 + nh * 32 + 4 * kh) * 2
                    : kDrop;
    }
    u32x2 zp[4][4];
    if (MODE == kBwd) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          zp[j][q] = __builtin_amdgcn_raw_buffer_load_b64(r_z, base[j] + 16 * q, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        bf16x4 o4;
        if (MODE == kFwd) {
          bf16x4 z4;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float z = acc[j][4 * q + e];
            z4[e] = static_cast<bf16>(z);
            o4[e] = static_cast<bf16>(gelu_f(z));
          }
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, z4), r_z,
                                                base[j] + 16 * q, 0, 0);
        } else {
          const bf16x4 zq = __builtin_bit_cast(bf16x4, zp[j][q]);
#pragma unroll
          for (int e = 0; e < 4; ++e)
            o4[e] = static_cast<bf16>(acc[j][4 * q + e] * gelu_grad_f(static_cast<float>(zq[e])));
        }
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, o4), r_out,
                                              base[j] + 16 * q, 0, 0);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// wgrad64 (MFMA): dw[t][ci][co] += sum_p x[p + off_t][ci] dz[p][co] and,
// optionally, db[co] += sum_p dz[p][co].  M = input channels, N = output
// channels, K = pixels.  Pixel tiles of kGR = 8 rows x kGC = 32 columns;
// the MFMA's K = 16 covers two pixel columns of 8 rows.  The input patch
// (10 x 34 pixels) and the dz tile are staged channel-major in LDS: a
// 16-byte chunk holds one channel at 8 consecutive rows of one column, and
// the patch is kept as three copies shifted by the kernel row (copy g holds
// patch rows g .. g + 7), so every MFMA operand is one aligned 16-byte read.
// Channel rows are padded to an odd number of chunks, which makes 16
// consecutive channels fall into distinct banks.  The transposition is done
// in registers while staging: a thread loads the 16-byte chunks (8 channels)
// of 4 consecutive rows and writes 8 channel runs of 4 rows.
// Twelve wavefronts: wave w owns the kernel row g = w / 4, input channels
// 32 (w % 2) .. + 31, output channels 32 ((w / 2) % 2) .. + 31 and the
// three taps of that row; the two wavefronts with g = 0 and w % 2 = 0 also
// sum their dz operands (VALU) for the bias gradient.
constexpr int kGR = 8, kGC = 32;
constexpr int kGXC = kGC + 2;                 // patch columns
constexpr int kGXPitch = 35 * 16;             // bytes per channel of a shifted patch copy
constexpr int kGDPitch = 33 * 16;             // bytes per channel of the dz tile
constexpr int kGXBytes = 64 * kGXPitch;
constexpr int kGDBytes = 64 * kGDPitch;
constexpr size_t kWgradLds = 3 * kGXBytes + kGDBytes;
constexpr int kWgradThreads = 768;
constexpr int kGXTasks = 3 * 2 * 8 * kGXC;    // (copy, row group, chunk, column)
constexpr int kGXWaves = (kGXTasks + 63) / 64;
constexpr int kGDTasks = 2 * 8 * kGC;         // (row group, chunk, column)
constexpr int kGDWaves = kGDTasks / 64;
constexpr int kGSlots = (kGXWaves + kGDWaves + 11) / 12;
static_assert(kWgradLds <= 160 * 1024, "wgrad64 LDS budget");
static_assert(kGDTasks % 64 == 0, "dz staging fills whole wavefronts");

struct WgradArgs {
  const bf16 *x;   // [B, H, W, 64] layer input
  const bf16 *dz;  // [B, H, W, 64] gradient of the pre-activation
  float *dw;       // [9][64 ci][64 co], accumulated
  float *db;       // [64], accumulated (may be null)
  float *part;     // per-block partial sums (kWgradPart floats a block), or null
  int B, H, W;
  int tiles_h, tiles_w, n_tiles;
  unsigned bytes;  // bytes of each [B, H, W, 64] tensor (< 2^31)
};
// Per-block partial sums of the weight-gradient kernels (deterministic
// mode): the weight gradient, then the bias gradient.
constexpr int kWgradDw = 9 * 64 * 64;
constexpr int kWgradPart = kWgradDw + 64;

__global__ __launch_bounds__(kWgradThreads) void k_wgrad64(WgradArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int i = lane & 31, kh = lane >> 5;
  const __amdgpu_buffer_rsrc_t r_x = conv_rsrc(a.x, a.bytes);
  const __amdgpu_buffer_rsrc_t r_dz = conv_rsrc(a.dz, a.bytes);

  // Staging slots: wave-task s * 12 + wave; the first kGXWaves move patch
  // chunks, the next kGDWaves dz chunks.  Per slot: row/column offset of
  // the task's first row from the tile origin, and its LDS byte address.
  int sh[kGSlots], sw[kGSlots], sl[kGSlots];
#pragma unroll
  for (int s = 0; s < kGSlots; ++s) {
    const int wt = s * 12 + wave;
    sh[s] = -1000000;  // no task: never inside the image
    sw[s] = 0;
    sl[s] = 0;
    if (wt < kGXWaves) {
      const int t = wt * 64 + lane;
      if (t < kGXTasks) {
        const int c = t % kGXC, r1 = t / kGXC;
        const int part = r1 & 7, kk = (r1 >> 3) & 1, g = r1 >> 4;
        sh[s] = g + 4 * kk - 1;
        sw[s] = c - 1;
        sl[s] = g * kGXBytes + part * 8 * kGXPitch + c * 16 + kk * 8;
        sw[s] = (sw[s] << 4) | part;
      }
    } else if (wt < kGXWaves + kGDWaves) {
      const int t = (wt - kGXWaves) * 64 + lane;
      const int c = t % kGC, r1 = t / kGC;
      const int part = r1 & 7, kk = r1 >> 3;
      sh[s] = 4 * kk;
      sw[s] = (c << 4) | part;
      sl[s] = 3 * kGXBytes + part * 8 * kGDPitch + c * 16 + kk * 8;
    }
  }
  u32x4 pre[kGSlots][4];
  auto fetch = [&](int tile) {
    const int per_img = a.tiles_h * a.tiles_w;
    const int b = tile / per_img, rem = tile - b * per_img;
    const int h0 = (rem / a.tiles_w) * kGR, w0 = (rem % a.tiles_w) * kGC;
#pragma unroll
    for (int s = 0; s < kGSlots; ++s) {
      const bool from_x = s * 12 + wave < kGXWaves;  // wave-uniform
      const int w = w0 + (sw[s] >> 4), part = sw[s] & 15;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int h = h0 + sh[s] + k;
        const bool ok = static_cast<unsigned>(h) < static_cast<unsigned>(a.H) &&
                        static_cast<unsigned>(w) < static_cast<unsigned>(a.W);
        const unsigned off =
            ok ? (static_cast<unsigned>((b * a.H + h) * a.W + w) * 64 + part * 8) * 2 : kDrop;
        pre[s][k] = __builtin_amdgcn_raw_buffer_load_b128(from_x ? r_x : r_dz, off, 0, 0);
      }
    }
  };

  const int g = wave >> 2, m = wave & 1, n = (wave >> 1) & 1;
  const bool bias_wave = g == 0 && m == 0 && a.db != nullptr;
  const int abase = g * kGXBytes + (m * 32 + i) * kGXPitch + kh * 16;
  const int bbase = 3 * kGXBytes + (n * 32 + i) * kGDPitch + kh * 16;
  f32x16 acc[3];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[0][r] = acc[1][r] = acc[2][r] = 0.0f;
  float bsum = 0.0f;  // dz sum of output channel 32 n + i (bias wavefronts)

  int tile = blockIdx.x;
  if (tile < a.n_tiles) fetch(tile);
  for (; tile < a.n_tiles; tile += gridDim.x) {
    __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
    for (int s = 0; s < kGSlots; ++s) {
      if (sh[s] > -1000000) {
        const int pitch = s * 12 + wave < kGXWaves ? kGXPitch : kGDPitch;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int d = e >> 1;
          u32x2 v;
          if (e & 1) {
            v[0] = (pre[s][0][d] >> 16) | (pre[s][1][d] & 0xffff0000u);
            v[1] = (pre[s][2][d] >> 16) | (pre[s][3][d] & 0xffff0000u);
          } else {
            v[0] = (pre[s][0][d] & 0xffffu) | (pre[s][1][d] << 16);
            v[1] = (pre[s][2][d] & 0xffffu) | (pre[s][3][d] << 16);
          }
          *reinterpret_cast<u32x2 *>(smem + sl[s] + e * pitch) = v;
        }
      }
    }
    __syncthreads();
    if (tile + static_cast<int>(gridDim.x) < a.n_tiles) fetch(tile + gridDim.x);

    // 16 K-steps of two pixel columns; the three wavefronts of each SIMD
    // hide one another's LDS latency.
#pragma unroll
    for (int st = 0; st < kGC / 2; ++st) {
      bf16x8 av[3];
#pragma unroll
      for (int tw = 0; tw < 3; ++tw)
        av[tw] = *reinterpret_cast<const bf16x8 *>(smem + abase + (2 * st + tw) * 16);
      const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(smem + bbase + 2 * st * 16);
#pragma unroll
      for (int tw = 0; tw < 3; ++tw)
        acc[tw] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[tw], bv, acc[tw], 0, 0, 0);
      if (bias_wave) {
#pragma unroll
        for (int e = 0; e < 8; ++e) bsum += static_cast<float>(bv[e]);
      }
    }
  }
  // dw[g * 3 + tw][ci][co] += acc[tw]: element r of lane l is
  // (ci = 32 m + 8 (r / 4) + 4 (l / 32) + r % 4, co = 32 n + l % 32).
  // With a partial buffer every block writes its sums (each element once;
  // k_sum_blocks adds them in block order), else it adds them atomically.
  float *pb = a.part ? a.part + static_cast<size_t>(blockIdx.x) * kWgradPart : nullptr;
#pragma unroll
  for (int tw = 0; tw < 3; ++tw)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int ci = m * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
      const int idx = ((g * 3 + tw) * 64 + ci) * 64 + n * 32 + i;
      if (pb) pb[idx] = acc[tw][r];
      else atomicAdd(&a.dw[idx], acc[tw][r]);
    }
  // Lanes i and i + 32 summed the two column halves of channel 32 n + i.
  if (bias_wave) {
    bsum += __shfl_xor(bsum, 32);
    if (kh == 0) {
      if (pb) pb[kWgradDw + n * 32 + i] = bsum;
      else atomicAdd(&a.db[n * 32 + i], bsum);
    }
  }
}

// The partials of a weight-gradient launch are added in a fixed order, so
// the result does not depend on the order the workgroups ran in: pass 1
// sums segments of kSumSeg consecutive workgroups (seg[y][i] = sum over
// b in segment y of part[b][i]), pass 2 adds the segments in order into
// dw[i] (i < n_dw) and db[i - n_dw] (n_dw <= i < n).
constexpr int kSumSeg = 16;
__global__ void k_sum_segments(const float *part, int n_blocks, int per_block, int n,
                               float *seg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b0 = blockIdx.y * kSumSeg, b1 = min(b0 + kSumSeg, n_blocks);
  float s = 0.0f;
  for (int b = b0; b < b1; ++b) s += part[static_cast<size_t>(b) * per_block + i];
  seg[static_cast<size_t>(blockIdx.y) * n + i] = s;
}
__global__ void k_sum_blocks(const float *seg, int n_segs, int n, float *dw, int n_dw,
                             float *db) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int y = 0; y < n_segs; ++y) s += seg[static_cast<size_t>(y) * n + i];
  if (i < n_dw) dw[i] += s;
  else db[i - n_dw] += s;
}

// ---------------------------------------------------------------------------
// Small layers (VALU).  NHWC rows of 64 BF16 channels are 128 bytes; every
// kernel gives a pixel a group of lanes that each own a contiguous run of
// channels, so the 64-channel traffic moves as 16-byte (8 channels) or
// 8-byte (4 channels) accesses and a wavefront covers 8 or 4 neighbouring
// pixels.  Weights live in LDS as FP32.  A pixel's nine taps are addressed
// as its own byte offset plus a per-tap constant; taps outside the image
// (from a 9-bit mask computed once per pixel) read through the drop offset,
// which returns zero.  With one field a thread handles two pixels per
// iteration, all their loads issued before the arithmetic.
struct SmallArgs {
  const bf16 *in;      // [B, H, W, Ci] (64 -> Co layers)
  const bf16 *w;       // [9][Ci][Co] (64 -> Co layers)
  const float *in32;   // [B, H, W, Ci] FP32 (Ci -> 64 layers)
  const float *w32;    // [9][Ci][64] FP32 (Ci -> 64 layers)
  const float *bias;   // [Co]
  const bf16 *zprev;   // backward: [B, H, W, Ci] pre-activation of the input
  bf16 *z_out;         // forward, Co == 64: pre-activation (may be null)
  bf16 *a_out;         // forward, Co == 64: activation / backward: gradient
  float *y_out;        // forward, Co < 64: output (no activation)
  int B, H, W, Ci, Co;
};

// Bit 3 th + tw is set when the tap at offset (th - 1, tw - 1) of pixel
// (h, w) lies in the image; MIRROR uses offset (1 - th, 1 - tw).
template <bool MIRROR>
__device__ __forceinline__ unsigned tap_mask(int h, int w, int H, int W) {
  const unsigned lo_r = h > 0 ? 1u : 0u, hi_r = h + 1 < H ? 1u : 0u;
  const unsigned lo_c = w > 0 ? 1u : 0u, hi_c = w + 1 < W ? 1u : 0u;
  const unsigned r = MIRROR ? (hi_r | 2u | (lo_r << 2)) : (lo_r | 2u | (hi_r << 2));
  const unsigned c = MIRROR ? (hi_c | 2u | (lo_c << 2)) : (lo_c | 2u | (hi_c << 2));
  return ((r & 1u) ? c : 0u) | ((r & 2u) ? c << 3 : 0u) | ((r & 4u) ? c << 6 : 0u);
}
// Byte offset of tap t relative to the pixel, `bpp` bytes per pixel.
template <bool MIRROR>
__device__ __forceinline__ int tap_delta(int t, int W, int bpp) {
  const int dh = t / 3 - 1, dw = t % 3 - 1;
  return (MIRROR ? -(dh * W + dw) : dh * W + dw) * bpp;
}
__device__ __forceinline__ unsigned tap_off(unsigned mask, int t, unsigned base, int delta) {
  return (mask >> t) & 1u ? base + static_cast<unsigned>(delta) : kDrop;
}
__device__ __forceinline__ float bf16_bits(unsigned short u) {
  return __builtin_bit_cast(float, static_cast<unsigned>(u) << 16);
}

template <int N, typename T>
__device__ __forceinline__ void lds_fill(float *dst, const T *src) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = static_cast<float>(src[i]);
}

// Forward of a CI -> 64 layer with GELU: 8 lanes per pixel, 8 output
// channels per lane.  Input and weights are FP32 (the normalised state and
// the master weights): small differences of the input survive into the
// first features, which BF16 then rounds relative to their own size.
template <int CI>
__global__ __launch_bounds__(256) void k_in_fwd(SmallArgs a) {
  __shared__ float wl[9 * CI * 64];
  __shared__ float bl[64];
  lds_fill<9 * CI * 64>(wl, a.w32);
  if (threadIdx.x < 64) bl[threadIdx.x] = a.bias[threadIdx.x];
  __syncthreads();
  constexpr int NP = CI == 1 ? 2 : 1;  // pixels per iteration
  const int j = threadIdx.x & 7;
  const int P = a.B * a.H * a.W;
  const __amdgpu_buffer_rsrc_t r_in = conv_rsrc(a.in32, P * CI * 4);
  const __amdgpu_buffer_rsrc_t r_z = conv_rsrc(a.z_out, a.z_out ? P * 128 : 0);
  const __amdgpu_buffer_rsrc_t r_a = conv_rsrc(a.a_out, P * 128);
  const int stride = gridDim.x * 32;
  for (int p0 = (blockIdx.x * 256 + threadIdx.x) >> 3; p0 < P; p0 += NP * stride) {
    float xv[NP][9][CI];
    unsigned obase[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const int p = p0 + s * stride;
      const bool in = p < P;
      const int w = p % a.W, h = (p / a.W) % a.H;
      const unsigned mask = in ? tap_mask<false>(h, w, a.H, a.W) : 0u;
      obase[s] = in ? static_cast<unsigned>(p) * 128 + 16 * j : kDrop;
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const unsigned off = tap_off(mask, t, static_cast<unsigned>(p) * CI * 4,
                                     tap_delta<false>(t, a.W, CI * 4));
#pragma unroll
        for (int c = 0; c < CI; ++c)
          xv[s][t][c] = __builtin_bit_cast(
              float, __builtin_amdgcn_raw_buffer_load_b32(r_in, off + 4 * c, 0, 0));
      }
    }
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      float acc[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = bl[8 * j + k];
#pragma unroll
      for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int c = 0; c < CI; ++c) {
          const float *wt = wl + (t * CI + c) * 64 + 8 * j;
#pragma unroll
          for (int k = 0; k < 8; ++k) acc[k] += xv[s][t][c] * wt[k];
        }
      bf16x8 zv, av;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        zv[k] = static_cast<bf16>(acc[k]);
        av[k] = static_cast<bf16>(gelu_f(acc[k]));
      }
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, zv), r_z, obase[s], 0, 0);
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, av), r_a, obase[s], 0, 0);
    }
  }
}

// Forward of a 64 -> CO layer (no activation, FP32 output): 8 lanes per
// pixel, each summing its 8 input channels over the 9 taps; the partial
// sums meet through lane shuffles.
template <int CO>
__global__ __launch_bounds__(256) void k_out_fwd(SmallArgs a) {
  __shared__ float wl[9 * 64 * CO];
  lds_fill<9 * 64 * CO>(wl, a.w);
  __syncthreads();
  constexpr int NP = CO == 1 ? 2 : 1;  // pixels per iteration
  const int j = threadIdx.x & 7;
  const int P = a.B * a.H * a.W;
  const __amdgpu_buffer_rsrc_t r_in = conv_rsrc(a.in, P * 128);
  const int stride = gridDim.x * 32;
  for (int p0 = (blockIdx.x * 256 + threadIdx.x) >> 3; p0 < P; p0 += NP * stride) {
    bf16x8 v[NP][9];
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const int p = p0 + s * stride;
      const bool in = p < P;
      const int w = p % a.W, h = (p / a.W) % a.H;
      const unsigned mask = in ? tap_mask<false>(h, w, a.H, a.W) : 0u;
#pragma unroll
      for (int t = 0; t < 9; ++t)
        v[s][t] = __builtin_bit_cast(
            bf16x8, __builtin_amdgcn_raw_buffer_load_b128(
                        r_in, tap_off(mask, t, static_cast<unsigned>(p) * 128 + 16 * j,
                                      tap_delta<false>(t, a.W, 128)),
                        0, 0));
    }
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const int p = p0 + s * stride;
      float acc[CO];
#pragma unroll
      for (int o = 0; o < CO; ++o) acc[o] = 0.0f;
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const float *wt = wl + (t * 64 + 8 * j) * CO;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const float xv = static_cast<float>(v[s][t][k]);
#pragma unroll
          for (int o = 0; o < CO; ++o) acc[o] += xv * wt[k * CO + o];
        }
      }
#pragma unroll
      for (int o = 0; o < CO; ++o) {
        acc[o] += __shfl_xor(acc[o], 1);
        acc[o] += __shfl_xor(acc[o], 2);
        acc[o] += __shfl_xor(acc[o], 4);
      }
      if (j == 0 && p < P) {
#pragma unroll
        for (int o = 0; o < CO; ++o)
          a.y_out[static_cast<int64_t>(p) * CO + o] = acc[o] + a.bias[o];
      }
    }
  }
}

// Input gradient of a 64 -> CO layer, times GELU' of the input's
// pre-activation: 8 lanes per pixel, 8 channels per lane.
template <int CO>
__global__ __launch_bounds__(256) void k_out_bwd(SmallArgs a) {
  __shared__ float wl[9 * 64 * CO];
  lds_fill<9 * 64 * CO>(wl, a.w);
  __syncthreads();
  constexpr int NP = CO == 1 ? 2 : 1;  // pixels per iteration
  const int j = threadIdx.x & 7;
  const int P = a.B * a.H * a.W;
  const __amdgpu_buffer_rsrc_t r_g = conv_rsrc(a.in, P * CO * 2);
  const __amdgpu_buffer_rsrc_t r_zp = conv_rsrc(a.zprev, P * 128);
  const __amdgpu_buffer_rsrc_t r_out = conv_rsrc(a.a_out, P * 128);
  const int stride = gridDim.x * 32;
  for (int p0 = (blockIdx.x * 256 + threadIdx.x) >> 3; p0 < P; p0 += NP * stride) {
    float gv[NP][9][CO];
    bf16x8 zp[NP];
    unsigned obase[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      const int p = p0 + s * stride;
      const bool in = p < P;
      const int w = p % a.W, h = (p / a.W) % a.H;
      // input (h, w) feeds output (h - th + 1, w - tw + 1)
      const unsigned mask = in ? tap_mask<true>(h, w, a.H, a.W) : 0u;
      obase[s] = in ? static_cast<unsigned>(p) * 128 + 16 * j : kDrop;
      zp[s] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r_zp, obase[s], 0, 0));
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const unsigned off = tap_off(mask, t, static_cast<unsigned>(p) * CO * 2,
                                     tap_delta<true>(t, a.W, CO * 2));
#pragma unroll
        for (int o = 0; o < CO; ++o)
          gv[s][t][o] = bf16_bits(__builtin_amdgcn_raw_buffer_load_b16(r_g, off + 2 * o, 0, 0));
      }
    }
#pragma unroll
    for (int s = 0; s < NP; ++s) {
      float acc[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = 0.0f;
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const float *wt = wl + (t * 64 + 8 * j) * CO;
#pragma unroll
        for (int o = 0; o < CO; ++o)
#pragma unroll
          for (int k = 0; k < 8; ++k) acc[k] += gv[s][t][o] * wt[k * CO + o];
      }
      bf16x8 dv;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        dv[k] = static_cast<bf16>(acc[k] * gelu_grad_f(static_cast<float>(zp[s][k])));
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, dv), r_out, obase[s], 0, 0);
    }
  }
}

// Sums NV per-lane partials over the four pixel slots of each wavefront
// (lanes l, l + 16, l + 32, l + 48 own the same channels) and the four
// wavefronts of the block; returns with red[i][w][l] (l < 16) holding the
// partial sums of wavefront w.
template <int NV>
__device__ __forceinline__ void block_reduce16(float (&v)[NV], float (*red)[4][16]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    v[i] += __shfl_xor(v[i], 16);
    v[i] += __shfl_xor(v[i], 32);
  }
  if (lane < 16) {
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i][wave][lane] = v[i];
  }
  __syncthreads();
}

// Weight and bias gradients of a CI -> 64 layer: dw[t][ci][co] and db[co]
// summed over pixels.  16 lanes per pixel, 4 output channels per lane.
template <int CI>
__global__ __launch_bounds__(256) void k_in_wgrad(const float *x, const bf16 *dz, float *dw,
                                                  float *db, int B, int H, int W, float *part) {
  constexpr int NA = 9 * CI * 4, NV = NA + 4;
  __shared__ float red[NV][4][16];
  const int j = threadIdx.x & 15;
  float v[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = 0.0f;
  const int P = B * H * W;
  const __amdgpu_buffer_rsrc_t r_x = conv_rsrc(x, P * CI * 4);
  const __amdgpu_buffer_rsrc_t r_dz = conv_rsrc(dz, P * 128);
  const int stride = gridDim.x * 16;
  for (int p0 = (blockIdx.x * 256 + threadIdx.x) >> 4; p0 < P; p0 += 2 * stride) {
    float g[2][4], xv[2][9][CI];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int p = p0 + s * stride;
      const bool in = p < P;
      const int w = p % W, h = (p / W) % H;
      const unsigned mask = in ? tap_mask<false>(h, w, H, W) : 0u;
      const bf16x4 gv = __builtin_bit_cast(
          bf16x4, __builtin_amdgcn_raw_buffer_load_b64(
                      r_dz, in ? static_cast<unsigned>(p) * 128 + 8 * j : kDrop, 0, 0));
#pragma unroll
      for (int k = 0; k < 4; ++k) g[s][k] = static_cast<float>(gv[k]);
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const unsigned off = tap_off(mask, t, static_cast<unsigned>(p) * CI * 4,
                                     tap_delta<false>(t, W, CI * 4));
#pragma unroll
        for (int c = 0; c < CI; ++c)
          xv[s][t][c] = __builtin_bit_cast(
              float, __builtin_amdgcn_raw_buffer_load_b32(r_x, off + 4 * c, 0, 0));
      }
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
      for (int k = 0; k < 4; ++k) v[NA + k] += g[s][k];
#pragma unroll
      for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int c = 0; c < CI; ++c)
#pragma unroll
          for (int k = 0; k < 4; ++k) v[(t * CI + c) * 4 + k] += xv[s][t][c] * g[s][k];
    }
  }
  block_reduce16<NV>(v, red);
  // per-block partials: dw [9 CI 64], then db [64]
  float *pb = part ? part + static_cast<size_t>(blockIdx.x) * (9 * CI * 64 + 64) : nullptr;
  for (int e = threadIdx.x; e < NV * 16; e += 256) {
    const int i = e >> 4, l = e & 15;
    const float s = red[i][0][l] + red[i][1][l] + red[i][2][l] + red[i][3][l];
    const int idx = i < NA ? (i >> 2) * 64 + 4 * l + (i & 3)  // [t][ci][co = 4l + k]
                           : 9 * CI * 64 + 4 * l + (i - NA);
    if (pb) pb[idx] = s;
    else if (i < NA) atomicAdd(&dw[idx], s);
    else atomicAdd(&db[idx - 9 * CI * 64], s);
  }
}

// Weight and bias gradients of a 64 -> CO layer: dw[t][ci][o] and db[o].
// 16 lanes per pixel, 4 input channels per lane.
template <int CO>
__global__ __launch_bounds__(256) void k_out_wgrad(const bf16 *x, const bf16 *dz, float *dw,
                                                   float *db, int B, int H, int W, float *part) {
  constexpr int NA = 9 * 4 * CO, NV = NA + CO;
  __shared__ float red[NV][4][16];
  const int j = threadIdx.x & 15;
  float v[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = 0.0f;
  const int P = B * H * W;
  const __amdgpu_buffer_rsrc_t r_x = conv_rsrc(x, P * 128);
  const __amdgpu_buffer_rsrc_t r_g = conv_rsrc(dz, P * CO * 2);
  const int stride = gridDim.x * 16;
  for (int p0 = (blockIdx.x * 256 + threadIdx.x) >> 4; p0 < P; p0 += 2 * stride) {
    float g[2][CO];
    bf16x4 xv[2][9];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int p = p0 + s * stride;
      const bool in = p < P;
      const int w = p % W, h = (p / W) % H;
      const unsigned mask = in ? tap_mask<false>(h, w, H, W) : 0u;
#pragma unroll
      for (int o = 0; o < CO; ++o)
        g[s][o] = bf16_bits(__builtin_amdgcn_raw_buffer_load_b16(
            r_g, in ? static_cast<unsigned>(p) * CO * 2 + 2 * o : kDrop, 0, 0));
#pragma unroll
      for (int t = 0; t < 9; ++t)
        xv[s][t] = __builtin_bit_cast(
            bf16x4, __builtin_amdgcn_raw_buffer_load_b64(
                        r_x, tap_off(mask, t, static_cast<unsigned>(p) * 128 + 8 * j,
                                     tap_delta<false>(t, W, 128)),
                        0, 0));
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (j == 0) {
#pragma unroll
        for (int o = 0; o < CO; ++o) v[NA + o] += g[s][o];
      }
#pragma unroll
      for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float xk = static_cast<float>(xv[s][t][k]);
#pragma unroll
          for (int o = 0; o < CO; ++o) v[(t * 4 + k) * CO + o] += xk * g[s][o];
        }
    }
  }
  block_reduce16<NV>(v, red);
  // per-block partials: dw [9 64 CO], then db [CO]
  float *pb = part ? part + static_cast<size_t>(blockIdx.x) * (9 * 64 * CO + CO) : nullptr;
  for (int e = threadIdx.x; e < NV * 16; e += 256) {
    const int i = e >> 4, l = e & 15;
    const float s = red[i][0][l] + red[i][1][l] + red[i][2][l] + red[i][3][l];
    if (i < NA) {
      const int t = i / (4 * CO), k = (i / CO) & 3, o = i % CO;
      const int idx = (t * 64 + 4 * l + k) * CO + o;  // [t][ci = 4l + k][o]
      if (pb) pb[idx] = s;
      else atomicAdd(&dw[idx], s);
    } else if (l == 0) {
      if (pb) pb[9 * 64 * CO + (i - NA)] = s;
      else atomicAdd(&db[i - NA], s);
    }
  }
}

// Bias gradient of a 64-channel layer: db[c] += sum over pixels of dz[p][c].
// 8 lanes per pixel, 8 channels per lane, two pixels per iteration.
__global__ __launch_bounds__(256) void k_bias_grad64(const bf16 *dz, float *db, int64_t P) {
  const int j = threadIdx.x & 7, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float s[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = 0.0f;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 32;
  for (int64_t p0 = (blockIdx.x * 256 + threadIdx.x) >> 3; p0 < P; p0 += 2 * stride) {
    const int64_t p1 = p0 + stride;
    const bf16x8 g0 = *reinterpret_cast<const bf16x8 *>(dz + p0 * 64 + 8 * j);
    const bf16x8 g1 = *reinterpret_cast<const bf16x8 *>(dz + (p1 < P ? p1 : p0) * 64 + 8 * j);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      s[k] += static_cast<float>(g0[k]) + (p1 < P ? static_cast<float>(g1[k]) : 0.0f);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    s[k] += __shfl_xor(s[k], 8);
    s[k] += __shfl_xor(s[k], 16);
    s[k] += __shfl_xor(s[k], 32);
  }
  __shared__ float red[4][64];
  if (lane < 8) {
#pragma unroll
    for (int k = 0; k < 8; ++k) red[wave][8 * lane + k] = s[k];
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int c = threadIdx.x;
    atomicAdd(&db[c], red[0][c] + red[1][c] + red[2][c] + red[3][c]);
  }
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
// Per-sample normalisation x = u / s, s = max |u| (at least 1e-30), field-
// major u -> NHWC FP32 x, in three launches over the whole batch: clear s,
// per-block maxima merged with an integer atomicMax on the float bits
// (non-negative floats order like their bit patterns; fmaxf drops NaNs),
// then the scaling pass.
__global__ void k_norm_clear(float *s, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < B) s[b] = 0.0f;
}

template <typename T>
__global__ __launch_bounds__(256) void k_norm_max(const T *u, float *s, int C, int64_t HW) {
  const int b = blockIdx.y;
  const int64_t n = C * HW;
  const T *ub = u + static_cast<int64_t>(b) * n;
  float m = 0.0f;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256)
    m = fmaxf(m, fabsf(static_cast<float>(ub[i])));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    atomicMax(reinterpret_cast<unsigned *>(s) + b, __float_as_uint(m));
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_norm_scale(const T *u, float *x, float *s, int C,
                                                    int64_t HW) {
  const int b = blockIdx.y;
  const float sc = fmaxf(s[b], 1e-30f);
  // Every block stores the same clamped value, so the order of these
  // stores and the other blocks' reads does not matter.
  if (blockIdx.x == 0 && threadIdx.x == 0) s[b] = sc;
  const int64_t n = C * HW;
  const T *ub = u + static_cast<int64_t>(b) * n;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * 256) {
    const int c = static_cast<int>(i / HW);
    const int64_t px = i - c * HW;
    x[(static_cast<int64_t>(b) * HW + px) * C + c] = static_cast<float>(ub[i]) / sc;
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

extern "C" int pml_conv64(int mode, const void *in, const void *wk, const float *bias,
                          const void *zprev, void *z_out, void *a_out, int B, int H, int W,
                          void *stream) {
  if (B <= 0 || H <= 0 || W <= 0 || !in || !wk || !a_out)
    return fail(PML_E_INVALID, "bad conv64 arguments");
  if (mode == kFwd && !bias) return fail(PML_E_INVALID, "forward needs a bias");
  if (mode == kBwd && !zprev) return fail(PML_E_INVALID, "backward needs the pre-activation");
  if (mode != kFwd && mode != kBwd)
    return fail(PML_E_INVALID, "mode must be 0 (forward) or 1 (input gradient)");
  // Each launch covers whole images whose tensors stay below 2^31 bytes,
  // the range of the kernel's 32-bit buffer offsets.
  const int64_t img_elems = static_cast<int64_t>(H) * W * 64;
  const int64_t img_bytes = img_elems * 2;
  if (img_bytes >= (int64_t(1) << 31)) return fail(PML_E_INVALID, "image too large for conv64");
  const int per = static_cast<int>(std::max<int64_t>(1, ((int64_t(1) << 31) - 1) / img_bytes));
  const void *kern = mode == kFwd ? reinterpret_cast<const void *>(&k_conv64<kFwd>)
                                  : reinterpret_cast<const void *>(&k_conv64<kBwd>);
  PML_HIP_CHECK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, kConvLds));
  hipStream_t st = as_stream(stream);
  for (int b0 = 0; b0 < B; b0 += per) {
    const int nb = std::min(per, B - b0);
    const int64_t off = b0 * img_elems;
    ConvArgs a{};
    a.in = static_cast<const bf16 *>(in) + off;
    a.wk = static_cast<const bf16 *>(wk);
    a.bias = bias;
    a.zprev = zprev ? static_cast<const bf16 *>(zprev) + off : nullptr;
    a.z_out = z_out ? static_cast<bf16 *>(z_out) + off : nullptr;
    a.a_out = static_cast<bf16 *>(a_out) + off;
    a.B = nb;
    a.H = H;
    a.W = W;
    a.tiles_h = (H + kTR - 1) / kTR;
    a.tiles_w = (W + kTC - 1) / kTC;
    a.n_tiles = nb * a.tiles_h * a.tiles_w;
    a.bytes = static_cast<unsigned>(nb * img_bytes);
    const int grid = std::min(a.n_tiles, cu_count());
    if (mode == kFwd)
      hipLaunchKernelGGL(k_conv64<kFwd>, dim3(grid), dim3(kConvThreads), kConvLds, st, a);
    else
      hipLaunchKernelGGL(k_conv64<kBwd>, dim3(grid), dim3(kConvThreads), kConvLds, st, a);
    PML_LAUNCH_CHECK();
  }
  return PML_OK;
}

// Floats of partial-sum buffer a launch of n_blocks workgroups of
// per_block sums needs: the partials, then one row per segment.
static int64_t partial_floats(int n_blocks, int per_block) {
  const int64_t segs = (n_blocks + kSumSeg - 1) / kSumSeg;
  return (static_cast<int64_t>(n_blocks) + segs) * per_block;
}

// Adds a launch's per-block partials (`n_blocks` blocks of `per_block`
// floats: n_dw weight entries, then the bias entries when db is set) to dw
// and db in a fixed order.
static int sum_blocks(float *part, int n_blocks, int per_block, float *dw, int n_dw, float *db,
                      hipStream_t st) {
  const int n = db ? per_block : n_dw;
  const int segs = (n_blocks + kSumSeg - 1) / kSumSeg;
  float *seg = part + static_cast<size_t>(n_blocks) * per_block;
  hipLaunchKernelGGL(k_sum_segments, dim3((n + 255) / 256, segs), dim3(256), 0, st, part,
                     n_blocks, per_block, n, seg);
  PML_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_blocks, dim3((n + 255) / 256), dim3(256), 0, st, seg, segs, n, dw,
                     n_dw, db);
  PML_LAUNCH_CHECK();
  return PML_OK;
}

// Blocks of one pml_wgrad64 launch over nb images of H x W.
static int wgrad64_blocks(int nb, int H, int W) {
  const int n_tiles = nb * ((H + kGR - 1) / kGR) * ((W + kGC - 1) / kGC);
  return std::min(n_tiles, cu_count());
}

extern "C" int pml_wgrad64(const void *x, const void *dz, float *dw, float *db, int B, int H,
                           int W, float *part, int64_t part_floats, void *stream) {
  if (B <= 0 || H <= 0 || W <= 0 || !x || !dz || !dw)
    return fail(PML_E_INVALID, "bad wgrad64 arguments");
  // Each launch covers whole images whose tensors stay below 2^31 bytes.
  const int64_t img_elems = static_cast<int64_t>(H) * W * 64;
  const int64_t img_bytes = img_elems * 2;
  if (img_bytes >= (int64_t(1) << 31)) return fail(PML_E_INVALID, "image too large for wgrad64");
  const int per = static_cast<int>(std::max<int64_t>(1, ((int64_t(1) << 31) - 1) / img_bytes));
  PML_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_wgrad64),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kWgradLds));
  for (int b0 = 0; b0 < B; b0 += per) {
    const int nb = std::min(per, B - b0);
    WgradArgs a{};
    a.x = static_cast<const bf16 *>(x) + b0 * img_elems;
    a.dz = static_cast<const bf16 *>(dz) + b0 * img_elems;
    a.dw = dw;
    a.db = db;
    a.B = nb;
    a.H = H;
    a.W = W;
    a.tiles_h = (H + kGR - 1) / kGR;
    a.tiles_w = (W + kGC - 1) / kGC;
    a.n_tiles = nb * a.tiles_h * a.tiles_w;
    a.bytes = static_cast<unsigned>(nb * img_bytes);
    const int blocks = wgrad64_blocks(nb, H, W);
    if (part && part_floats < partial_floats(blocks, kWgradPart))
      return fail(PML_E_INVALID, "partial-sum buffer too small (see pml_grad_partials_floats)");
    a.part = part;
    hipLaunchKernelGGL(k_wgrad64, dim3(blocks), dim3(kWgradThreads), kWgradLds,
                       as_stream(stream), a);
    PML_LAUNCH_CHECK();
    if (part) {
      const int rc = sum_blocks(part, blocks, kWgradPart, dw, kWgradDw, db, as_stream(stream));
      if (rc) return rc;
    }
  }
  return PML_OK;
}

// Blocks for a grid-stride loop that covers `units` lane groups of
// `per_block` per block, capped at `per_cu` blocks per CU.
static int small_grid(int64_t units, int per_block, int per_cu) {
  const int64_t need = (units + per_block - 1) / per_block;
  const int64_t cap = static_cast<int64_t>(per_cu) * cu_count();
  return static_cast<int>(std::max<int64_t>(1, std::min(need, cap)));
}

// Images per launch of the small-layer kernels: their 64-channel tensors
// stay below 2^31 bytes, the range of the 32-bit buffer offsets.
static int small_chunk(int H, int W) {
  const int64_t img_bytes = static_cast<int64_t>(H) * W * 128;
  return static_cast<int>(std::max<int64_t>(1, ((int64_t(1) << 31) - 1) / img_bytes));
}

extern "C" int pml_conv_small(int mode, const void *in, const void *w, const float *bias,
                              const void *zprev, void *z_out, void *a_out, float *y_out, int Ci,
                              int Co, int B, int H, int W, void *stream) {
  if (B <= 0 || H <= 0 || W <= 0 || !in || !w) return fail(PML_E_INVALID, "bad conv arguments");
  if (static_cast<int64_t>(H) * W * 128 >= (int64_t(1) << 31))
    return fail(PML_E_INVALID, "image too large");
  const bool in_layer = Co == 64 && Ci >= 1 && Ci <= 3;
  const bool out_layer = Ci == 64 && Co >= 1 && Co <= 3;
  if (!in_layer && !out_layer) return fail(PML_E_INVALID, "small layers are C (1..3) -> 64 or 64 -> C");
  if (mode == kFwd) {
    if (!bias) return fail(PML_E_INVALID, "forward needs a bias");
    if (in_layer && !a_out) return fail(PML_E_INVALID, "missing activation output");
    if (out_layer && !y_out) return fail(PML_E_INVALID, "missing output");
  } else if (mode == kBwd) {
    // input gradient of a 64 -> Co layer (the caller passes Ci = 64)
    if (!out_layer || !zprev || !a_out) return fail(PML_E_INVALID, "bad input-gradient arguments");
  } else {
    return fail(PML_E_INVALID, "mode must be 0 or 1");
  }
  hipStream_t st = as_stream(stream);
  const int per = small_chunk(H, W);
  const int64_t hw = static_cast<int64_t>(H) * W;
  for (int b0 = 0; b0 < B; b0 += per) {
    const int nb = std::min(per, B - b0);
    const int64_t px = b0 * hw;  // first pixel of the chunk
    SmallArgs a{};
    if (in_layer) {
      a.in32 = static_cast<const float *>(in) + px * Ci;
      a.w32 = static_cast<const float *>(w);
    } else {
      a.in = static_cast<const bf16 *>(in) + px * Ci;
      a.w = static_cast<const bf16 *>(w);
    }
    a.bias = bias;
    a.zprev = zprev ? static_cast<const bf16 *>(zprev) + px * 64 : nullptr;
    a.z_out = z_out ? static_cast<bf16 *>(z_out) + px * 64 : nullptr;
    a.a_out = a_out ? static_cast<bf16 *>(a_out) + px * 64 : nullptr;
    a.y_out = y_out ? y_out + px * Co : nullptr;
    a.B = nb;
    a.H = H;
    a.W = W;
    a.Ci = Ci;
    a.Co = Co;
    // two pixels per lane group and iteration: 64 pixels per block
    const dim3 grid(small_grid(nb * hw, 64, 8)), block(256);
    if (mode == kFwd) {
      if (Ci == 1) hipLaunchKernelGGL(k_in_fwd<1>, grid, block, 0, st, a);
      else if (Ci == 2) hipLaunchKernelGGL(k_in_fwd<2>, grid, block, 0, st, a);
      else if (Ci == 3) hipLaunchKernelGGL(k_in_fwd<3>, grid, block, 0, st, a);
      else if (Co == 1) hipLaunchKernelGGL(k_out_fwd<1>, grid, block, 0, st, a);
      else if (Co == 2) hipLaunchKernelGGL(k_out_fwd<2>, grid, block, 0, st, a);
      else hipLaunchKernelGGL(k_out_fwd<3>, grid, block, 0, st, a);
    } else {
      if (Co == 1) hipLaunchKernelGGL(k_out_bwd<1>, grid, block, 0, st, a);
      else if (Co == 2) hipLaunchKernelGGL(k_out_bwd<2>, grid, block, 0, st, a);
      else hipLaunchKernelGGL(k_out_bwd<3>, grid, block, 0, st, a);
    }
    PML_LAUNCH_CHECK();
  }
  return PML_OK;
}

extern "C" int pml_conv_small_wgrad(const void *x, const void *dz, float *dw, float *db, int Ci,
                                    int Co, int B, int H, int W, float *part,
                                    int64_t part_floats, void *stream) {
  if (B <= 0 || H <= 0 || W <= 0 || !x || !dz || !dw || !db)
    return fail(PML_E_INVALID, "bad weight-gradient arguments");
  if (static_cast<int64_t>(H) * W * 128 >= (int64_t(1) << 31))
    return fail(PML_E_INVALID, "image too large");
  const bool in_layer = Co == 64 && Ci >= 1 && Ci <= 3;
  const bool out_layer = Ci == 64 && Co >= 1 && Co <= 3;
  if (!in_layer && !out_layer) return fail(PML_E_INVALID, "small layers are C (1..3) -> 64 or 64 -> C");
  hipStream_t st = as_stream(stream);
  const int per = small_chunk(H, W);
  const int64_t hw = static_cast<int64_t>(H) * W;
  for (int b0 = 0; b0 < B; b0 += per) {
    const int nb = std::min(per, B - b0);
    const int64_t px = b0 * hw;
    const bf16 *xb = static_cast<const bf16 *>(x) + px * Ci;
    const float *xf = static_cast<const float *>(x) + px * Ci;  // Ci -> 64 layers read FP32
    const bf16 *dzb = static_cast<const bf16 *>(dz) + px * Co;
    // two pixels per lane group and iteration: 32 pixels per block
    const dim3 grid(small_grid(nb * hw, 32, 4)), block(256);
    const int n_dw = 9 * Ci * Co, per_block = n_dw + Co;
    if (part && part_floats < partial_floats(static_cast<int>(grid.x), per_block))
      return fail(PML_E_INVALID, "partial-sum buffer too small (see pml_grad_partials_floats)");
    if (Ci == 1) hipLaunchKernelGGL(k_in_wgrad<1>, grid, block, 0, st, xf, dzb, dw, db, nb, H, W, part);
    else if (Ci == 2) hipLaunchKernelGGL(k_in_wgrad<2>, grid, block, 0, st, xf, dzb, dw, db, nb, H, W, part);
    else if (Ci == 3) hipLaunchKernelGGL(k_in_wgrad<3>, grid, block, 0, st, xf, dzb, dw, db, nb, H, W, part);
    else if (Co == 1) hipLaunchKernelGGL(k_out_wgrad<1>, grid, block, 0, st, xb, dzb, dw, db, nb, H, W, part);
    else if (Co == 2) hipLaunchKernelGGL(k_out_wgrad<2>, grid, block, 0, st, xb, dzb, dw, db, nb, H, W, part);
    else hipLaunchKernelGGL(k_out_wgrad<3>, grid, block, 0, st, xb, dzb, dw, db, nb, H, W, part);
    PML_LAUNCH_CHECK();
    if (part) {
      const int rc = sum_blocks(part, static_cast<int>(grid.x), per_block, dw, n_dw, db, st);
      if (rc) return rc;
    }
  }
  return PML_OK;
}

extern "C" int pml_grad_partials_floats(int Ci, int Co, int B, int H, int W, int64_t *floats) {
  if (B <= 0 || H <= 0 || W <= 0 || !floats) return fail(PML_E_INVALID, "bad arguments");
  if (Ci == 64 && Co == 64) {
    const int64_t img_bytes = static_cast<int64_t>(H) * W * 128;
    if (img_bytes >= (int64_t(1) << 31)) return fail(PML_E_INVALID, "image too large");
    const int per = static_cast<int>(std::max<int64_t>(1, ((int64_t(1) << 31) - 1) / img_bytes));
    *floats = partial_floats(wgrad64_blocks(std::min(per, B), H, W), kWgradPart);
    return PML_OK;
  }
  const bool in_layer = Co == 64 && Ci >= 1 && Ci <= 3;
  const bool out_layer = Ci == 64 && Co >= 1 && Co <= 3;
  if (!in_layer && !out_layer) return fail(PML_E_INVALID, "unsupported layer shape");
  const int per = small_chunk(H, W);
  const int64_t hw = static_cast<int64_t>(H) * W;
  *floats = partial_floats(small_grid(std::min(per, B) * hw, 32, 4), 9 * Ci * Co + Co);
  return PML_OK;
}

extern "C" int pml_bias_grad64(const void *dz, float *db, int64_t pixels, void *stream) {
  if (pixels <= 0 || !dz || !db) return fail(PML_E_INVALID, "bad bias-gradient arguments");
  hipLaunchKernelGGL(k_bias_grad64, dim3(small_grid(pixels, 64, 4)), dim3(256), 0,
                     as_stream(stream), static_cast<const bf16 *>(dz), db, pixels);
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
  if (precision != PML_F64 && precision != PML_F32)
    return fail(PML_E_INVALID, "normalisation reads FP32 or FP64");
  if (B > 65535) return fail(PML_E_INVALID, "too many samples");
  hipStream_t st = as_stream(stream);
  // about 8 elements per thread, spread over the whole chip
  const int64_t n = static_cast<int64_t>(C) * HW;
  const int per_sample = static_cast<int>(
      std::max<int64_t>(1, std::min<int64_t>((n + 2047) / 2048, (8LL * cu_count() + B - 1) / B)));
  const dim3 grid(per_sample, B);
  hipLaunchKernelGGL(k_norm_clear, dim3((B + 255) / 256), dim3(256), 0, st, s, B);
  if (precision == PML_F64) {
    const double *ud = static_cast<const double *>(u);
    hipLaunchKernelGGL(k_norm_max<double>, grid, dim3(256), 0, st, ud, s, C, HW);
    hipLaunchKernelGGL(k_norm_scale<double>, grid, dim3(256), 0, st, ud, static_cast<float *>(x), s,
                       C, HW);
  } else {
    const float *uf = static_cast<const float *>(u);
    hipLaunchKernelGGL(k_norm_max<float>, grid, dim3(256), 0, st, uf, s, C, HW);
    hipLaunchKernelGGL(k_norm_scale<float>, grid, dim3(256), 0, st, uf, static_cast<float *>(x), s,
                       C, HW);
  }
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
