// mlp.hip -- BF16 MFMA perceptron for the chemistry surrogate: training GEMMs
// with fused epilogues, a fused all-layer inference kernel, loss and Adam.
//
// MFMA: v_mfma_f32_32x32x16_bf16 (gfx950).  Lane l of a wave holds
//   A fragment: A[row = l%32][k = 8*(l/32) + v], v = 0..7
//   B fragment: B[k = 8*(l/32) + v][col = l%32]
//   accumulator r = 0..15: C[row = 8*(r/4) + 4*(l/32) + r%4][col = l%32]
// Everything is padded so that GEMM dimensions are multiples of 32 (M, N)
// and 16 (K): the input to 32 features, the output to 32 columns; padded
// weights and biases are zero and stay zero under training (zero gradients).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/nrg_mlp.h"
#include "mlp_internal.h"
#include "nrg_common.h"

namespace nrg {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int WAVE = 64;

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c)
{
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// GELU, tanh form (Hendrycks & Gimpel 2016):
//   g(z) = z/2 (1 + tanh(u)),  u = sqrt(2/pi) (z + 0.044715 z^3)
// evaluated as z * sigma(2u) (1/2 (1 + tanh(u)) = 1 / (1 + e^(-2u))): one
// exp2 and one reciprocal, with -2 sqrt(2/pi) log2(e) folded into the cubic.
__device__ __forceinline__ float sigma2u(float z)
{
  const float a0 = -2.0f * 0.7978845608028654f * 1.4426950408889634f;
  const float a1 = a0 * 0.044715f;
  const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(z * z, a1, a0));
  return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu(float z) { return z * sigma2u(z); }
__device__ __forceinline__ float gelu_grad(float z)
{
  const float c = 0.7978845608028654f, a = 0.044715f;
  const float sg = sigma2u(z);
  const float dudz = c * (1.0f + 3.0f * a * z * z);
  return sg + 2.0f * z * sg * (1.0f - sg) * dudz;
}

// ---- training GEMM: one wave per 32x32 tile of C, optional split over K ----------
enum { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_BACK = 3 };

struct GemmArgs {
  const bf16 *A;
  long long sAm, sAk;          // A(m, k) = A[m*sAm + k*sAk]
  const bf16 *B;
  long long sBk, sBn;          // B(k, n) = B[k*sBk + n*sBn]
  int M, N, K;
  int mode;                    // EPI_*
  int ksplit;                  // EPI_STORE only: K ranges per tile, atomically added
  long long ldc;
  float *C;                    // EPI_STORE / EPI_BIAS out (EPI_STORE with ksplit > 1: pre-zeroed)
  const float *bias;           // EPI_BIAS_GELU / EPI_BIAS
  float *Z;                    // EPI_BIAS_GELU: pre-activation
  bf16 *H;                     // EPI_BIAS_GELU: gelu(z); EPI_GELU_BACK: acc * gelu'(Zin)
  const float *Zin;            // EPI_GELU_BACK
};

__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g)
{
  const int l = threadIdx.x & (WAVE - 1);
  const int tiles_n = g.N / 32;
  const int ntiles = (g.M / 32) * tiles_n;
  const long long wid = (long long)blockIdx.x * 4 + threadIdx.x / WAVE;
  if (wid >= (long long)ntiles * g.ksplit) return;
  const int tile = (int)(wid % ntiles), part = (int)(wid / ntiles);
  const int m0 = (tile / tiles_n) * 32, n0 = (tile % tiles_n) * 32;
  const int r32 = l & 31, hi = l >> 5;
  const int kchunk = ((g.K / 16 + g.ksplit - 1) / g.ksplit) * 16;
  const int kbeg = part * kchunk, kend = min(g.K, kbeg + kchunk);

  f32x16 acc = {};
  const bf16 *Ap = g.A + (long long)(m0 + r32) * g.sAm;
  const bf16 *Bp = g.B + (long long)(n0 + r32) * g.sBn;
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    bf16x8 a, b;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const long long k = k0 + 8 * hi + v;
      a[v] = Ap[k * g.sAk];
      b[v] = Bp[k * g.sBk];
    }
    acc = mfma(a, b, acc);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long long row = m0 + 8 * (r / 4) + 4 * hi + (r % 4);
    const int col = n0 + r32;
    const long long o = row * g.ldc + col;
    if (g.mode == EPI_STORE) {
      if (g.ksplit > 1) atomicAdd(&g.C[o], acc[r]);
      else g.C[o] = acc[r];
    } else if (g.mode == EPI_BIAS_GELU) {
      const float z = acc[r] + g.bias[col];
      g.Z[o] = z;
      g.H[o] = (bf16)gelu(z);
    } else if (g.mode == EPI_BIAS) {
      g.C[o] = acc[r] + g.bias[col];
    } else {
      g.H[o] = (bf16)(acc[r] * gelu_grad(g.Zin[o]));
    }
  }
}

// ---- small element-wise kernels ---------------------------------------------------

__global__ void fill_f32_kernel(float *p, long long n, float v)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// x (n x n_in FP32) -> X (n x 32 bf16, zero padded)
__global__ void pack_input_kernel(const float *x, long long n, int n_in, bf16 *X)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * MLP_IN_PAD) return;
  const long long row = i / MLP_IN_PAD;
  const int c = (int)(i % MLP_IN_PAD);
  X[i] = (bf16)(c < n_in ? x[row * n_in + c] : 0.0f);
}

// loss = sum_rows sum_c w_c (y - t)^2 / n and dY = 2 w_c (y - t) / n (bf16, padded)
__global__ void loss_kernel(const float *Y, const float *T, const float *w, long long n, int n_out,
                            bf16 *dY, float *loss)
{
  const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.0f;
  if (row < n) {
    const float inv_n = 1.0f / (float)n;
    for (int c = 0; c < MLP_OUT_PAD; ++c) {
      float g = 0.0f;
      if (c < n_out) {
        const float d = Y[row * MLP_OUT_PAD + c] - T[row * n_out + c];
        const float wc = w ? w[c] : 1.0f;
        s += wc * d * d;
        g = 2.0f * wc * d * inv_n;
      }
      dY[row * MLP_OUT_PAD + c] = (bf16)g;
    }
    s *= inv_n;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (loss && (threadIdx.x & (WAVE - 1)) == 0) atomicAdd(loss, s);
}

// db[c] += sum over a 256-row slab of dZ[row][c]
__global__ void bias_grad_kernel(const bf16 *dZ, long long n, int N, float *db)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const long long r0 = (long long)blockIdx.y * 256;
  const long long r1 = r0 + 256 < n ? r0 + 256 : n;
  float s = 0.0f;
  for (long long r = r0; r < r1; ++r) s += (float)dZ[r * N + c];
  atomicAdd(&db[c], s);
}

__global__ void adam_kernel(float *P, bf16 *Pbf, const float *G, float *M1, float *M2, long long n,
                            float lr, float bc1, float bc2)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float g = G[i];
  const float m = b1 * M1[i] + (1.0f - b1) * g;
  const float v = b2 * M2[i] + (1.0f - b2) * g * g;
  M1[i] = m;
  M2[i] = v;
  const float w = P[i] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
  P[i] = w;
  Pbf[i] = (bf16)w;
}

__global__ void to_bf16_kernel(const float *P, bf16 *Pbf, long long n)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Pbf[i] = (bf16)P[i];
}

__global__ void gather_kernel(const float *src, const int *idx, float *dst, long long n)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// Inference layout of layer weights: Wsw[((kb*NT + nt)*64 + lane)*8 + v] =
// W[k = 16 kb + 8 (lane/32) + v][n = 32 nt + lane%32], so a lane's B fragment
// is one 16-byte load.
__global__ void swizzle_kernel(const bf16 *W, int K, int N, bf16 *Wsw)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)K * N) return;
  const int v = (int)(i % 8);
  const int lane = (int)((i / 8) % WAVE);
  const long long f = i / (8 * WAVE);
  const int NT = N / 32;
  const int nt = (int)(f % NT), kb = (int)(f / NT);
  const int k = 16 * kb + 8 * (lane / 32) + v, n = 32 * nt + lane % 32;
  Wsw[i] = W[(long long)k * N + n];
}

// ---- fused inference: all layers, 32 rows per wave, 128 rows per block -----------
// Hidden activations of a wave's rows stay in LDS as bf16 between layers;
// weights stream from L2 as pre-swizzled B fragments.
constexpr int ACT_LD = MLP_HIDDEN_MAX + 8;   // row pitch (bf16): 16 B skew against bank conflicts

struct InferArgs {
  const bf16 *X;               // n x 32
  long long n;
  int n_out;
  int hidden, layers;          // hidden width (multiple of 32, <= MLP_HIDDEN_MAX), hidden layers
  const bf16 *Wsw[MLP_MAX_LAYERS];
  const float *bias[MLP_MAX_LAYERS];
  float *Y;                    // n x n_out
};

template <int NT>
__device__ __forceinline__ void act_store(bf16 (*act)[ACT_LD], const f32x16 (&acc)[NT],
                                          const float *bias, int lane)
{
  const int r32 = lane & 31, hi = lane >> 5;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 8 * (r / 4) + 4 * hi + (r % 4), col = 32 * nt + r32;
      act[row][col] = (bf16)gelu(acc[nt][r] + bias[col]);
    }
}

template <int NT>
__global__ void __launch_bounds__(256) mlp_infer_kernel(InferArgs a)
{
  __shared__ bf16 act_all[4][32][ACT_LD];
  const int w = threadIdx.x / WAVE, l = threadIdx.x & (WAVE - 1);
  const int r32 = l & 31, hi = l >> 5;
  bf16 (*act)[ACT_LD] = act_all[w];
  const long long row0 = (long long)blockIdx.x * 128 + w * 32;
  const long long myrow = row0 + r32 < a.n ? row0 + r32 : a.n - 1;

  f32x16 acc[NT];
  // layer 0: K = 32 (two k-blocks) from the packed input
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x16{};
#pragma unroll
  for (int kb = 0; kb < MLP_IN_PAD / 16; ++kb) {
    const bf16x8 av = *reinterpret_cast<const bf16x8 *>(a.X + myrow * MLP_IN_PAD + 16 * kb + 8 * hi);
    const bf16x8 *Wb = reinterpret_cast<const bf16x8 *>(a.Wsw[0]) + (kb * NT) * WAVE + l;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = mfma(av, Wb[nt * WAVE], acc[nt]);
  }
  act_store<NT>(act, acc, a.bias[0], l);
  __syncthreads();

  // hidden layers: K = N = hidden
  for (int layer = 1; layer < a.layers; ++layer) {
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x16{};
    const bf16x8 *W = reinterpret_cast<const bf16x8 *>(a.Wsw[layer]);
    for (int kb = 0; kb < NT * 2; ++kb) {
      const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&act[r32][16 * kb + 8 * hi]);
      const bf16x8 *Wb = W + (kb * NT) * WAVE + l;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[nt] = mfma(av, Wb[nt * WAVE], acc[nt]);
    }
    __syncthreads();
    act_store<NT>(act, acc, a.bias[layer], l);
    __syncthreads();
  }

  // output layer: K = hidden, N = 32 (padded)
  f32x16 o = {};
  const bf16x8 *Wo = reinterpret_cast<const bf16x8 *>(a.Wsw[a.layers]);
  for (int kb = 0; kb < NT * 2; ++kb) {
    const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&act[r32][16 * kb + 8 * hi]);
    o = mfma(av, Wo[kb * WAVE + l], o);
  }
  const float *bo = a.bias[a.layers];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long long row = row0 + 8 * (r / 4) + 4 * hi + (r % 4);
    if (r32 < a.n_out && row < a.n) a.Y[row * a.n_out + r32] = o[r] + bo[r32];
  }
}

// Inference, hidden width 32*NT with staged weights: 64 rows per wave (two
// 32-row MFMA tiles share every B fragment), 256 rows per block, and each
// 16-row k-block of a hidden layer's weights copied once per block into a
// double-buffered LDS slot that the four waves read.  The per-tile
// accumulation order equals mlp_infer_kernel's, so the results are identical.
constexpr int ROWS2 = 64;

template <int NT>
__global__ void __launch_bounds__(256) mlp_infer2_kernel(InferArgs a)
{
  __shared__ bf16 act_all[4][ROWS2][ACT_LD];
  __shared__ bf16x8 wbuf[2][NT * WAVE];
  const int w = threadIdx.x / WAVE, l = threadIdx.x & (WAVE - 1);
  const int r32 = l & 31, hi = l >> 5;
  bf16 (*act)[ACT_LD] = act_all[w];
  const long long row0 = (long long)blockIdx.x * (4 * ROWS2) + w * ROWS2;
  const long long ra = row0 + r32 < a.n ? row0 + r32 : a.n - 1;
  const long long rb = row0 + 32 + r32 < a.n ? row0 + 32 + r32 : a.n - 1;

  f32x16 acc[2][NT];
  // layer 0: K = 32 from the packed input; its B fragments (16 KB) from L2
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) acc[0][nt] = acc[1][nt] = f32x16{};
#pragma unroll
  for (int kb = 0; kb < MLP_IN_PAD / 16; ++kb) {
    const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(a.X + ra * MLP_IN_PAD + 16 * kb + 8 * hi);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(a.X + rb * MLP_IN_PAD + 16 * kb + 8 * hi);
    const bf16x8 *Wb = reinterpret_cast<const bf16x8 *>(a.Wsw[0]) + (kb * NT) * WAVE + l;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const bf16x8 b = Wb[nt * WAVE];
      acc[0][nt] = mfma(a0, b, acc[0][nt]);
      acc[1][nt] = mfma(a1, b, acc[1][nt]);
    }
  }
  act_store<NT>(act, acc[0], a.bias[0], l);
  act_store<NT>(act + 32, acc[1], a.bias[0], l);
  __syncthreads();

  constexpr int NV = NT * WAVE;            // bf16x8 vectors per k-block
  static_assert(NV % 256 == 0, "k-block staging assumes whole vectors per thread");
  for (int layer = 1; layer < a.layers; ++layer) {
    const bf16x8 *W = reinterpret_cast<const bf16x8 *>(a.Wsw[layer]);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[0][nt] = acc[1][nt] = f32x16{};
#pragma unroll
    for (int v = 0; v < NV / 256; ++v) wbuf[0][v * 256 + threadIdx.x] = W[v * 256 + threadIdx.x];
    __syncthreads();
    for (int kb = 0; kb < NT * 2; ++kb) {
      bf16x8 nxt[NV / 256];
      const bool more = kb + 1 < NT * 2;
      if (more) {
#pragma unroll
        for (int v = 0; v < NV / 256; ++v) nxt[v] = W[(kb + 1) * NV + v * 256 + threadIdx.x];
      }
      const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(&act[r32][16 * kb + 8 * hi]);
      const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(&act[32 + r32][16 * kb + 8 * hi]);
      const bf16x8 *Wb = wbuf[kb & 1] + l;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const bf16x8 b = Wb[nt * WAVE];
        acc[0][nt] = mfma(a0, b, acc[0][nt]);
        acc[1][nt] = mfma(a1, b, acc[1][nt]);
      }
      if (more) {
#pragma unroll
        for (int v = 0; v < NV / 256; ++v) wbuf[(kb + 1) & 1][v * 256 + threadIdx.x] = nxt[v];
      }
      __syncthreads();
    }
    act_store<NT>(act, acc[0], a.bias[layer], l);
    act_store<NT>(act + 32, acc[1], a.bias[layer], l);
    __syncthreads();
  }

  // output layer: K = hidden, N = 32 (padded); its fragments (16 KB) from L2
  f32x16 o0 = {}, o1 = {};
  const bf16x8 *Wo = reinterpret_cast<const bf16x8 *>(a.Wsw[a.layers]);
  for (int kb = 0; kb < NT * 2; ++kb) {
    const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(&act[r32][16 * kb + 8 * hi]);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(&act[32 + r32][16 * kb + 8 * hi]);
    const bf16x8 b = Wo[kb * WAVE + l];
    o0 = mfma(a0, b, o0);
    o1 = mfma(a1, b, o1);
  }
  const float *bo = a.bias[a.layers];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long long row = row0 + 8 * (r / 4) + 4 * hi + (r % 4);
    if (r32 < a.n_out && row < a.n) a.Y[row * a.n_out + r32] = o0[r] + bo[r32];
    if (r32 < a.n_out && row + 32 < a.n) a.Y[(row + 32) * a.n_out + r32] = o1[r] + bo[r32];
  }
}

// ---- inference with the activations in registers ---------------------------------
// Each layer is computed transposed, Z^T = W^T H^T: the weights are the MFMA A
// operand (32 output features x 16 inputs) and the activations the B operand
// (16 inputs x 32 cells).  A lane's accumulator then holds, for its cell, the
// output features 32 mt + 8 (r/4) + 4 hi + r%4 (r = 0..15, hi = lane/32): after
// bias and GELU, elements r = 0..7 and 8..15 are exactly the lane's B fragments
// of k-blocks 2 mt and 2 mt + 1 of the next layer, provided that layer's
// weights are stored with their K rows in that order (swizzle_t_kernel).  So
// hidden activations never leave the registers; the hidden layers' weights
// (H x H bf16, 128 KB at H = 256) are staged in LDS once per layer and block.
#ifndef T_NC
#define T_NC 1                      // 32-cell tiles per wave (they share each A fragment)
#endif
#ifndef T_WPB
#define T_WPB 8                     // waves per block: two per SIMD, so one wave's GELUs
#endif                              // overlap the other's MFMAs

// physical feature (output column of the previous layer) of logical K row kl
__host__ __device__ inline int t_phys(int kl)
{
  const int kb = kl / 16, hi = (kl / 8) % 2, v = kl % 8;
  const int r = 8 * (kb % 2) + v;
  return 32 * (kb / 2) + 8 * (r / 4) + 4 * hi + r % 4;
}

// WswT[((mt*KB + kb)*64 + lane)*8 + v] = W[k][n], n = 32 mt + lane%32,
// k = (permuted ? t_phys : identity)(16 kb + 8 (lane/32) + v).
__global__ void swizzle_t_kernel(const bf16 *W, int K, int N, int permuted, bf16 *WswT)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)K * N) return;
  const int v = (int)(i % 8);
  const int lane = (int)((i / 8) % WAVE);
  const long long f = i / (8 * WAVE);
  const int KB = K / 16;
  const int kb = (int)(f % KB), mt = (int)(f / KB);
  const int kl = 16 * kb + 8 * (lane / 32) + v;
  const int k = permuted ? t_phys(kl) : kl;
  WswT[i] = W[(long long)k * N + 32 * mt + lane % 32];
}

// Accumulator start value: the bias of the lane's 16 output features of tile mt
// (so the epilogue needs no separate add).
__device__ __forceinline__ f32x16 t_bias(const float *bias, int mt, int hi)
{
  f32x16 b;
#pragma unroll
  for (int r = 0; r < 16; ++r) b[r] = bias[32 * mt + 8 * (r / 4) + 4 * hi + r % 4];
  return b;
}

// GELU of one 32-feature tile, packed as the next layer's two B fragments
__device__ __forceinline__ void t_epilogue(const f32x16 &acc, bf16x8 &lo, bf16x8 &up)
{
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const bf16 g = (bf16)gelu(acc[r]);
    if (r < 8)
      lo[r] = g;
    else
      up[r - 8] = g;
  }
}

// NC 32-cell tiles per wave, WPB waves per block.
template <int NT, int NC, int WPB>
__global__ void __launch_bounds__(WPB * 64) mlp_infer_t_kernel(InferArgs a)
{
  constexpr int KB = 2 * NT;                      // k-blocks of a hidden layer
  constexpr int NTHR = WPB * WAVE;
  __shared__ bf16x8 wl[NT * KB * WAVE];           // one hidden layer's A fragments
  __shared__ float bl[32 * NT];                   // its bias
  const int w = threadIdx.x / WAVE, l = threadIdx.x & (WAVE - 1);
  const int r32 = l & 31, hi = l >> 5;
  const long long cell0 = ((long long)blockIdx.x * WPB + w) * (32 * NC);

  bf16x8 h[KB][NC];
  // layer 0 (K = 32): input B fragments straight from X, weights from L2
  {
    bf16x8 x[2][NC];
#pragma unroll
    for (int ct = 0; ct < NC; ++ct) {
      const long long c = cell0 + 32 * ct + r32 < a.n ? cell0 + 32 * ct + r32 : a.n - 1;
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
        x[kb][ct] = *reinterpret_cast<const bf16x8 *>(a.X + c * MLP_IN_PAD + 16 * kb + 8 * hi);
    }
    const bf16x8 *W0 = reinterpret_cast<const bf16x8 *>(a.Wsw[0]);
#pragma unroll
    for (int mt = 0; mt < NT; ++mt) {
      const f32x16 b = t_bias(a.bias[0], mt, hi);
      f32x16 acc[NC];
#pragma unroll
      for (int ct = 0; ct < NC; ++ct) acc[ct] = b;
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        const bf16x8 A = W0[(mt * 2 + kb) * WAVE + l];
#pragma unroll
        for (int ct = 0; ct < NC; ++ct) acc[ct] = mfma(A, x[kb][ct], acc[ct]);
      }
#pragma unroll
      for (int ct = 0; ct < NC; ++ct) t_epilogue(acc[ct], h[2 * mt][ct], h[2 * mt + 1][ct]);
    }
  }
  // hidden layers
  for (int layer = 1; layer < a.layers; ++layer) {
    __syncthreads();                               // previous layer's weights no longer read
    const bf16x8 *W = reinterpret_cast<const bf16x8 *>(a.Wsw[layer]);
#pragma unroll 4
    for (int i = threadIdx.x; i < NT * KB * WAVE; i += NTHR) wl[i] = W[i];
    for (int i = threadIdx.x; i < 32 * NT; i += NTHR) bl[i] = a.bias[layer][i];
    __syncthreads();
    bf16x8 hn[KB][NC];
#pragma unroll
    for (int mt = 0; mt < NT; ++mt) {
      const f32x16 b = t_bias(bl, mt, hi);
      f32x16 acc[NC];
#pragma unroll
      for (int ct = 0; ct < NC; ++ct) acc[ct] = b;
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        const bf16x8 A = wl[(mt * KB + kb) * WAVE + l];
#pragma unroll
        for (int ct = 0; ct < NC; ++ct) acc[ct] = mfma(A, h[kb][ct], acc[ct]);
      }
#pragma unroll
      for (int ct = 0; ct < NC; ++ct) t_epilogue(acc[ct], hn[2 * mt][ct], hn[2 * mt + 1][ct]);
    }
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int ct = 0; ct < NC; ++ct) h[kb][ct] = hn[kb][ct];
  }
  // output layer: one 32-feature tile, weights from L2
  const bf16x8 *Wo = reinterpret_cast<const bf16x8 *>(a.Wsw[a.layers]);
  const float *bo = a.bias[a.layers];
  f32x16 o[NC];
  {
    const f32x16 b = t_bias(bo, 0, hi);
#pragma unroll
    for (int ct = 0; ct < NC; ++ct) o[ct] = b;
  }
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    const bf16x8 A = Wo[kb * WAVE + l];
#pragma unroll
    for (int ct = 0; ct < NC; ++ct) o[ct] = mfma(A, h[kb][ct], o[ct]);
  }
#pragma unroll
  for (int ct = 0; ct < NC; ++ct) {
    const long long c = cell0 + 32 * ct + r32;
    if (c >= a.n) continue;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int f = 8 * (r / 4) + 4 * hi + r % 4;
      if (f < a.n_out) a.Y[c * a.n_out + f] = o[ct][r];
    }
  }
}

unsigned blocks_for(long long n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ---- handle ---------------------------------------------------------------------

struct MlpLayer {
  int K, N;                    // padded fan-in, fan-out
  int Kr, Nr;                  // real fan-in, fan-out
  long long w_off, b_off;      // offsets in the padded parameter arrays
  long long sw_off;            // offset in the swizzled bf16 array
};

}  // namespace nrg

struct nrg_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_in = 0, hidden = 0, layers = 0, n_out = 0;
  std::vector<nrg::MlpLayer> L;            // layers + 1 entries
  long long n_pad = 0, n_packed = 0, n_sw = 0;
  float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;
  nrg::bf16 *Pbf = nullptr, *Wsw = nullptr, *WswT = nullptr;   // WswT: transposed-kernel layout
  int *gidx = nullptr;                      // packed index -> padded index
  long long adam_t = 0;
  bool sw_dirty = true;
  // batch workspace
  long long cap = 0;
  nrg::bf16 *X = nullptr, *dY = nullptr;
  float *Yb = nullptr, *loss_tmp = nullptr;
  std::vector<nrg::bf16 *> H;               // per hidden layer, n x hidden
  std::vector<float *> Z;
  nrg::bf16 *dZ[2] = {nullptr, nullptr};
};

namespace nrg {
namespace {

int launch_gemm(nrg_mlp_t *h, const GemmArgs &g0)
{
  GemmArgs g = g0;
  if (g.M % 32 || g.N % 32 || g.K % 16 || g.M <= 0 || g.N <= 0 || g.K <= 0 || g.ksplit < 1)
    return set_error(NRG_CHEM_E_ARG, "gemm shape %dx%dx%d not supported", g.M, g.N, g.K);
  const long long waves = (long long)(g.M / 32) * (g.N / 32) * g.ksplit;
  hipLaunchKernelGGL(gemm_kernel, dim3(blocks_for(waves, 4)), dim3(256), 0, h->stream, g);
  HIP_TRY(hipGetLastError());
  return NRG_CHEM_OK;
}

int ensure_batch(nrg_mlp_t *h, long long n)
{
  if (n <= h->cap) return NRG_CHEM_OK;
  auto fr = [](void *p) { if (p) (void)hipFree(p); };
  fr(h->X); fr(h->dY); fr(h->Yb); fr(h->dZ[0]); fr(h->dZ[1]);
  for (auto p : h->H) fr(p);
  for (auto p : h->Z) fr(p);
  h->H.assign(h->layers, nullptr);
  h->Z.assign(h->layers, nullptr);
  h->X = h->dY = h->dZ[0] = h->dZ[1] = nullptr;
  h->Yb = nullptr;
  h->cap = 0;
  const long long m = ((n + 127) / 128) * 128;
  HIP_TRY(hipMalloc(&h->X, m * MLP_IN_PAD * sizeof(bf16)));
  HIP_TRY(hipMalloc(&h->dY, m * MLP_OUT_PAD * sizeof(bf16)));
  HIP_TRY(hipMalloc(&h->Yb, m * MLP_OUT_PAD * sizeof(float)));
  for (int i = 0; i < 2; ++i) HIP_TRY(hipMalloc(&h->dZ[i], m * h->hidden * sizeof(bf16)));
  for (int i = 0; i < h->layers; ++i) {
    HIP_TRY(hipMalloc(&h->H[i], m * h->hidden * sizeof(bf16)));
    HIP_TRY(hipMalloc(&h->Z[i], m * h->hidden * sizeof(float)));
  }
  h->cap = m;
  return NRG_CHEM_OK;
}

int refresh_swizzle(nrg_mlp_t *h)
{
  if (!h->sw_dirty) return NRG_CHEM_OK;
  for (size_t i = 0; i < h->L.size(); ++i) {
    const MlpLayer &ly = h->L[i];
    const long long n = (long long)ly.K * ly.N;
    hipLaunchKernelGGL(swizzle_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream,
                       h->Pbf + ly.w_off, ly.K, ly.N, h->Wsw + ly.sw_off);
    hipLaunchKernelGGL(swizzle_t_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream,
                       h->Pbf + ly.w_off, ly.K, ly.N, i > 0 ? 1 : 0, h->WswT + ly.sw_off);
  }
  HIP_TRY(hipGetLastError());
  // inference may run on another stream (the chemistry handle's)
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->sw_dirty = false;
  return NRG_CHEM_OK;
}

// Forward of the training path (rows padded to a multiple of 32 inside the workspace).
int forward_train(nrg_mlp_t *h, long long n, const float *x)
{
  int rc = ensure_batch(h, n);
  if (rc) return rc;
  const long long m = ((n + 31) / 32) * 32;
  if (m > n) HIP_TRY(hipMemsetAsync(h->X + n * MLP_IN_PAD, 0, (m - n) * MLP_IN_PAD * sizeof(bf16), h->stream));
  hipLaunchKernelGGL(pack_input_kernel, dim3(blocks_for(n * MLP_IN_PAD, 256)), dim3(256), 0,
                     h->stream, x, n, h->n_in, h->X);
  const bf16 *in = h->X;
  int K = MLP_IN_PAD;
  for (int i = 0; i <= h->layers; ++i) {
    const MlpLayer &ly = h->L[i];
    GemmArgs g{};
    g.A = in; g.sAm = K; g.sAk = 1;
    g.B = h->Pbf + ly.w_off; g.sBk = ly.N; g.sBn = 1;
    g.M = (int)m; g.N = ly.N; g.K = ly.K; g.ksplit = 1; g.ldc = ly.N;
    g.bias = h->P + ly.b_off;
    if (i < h->layers) {
      g.mode = EPI_BIAS_GELU; g.Z = h->Z[i]; g.H = h->H[i];
    } else {
      g.mode = EPI_BIAS; g.C = h->Yb;
    }
    if ((rc = launch_gemm(h, g))) return rc;
    in = h->H[i < h->layers ? i : 0];
    K = ly.N;
  }
  return NRG_CHEM_OK;
}

// Backward into h->G (padded layout); dY already in h->dY.
int backward(nrg_mlp_t *h, long long n)
{
  const long long m = ((n + 31) / 32) * 32;
  if (m > n) HIP_TRY(hipMemsetAsync(h->dY + n * MLP_OUT_PAD, 0, (m - n) * MLP_OUT_PAD * sizeof(bf16), h->stream));
  hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks_for(h->n_pad, 256)), dim3(256), 0, h->stream,
                     h->G, h->n_pad, 0.0f);
  const bf16 *dZ = h->dY;
  int cur = 0;
  int rc;
  const int ksplit = (int)std::max<long long>(1, std::min<long long>(64, m / 512));
  for (int i = h->layers; i >= 0; --i) {
    const MlpLayer &ly = h->L[i];
    const bf16 *Hin = i > 0 ? h->H[i - 1] : h->X;
    const int Kin = ly.K;
    // dW = Hin^T dZ   (Kin x N, reduction over the batch)
    GemmArgs g{};
    g.A = Hin; g.sAm = 1; g.sAk = Kin;
    g.B = dZ; g.sBk = ly.N; g.sBn = 1;
    g.M = Kin; g.N = ly.N; g.K = (int)m; g.mode = EPI_STORE; g.ksplit = ksplit; g.ldc = ly.N;
    g.C = h->G + ly.w_off;
    if ((rc = launch_gemm(h, g))) return rc;
    hipLaunchKernelGGL(bias_grad_kernel, dim3(blocks_for(ly.N, 256), (unsigned)blocks_for(m, 256)),
                       dim3(256), 0, h->stream, dZ, m, ly.N, h->G + ly.b_off);
    if (i == 0) break;
    // dZ_{i-1} = (dZ W^T) * gelu'(Z_{i-1})
    GemmArgs b{};
    b.A = dZ; b.sAm = ly.N; b.sAk = 1;
    b.B = h->Pbf + ly.w_off; b.sBk = 1; b.sBn = ly.N;
    b.M = (int)m; b.N = Kin; b.K = ly.N; b.mode = EPI_GELU_BACK; b.ksplit = 1; b.ldc = Kin;
    b.Zin = h->Z[i - 1]; b.H = h->dZ[cur];
    if ((rc = launch_gemm(h, b))) return rc;
    dZ = h->dZ[cur];
    cur ^= 1;
  }
  HIP_TRY(hipGetLastError());
  return NRG_CHEM_OK;
}

int loss_and_grad(nrg_mlp_t *h, long long n, const float *x, const float *t, const float *w,
                  float *loss)
{
  int rc = forward_train(h, n, x);
  if (rc) return rc;
  if (loss) hipLaunchKernelGGL(fill_f32_kernel, dim3(1), dim3(64), 0, h->stream, loss, 1, 0.0f);
  hipLaunchKernelGGL(loss_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, h->Yb, t, w,
                     n, h->n_out, h->dY, loss);
  return backward(h, n);
}

}  // namespace

int mlp_infer_bf16(nrg_mlp_t *h, long long n, const bf16 *X, float *Y, hipStream_t stream)
{
  if (n <= 0) return NRG_CHEM_OK;
  int rc = refresh_swizzle(h);
  if (rc) return rc;
  InferArgs a{};
  a.X = X; a.n = n; a.n_out = h->n_out; a.hidden = h->hidden; a.layers = h->layers; a.Y = Y;
  const bool transposed = h->hidden == 256;
  for (int i = 0; i <= h->layers; ++i) {
    a.Wsw[i] = (transposed ? h->WswT : h->Wsw) + h->L[i].sw_off;
    a.bias[i] = h->P + h->L[i].b_off;
  }
  const dim3 grid(blocks_for(n, 128)), block(256);
  switch (h->hidden) {
    case 64: hipLaunchKernelGGL(mlp_infer_kernel<2>, grid, block, 0, stream, a); break;
    case 128: hipLaunchKernelGGL(mlp_infer_kernel<4>, grid, block, 0, stream, a); break;
    case 256:
      hipLaunchKernelGGL((mlp_infer_t_kernel<8, T_NC, T_WPB>), dim3(blocks_for(n, 32 * T_NC * T_WPB)),
                         dim3(64 * T_WPB), 0, stream, a);
      break;
    default: return set_error(NRG_CHEM_E_ARG, "hidden width %d has no inference kernel", h->hidden);
  }
  HIP_TRY(hipGetLastError());
  return NRG_CHEM_OK;
}

hipStream_t mlp_stream(nrg_mlp_t *h) { return h->stream; }
int mlp_n_out(const nrg_mlp_t *h) { return h->n_out; }
int mlp_n_in(const nrg_mlp_t *h) { return h->n_in; }

}  // namespace nrg

using namespace nrg;

extern "C" {

int nrg_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out, int32_t device,
                   nrg_mlp_t **out)
{
  if (!out) return set_error(NRG_CHEM_E_ARG, "out is NULL");
  *out = nullptr;
  if (n_in < 1 || n_in > MLP_IN_PAD || n_out < 1 || n_out > MLP_OUT_PAD || n_layers < 1 ||
      n_layers > MLP_MAX_LAYERS - 1 || !(n_hidden == 64 || n_hidden == 128 || n_hidden == 256))
    return set_error(NRG_CHEM_E_ARG, "unsupported network %d -> %d x %d -> %d", n_in, n_hidden,
                     n_layers, n_out);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return set_error(NRG_CHEM_E_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= ndev) return set_error(NRG_CHEM_E_NO_DEVICE, "device %d out of range", device);
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(NRG_CHEM_E_NO_DEVICE, "device %d is %s, kernels are built for gfx950", device,
                     prop.gcnArchName);
  nrg_mlp_t *h = new (std::nothrow) nrg_mlp();
  if (!h) return set_error(NRG_CHEM_E_ARG, "out of host memory");
  h->device = device;
  h->n_in = n_in; h->hidden = n_hidden; h->layers = n_layers; h->n_out = n_out;
  long long off = 0, sw = 0, packed = 0;
  for (int i = 0; i <= n_layers; ++i) {
    MlpLayer ly;
    ly.K = i == 0 ? MLP_IN_PAD : n_hidden;
    ly.N = i == n_layers ? MLP_OUT_PAD : n_hidden;
    ly.Kr = i == 0 ? n_in : n_hidden;
    ly.Nr = i == n_layers ? n_out : n_hidden;
    ly.w_off = off; off += (long long)ly.K * ly.N;
    ly.b_off = off; off += ly.N;
    ly.sw_off = sw; sw += (long long)ly.K * ly.N;
    packed += (long long)ly.Kr * ly.Nr + ly.Nr;
    h->L.push_back(ly);
  }
  h->n_pad = off; h->n_sw = sw; h->n_packed = packed;
  auto fail = [&](int rc) { nrg_mlp_destroy(h); return rc; };
  if (hipSetDevice(device) != hipSuccess) return fail(set_error(NRG_CHEM_E_HIP, "hipSetDevice"));
  if (hipStreamCreate(&h->stream) != hipSuccess) return fail(set_error(NRG_CHEM_E_HIP, "hipStreamCreate"));
  const size_t fb = h->n_pad * sizeof(float);
  if (hipMalloc(&h->P, fb) || hipMalloc(&h->G, fb) || hipMalloc(&h->M1, fb) || hipMalloc(&h->M2, fb) ||
      hipMalloc(&h->Pbf, h->n_pad * sizeof(bf16)) || hipMalloc(&h->Wsw, h->n_sw * sizeof(bf16)) ||
      hipMalloc(&h->WswT, h->n_sw * sizeof(bf16)) ||
      hipMalloc(&h->gidx, h->n_packed * sizeof(int)) || hipMalloc(&h->loss_tmp, sizeof(float)))
    return fail(set_error(NRG_CHEM_E_HIP, "hipMalloc parameters"));
  std::vector<int> gi;
  gi.reserve(h->n_packed);
  for (const MlpLayer &ly : h->L) {
    for (int k = 0; k < ly.Kr; ++k)
      for (int n = 0; n < ly.Nr; ++n) gi.push_back((int)(ly.w_off + (long long)k * ly.N + n));
    for (int n = 0; n < ly.Nr; ++n) gi.push_back((int)(ly.b_off + n));
  }
  if (hipMemcpy(h->gidx, gi.data(), gi.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_error(NRG_CHEM_E_HIP, "hipMemcpy index map"));
  std::vector<float> zero(h->n_packed, 0.0f);
  int rc = nrg_mlp_set_params(h, zero.data());
  if (rc) return fail(rc);
  *out = h;
  return NRG_CHEM_OK;
}

void nrg_mlp_destroy(nrg_mlp_t *h)
{
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void *p : {(void *)h->P, (void *)h->G, (void *)h->M1, (void *)h->M2, (void *)h->Pbf,
                  (void *)h->Wsw, (void *)h->WswT, (void *)h->gidx, (void *)h->loss_tmp, (void *)h->X, (void *)h->dY,
                  (void *)h->Yb, (void *)h->dZ[0], (void *)h->dZ[1]})
    if (p) (void)hipFree(p);
  for (auto p : h->H) if (p) (void)hipFree(p);
  for (auto p : h->Z) if (p) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int64_t nrg_mlp_num_params(const nrg_mlp_t *h) { return h ? h->n_packed : 0; }

int nrg_mlp_set_params(nrg_mlp_t *h, const float *params)
{
  if (!h || !params) return set_error(NRG_CHEM_E_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(h->device));
  std::vector<float> pad(h->n_pad, 0.0f);
  long long q = 0;
  for (const MlpLayer &ly : h->L) {
    for (int k = 0; k < ly.Kr; ++k)
      for (int n = 0; n < ly.Nr; ++n) pad[ly.w_off + (long long)k * ly.N + n] = params[q++];
    for (int n = 0; n < ly.Nr; ++n) pad[ly.b_off + n] = params[q++];
  }
  HIP_TRY(hipMemcpyAsync(h->P, pad.data(), h->n_pad * sizeof(float), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks_for(h->n_pad, 256)), dim3(256), 0, h->stream,
                     h->M1, h->n_pad, 0.0f);
  hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks_for(h->n_pad, 256)), dim3(256), 0, h->stream,
                     h->M2, h->n_pad, 0.0f);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(h->n_pad, 256)), dim3(256), 0, h->stream,
                     h->P, h->Pbf, h->n_pad);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->adam_t = 0;
  h->sw_dirty = true;
  return NRG_CHEM_OK;
}

int nrg_mlp_get_params(nrg_mlp_t *h, float *params)
{
  if (!h || !params) return set_error(NRG_CHEM_E_ARG, "NULL argument");
  HIP_TRY(hipSetDevice(h->device));
  std::vector<float> pad(h->n_pad);
  HIP_TRY(hipMemcpyAsync(pad.data(), h->P, h->n_pad * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  long long q = 0;
  for (const MlpLayer &ly : h->L) {
    for (int k = 0; k < ly.Kr; ++k)
      for (int n = 0; n < ly.Nr; ++n) params[q++] = pad[ly.w_off + (long long)k * ly.N + n];
    for (int n = 0; n < ly.Nr; ++n) params[q++] = pad[ly.b_off + n];
  }
  return NRG_CHEM_OK;
}

int nrg_mlp_forward_dev(nrg_mlp_t *h, int64_t n, const float *x, float *y)
{
  if (!h || (n > 0 && (!x || !y)) || n < 0) return set_error(NRG_CHEM_E_ARG, "bad argument");
  if (n == 0) return NRG_CHEM_OK;
  HIP_TRY(hipSetDevice(h->device));
  int rc = ensure_batch(h, n);
  if (rc) return rc;
  hipLaunchKernelGGL(pack_input_kernel, dim3(blocks_for(n * MLP_IN_PAD, 256)), dim3(256), 0,
                     h->stream, x, n, h->n_in, h->X);
  if ((rc = mlp_infer_bf16(h, n, h->X, y, h->stream))) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  return NRG_CHEM_OK;
}

int nrg_mlp_grad_dev(nrg_mlp_t *h, int64_t n, const float *x, const float *t, const float *w,
                     float *grad, float *loss, float *y_out)
{
  if (!h || n <= 0 || !x || !t || !grad) return set_error(NRG_CHEM_E_ARG, "bad argument");
  HIP_TRY(hipSetDevice(h->device));
  int rc = loss_and_grad(h, n, x, t, w, loss);
  if (rc) return rc;
  hipLaunchKernelGGL(gather_kernel, dim3(blocks_for(h->n_packed, 256)), dim3(256), 0, h->stream,
                     h->G, h->gidx, grad, h->n_packed);
  if (y_out)
    HIP_TRY(hipMemcpy2DAsync(y_out, h->n_out * sizeof(float), h->Yb, MLP_OUT_PAD * sizeof(float),
                             h->n_out * sizeof(float), n, hipMemcpyDeviceToDevice, h->stream));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return NRG_CHEM_OK;
}

int nrg_mlp_train_step_dev(nrg_mlp_t *h, int64_t n, const float *x, const float *t, const float *w,
                           float lr, float *loss)
{
  if (!h || n <= 0 || !x || !t) return set_error(NRG_CHEM_E_ARG, "bad argument");
  HIP_TRY(hipSetDevice(h->device));
  int rc = loss_and_grad(h, n, x, t, w, loss);
  if (rc) return rc;
  h->adam_t += 1;
  const float bc1 = 1.0f - std::pow(0.9f, (float)h->adam_t);
  const float bc2 = 1.0f - std::pow(0.999f, (float)h->adam_t);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(h->n_pad, 256)), dim3(256), 0, h->stream, h->P,
                     h->Pbf, h->G, h->M1, h->M2, h->n_pad, lr, bc1, bc2);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->sw_dirty = true;
  return NRG_CHEM_OK;
}

void *nrg_mlp_stream(nrg_mlp_t *h) { return h ? (void *)h->stream : nullptr; }

}  // extern "C"
