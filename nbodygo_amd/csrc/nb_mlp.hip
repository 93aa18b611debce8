// nb_mlp.hip -- small MLP (d_in -> H -> ... -> H -> d_out, GELU) on gfx950 matrix cores.
//
// No counterpart in the reference.  It is the learned part of the far-field surrogate: BF16
// weights and activations feed the matrix cores with FP32 accumulation (gfx950's
// v_mfma_f32_16x16x32_bf16 in the forward pass and inference, v_mfma_f32_16x16x16_bf16 in the
// backward GEMMs); master weights, biases, pre-activations, gradients and the Adam state are FP32.
//
// GEMM: C[M x N] (FP32, row-major) = op(A)[M x K] * op(B)[K x N], A and B BF16 row-major, op =
// identity or transpose.  A 256-thread workgroup computes a 64 x 64 tile as 2 x 2 waves of 32 x 32,
// each wave 2 x 2 MFMA blocks of 16 x 16, stepping K by 16 (or 32) through LDS.  Operand layout of
// the 16 x 16 x 16 BF16 MFMA (checked against a host GEMM in tests): lane l holds A[l % 16][4*(l/16)
// .. +3] and B[4*(l/16) .. +3][l % 16]; the 16 x 16 x 32 form holds eight, 8*(l/16) .. +7; the
// accumulator holds D[4*(l/16) + v][l % 16] in register v.  Split-K over grid.z writes partial
// tiles that a second kernel sums in fixed order.  Dimensions are padded by the caller: M, N
// multiples of 64, K a multiple of 16.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "nb_internal.hpp"

namespace nb {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));  // gfx950's 16 x 16 x 32 MFMA operands
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even: one v_cvt_pk_bf16_f32 (for finite inputs the same bits as adding
  // 0x7fff + the kept lowest bit to the FP32 word and taking its upper half)
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// tanh(u) = 1 - 2 / (e^(2u) + 1), with the hardware base-2 exponential and reciprocal: an
// absolute error of a few 1e-8 (what GELU's 1 + tanh needs), saturating to +-1 for large |u|.
__device__ __forceinline__ float tanh_fast(float u) {
  const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * u);  // 2 log2(e) u
  return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}

// GELU, tanh form: 0.5 x (1 + tanh(z)), z = sqrt(2/pi) (x + 0.044715 x^3).  Since
// 1 + tanh(z) = 2 - 2 / (e^(2z) + 1), this is x - x / (e^(2z) + 1), and e^(2z) = 2^(x (c0 + c1 x^2))
// with c0 = 2 log2(e) sqrt(2/pi), c1 = 0.044715 c0: five VALU operations around the hardware
// exponential and reciprocal (the surrogate's inference evaluates it 96 times per lane and tile).
// Large x gives x, very negative x gives 0.
__device__ __forceinline__ float gelu(float x) {
  const float c0 = 2.302208198144325f, c1 = 0.1029432395800235f;
  const float e = __builtin_amdgcn_exp2f(x * __builtin_fmaf(c1, x * x, c0));
  return __builtin_fmaf(-x, __builtin_amdgcn_rcpf(e + 1.f), x);
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float k0 = 0.7978845608028654f, k1 = 0.044715f;
  const float u = k0 * (x + k1 * x * x * x);
  const float t = tanh_fast(u);
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * k0 * (1.f + 3.f * k1 * x * x);
}

enum { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2 };

// KS = 32 (no transposes; K and the split a multiple of 32): gfx950's v_mfma_f32_16x16x32_bf16,
// twice the BF16 rate of the 16 x 16 x 16 form, lane l holding A[l % 16][8 (l / 16) .. +7] and
// B[8 (l / 16) .. +7][l % 16] (the same K order for both operands, so the product is the
// GEMM's whatever order the hardware sums in).  The forward pass and the fused inference use it.
template <bool TA, bool TB, int KS = 16>
__global__ __launch_bounds__(256) void k_gemm_bf16(const uint16_t *__restrict__ A,
                                                   const uint16_t *__restrict__ B, float *C,
                                                   int M, int N, int K, int k_per_split,
                                                   int lda, int ldb, int epi,
                                                   const float *__restrict__ bias,
                                                   uint16_t *__restrict__ Hout) {
  static_assert(KS == 16 || (KS == 32 && !TA && !TB), "the 32-wide step has no transposes");
  __shared__ __attribute__((aligned(16))) uint16_t As[64][KS + (KS == 32 ? 8 : 4)];  // [m][k]
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64][KS + (KS == 32 ? 8 : 4)];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int kb = blockIdx.z * k_per_split;
  const int ke = min(K, kb + k_per_split);
  f32x4 acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  if constexpr (KS == 32) {
    for (int k0 = kb; k0 < ke; k0 += 32) {
      {  // A tile [m0 .. +63][k0 .. +31], 8 per thread
        const int r = tid >> 2, kk = (tid & 3) * 8;
        const uint16_t *p = A + (size_t)(m0 + r) * lda + k0 + kk;
        for (int e = 0; e < 8; e++) As[r][kk + e] = p[e];
      }
      {  // B tile [k0 .. +31][n0 .. +63] -> Bs[n][k], 8 per thread
        const int kk = tid >> 3, nn = (tid & 7) * 8;
        const uint16_t *p = B + (size_t)(k0 + kk) * ldb + n0 + nn;
        for (int e = 0; e < 8; e++) Bs[nn + e][kk] = p[e];
      }
      __syncthreads();
      const int kq = 8 * (lane >> 4);
      for (int a = 0; a < 2; a++) {
        const int row = wm * 32 + a * 16 + (lane & 15);
        const bf16x8 fa = *(const bf16x8 *)&As[row][kq];
        for (int b = 0; b < 2; b++) {
          const int col = wn * 32 + b * 16 + (lane & 15);
          const bf16x8 fb = *(const bf16x8 *)&Bs[col][kq];
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, acc[a][b], 0, 0, 0);
        }
      }
      __syncthreads();
    }
  } else
  for (int k0 = kb; k0 < ke; k0 += 16) {
    // A tile: op(A)[m0 .. m0+63][k0 .. k0+15]
    if (!TA) {
      const int r = tid >> 2, kk = (tid & 3) * 4;
      const uint16_t *p = A + (size_t)(m0 + r) * lda + k0 + kk;
      for (int e = 0; e < 4; e++) As[r][kk + e] = p[e];
    } else {
      const int kk = tid >> 4, mm = (tid & 15) * 4;  // A stored [K][M]
      const uint16_t *p = A + (size_t)(k0 + kk) * lda + m0 + mm;
      for (int e = 0; e < 4; e++) As[mm + e][kk] = p[e];
    }
    // B tile: op(B)[k0 .. k0+15][n0 .. n0+63] -> Bs[n][k]
    if (!TB) {
      const int kk = tid >> 4, nn = (tid & 15) * 4;  // B stored [K][N]
      const uint16_t *p = B + (size_t)(k0 + kk) * ldb + n0 + nn;
      for (int e = 0; e < 4; e++) Bs[nn + e][kk] = p[e];
    } else {
      const int nn = tid >> 2, kk = (tid & 3) * 4;   // B stored [N][K]
      const uint16_t *p = B + (size_t)(n0 + nn) * ldb + k0 + kk;
      for (int e = 0; e < 4; e++) Bs[nn][kk + e] = p[e];
    }
    __syncthreads();
    const int kq = 4 * (lane >> 4);
    for (int a = 0; a < 2; a++) {
      const int row = wm * 32 + a * 16 + (lane & 15);
      bf16x4 fa;
      for (int e = 0; e < 4; e++) fa[e] = (short)As[row][kq + e];
      for (int b = 0; b < 2; b++) {
        const int col = wn * 32 + b * 16 + (lane & 15);
        bf16x4 fb;
        for (int e = 0; e < 4; e++) fb[e] = (short)Bs[col][kq + e];
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa, fb, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float *Cz = C + (size_t)blockIdx.z * M * N;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int v = 0; v < 4; v++) {
        const int row = m0 + wm * 32 + a * 16 + 4 * (lane >> 4) + v;
        const int col = n0 + wn * 32 + b * 16 + (lane & 15);
        float val = acc[a][b][v];
        if (epi == EPI_STORE) {
          Cz[(size_t)row * N + col] = val;
        } else {
          val += bias[col];
          Cz[(size_t)row * N + col] = val;  // pre-activation (or output)
          if (epi == EPI_BIAS_GELU) Hout[(size_t)row * N + col] = f2bf(gelu(val));
        }
      }
}

// Sum split-K partials in fixed order: C[i] = sum_z P[z][i].
__global__ void k_splitk_sum(const float *P, float *C, int64_t count, int splits) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int z = 0; z < splits; z++) s += P[(size_t)z * count + i];
  C[i] = s;
}

__global__ void k_f32_to_bf16(const float *x, uint16_t *y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Rows of width `cols` (row stride ld) converted and zero-padded to width ldp.
__global__ void k_pad_rows_bf16(const float *x, int rows, int cols, int ld, uint16_t *y, int ldp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * ldp) return;
  const int r = (int)(i / ldp), c = (int)(i % ldp);
  y[i] = c < cols ? f2bf(x[(size_t)r * ld + c]) : 0;
}

// dZ = dH * gelu'(Z)  (BF16 out for the next GEMMs), keeping FP32 dZ for the bias gradient.
__global__ void k_gelu_back(const float *dH, const float *Z, float *dZ, uint16_t *dZb, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = dH[i] * gelu_grad(Z[i]);
  dZ[i] = g;
  dZb[i] = f2bf(g);
}

// Column sums of X[rows][cols] (bias gradients): a workgroup of 16 x 64 threads per 64 columns;
// thread (g, c) sums rows g, g+16, ... in order (coalesced across c), then the 16 partial sums
// meet in a fixed LDS tree.
__global__ __launch_bounds__(1024) void k_colsum(const float *X, int rows, int cols,
                                                 float *out) {
  __shared__ float sh[16][64];
  const int cl = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cl;
  float s = 0.f;
  if (c < cols)
    for (int r = g; r < rows; r += 16) s += X[(size_t)r * cols + c];
  sh[g][cl] = s;
  __syncthreads();
  for (int w = 8; w > 0; w >>= 1) {
    if (g < w) sh[g][cl] += sh[g + w][cl];
    __syncthreads();
  }
  if (g == 0 && c < cols) out[c] = sh[0][cl];
}

// MSE over the first dout columns: L = sum (y - t)^2 / (rows * dout); dY = 2 (y - t) / (rows *
// dout) in FP32 and BF16; padded columns get zero gradient.  Per-row losses to lrow.
__global__ void k_mse(const float *Y, int ldy, const float *T, int rows, int dout, float *dY,
                      uint16_t *dYb, float *lrow) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float l = 0.f;
  const float sc = 2.f / ((float)rows * (float)dout);
  for (int c = 0; c < ldy; c++) {
    float g = 0.f;
    if (c < dout) {
      const float d = Y[(size_t)r * ldy + c] - T[(size_t)r * dout + c];
      l += d * d;
      g = sc * d;
    }
    dY[(size_t)r * ldy + c] = g;
    dYb[(size_t)r * ldy + c] = f2bf(g);
  }
  lrow[r] = l;
}

__global__ void k_adam(float *w, const float *g, float *m1, float *m2, uint16_t *wb, int64_t n,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float wi = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

// [K][N] -> [N][K] (BF16), so that an MFMA B fragment (four consecutive k of one column) is one
// 8-byte load.
__global__ void k_transpose_bf16(const uint16_t *src, int K, int N, uint16_t *dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * N) return;
  const int k = i / N, n = i % N;
  dst[(size_t)n * K + k] = src[i];
}

// Fused inference: every layer of the MLP for 64 rows per workgroup, activations kept in LDS as
// BF16 (exactly what the layer-by-layer path stores between its GEMMs), weights read from the
// transposed BF16 copies.  The MFMA sequence per output element (gfx950's 16 x 16 x 32 form, K in
// steps of 32 from 0) and the epilogue (bias, GELU, BF16 rounding) are those of the forward
// k_gemm_bf16, so the outputs equal the layer-by-layer forward bit for bit.  Limits: padded input width and hidden width <= 128,
// hidden width a multiple of 64, d_out <= 16.
constexpr int kInferW = 128;
constexpr int kInferMaxLayers = 6;
struct InferArgs {
  const float *x;
  int rows, din, dinp, hid, layers, dout;
  const uint16_t *wt[kInferMaxLayers + 1];  // [out][in]
  const float *b[kInferMaxLayers + 1];
  float *y;  // [rows x dout]
};

// One hidden layer for the workgroup's 64 rows: Hout = BF16(GELU(Hin * W + b)).  Wave w owns
// columns [w HID/4, (w+1) HID/4); fb holds its B fragments of the layer and bc its biases.
template <int K, int HID>
__device__ __forceinline__ void infer_layer_w(const uint16_t (*Hin)[kInferW + 8],
                                              uint16_t (*Hout)[kInferW + 8],
                                              const bf16x8 (&fb)[K / 32][HID / 64],
                                              const float (&bc)[HID / 64], int w, int lr,
                                              int lq) {
  constexpr int NBW = HID / 64, KS = K / 32;
  f32x4 acc[4][NBW];
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int b = 0; b < NBW; b++) acc[m][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    bf16x8 fa[4];
#pragma unroll
    for (int m = 0; m < 4; m++) fa[m] = *(const bf16x8 *)&Hin[m * 16 + lr][ks * 32 + 2 * lq];
#pragma unroll
    for (int b = 0; b < NBW; b++)
#pragma unroll
      for (int m = 0; m < 4; m++)
        acc[m][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[m], fb[ks][b], acc[m][b], 0, 0, 0);
  }
#pragma unroll
  for (int b = 0; b < NBW; b++) {
    const int col = w * (HID / 4) + b * 16 + lr;
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int v = 0; v < 4; v++) Hout[m * 16 + lq + v][col] = f2bf(gelu(acc[m][b][v] + bc[b]));
  }
}

// The wave's B fragments and biases of one hidden layer.
template <int K, int HID>
__device__ __forceinline__ void infer_frags(bf16x8 (&fb)[K / 32][HID / 64],
                                            float (&bc)[HID / 64], const uint16_t *wt,
                                            const float *bias, int w, int lr, int lq) {
#pragma unroll
  for (int ks = 0; ks < K / 32; ks++)
#pragma unroll
    for (int b = 0; b < HID / 64; b++)
      fb[ks][b] = *(const bf16x8 *)(wt + (size_t)(w * (HID / 4) + b * 16 + lr) * K + ks * 32 +
                                    2 * lq);
#pragma unroll
  for (int b = 0; b < HID / 64; b++) bc[b] = bias[w * (HID / 4) + b * 16 + lr];
}

// All B fragments of the layer are fetched up front.
template <int K, int HID>
__device__ __forceinline__ void infer_layer(const uint16_t (*Hin)[kInferW + 8],
                                            uint16_t (*Hout)[kInferW + 8], const uint16_t *wt,
                                            const float *bias, int w, int lr, int lq) {
  bf16x8 fb[K / 32][HID / 64];
  float bc[HID / 64];
  infer_frags<K, HID>(fb, bc, wt, bias, w, lr, lq);
  infer_layer_w<K, HID>(Hin, Hout, fb, bc, w, lr, lq);
}

template <int DINP, int HID>
__global__ __launch_bounds__(256) void k_mlp_infer(InferArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t Hs[2][64][kInferW + 8];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r0 = blockIdx.x * 64;
  for (int i = tid; i < 64 * DINP; i += 256) {
    const int r = i / DINP, c = i % DINP, row = r0 + r;
    Hs[0][r][c] = (row < a.rows && c < a.din) ? f2bf(a.x[(size_t)row * a.din + c]) : 0;
  }
  __syncthreads();
  const int lr = lane & 15, lq = 4 * (lane >> 4);
  infer_layer<DINP, HID>(Hs[0], Hs[1], a.wt[0], a.b[0], w, lr, lq);
  __syncthreads();
  int cur = 1;
  for (int l = 1; l < a.layers; l++) {
    infer_layer<HID, HID>(Hs[cur], Hs[cur ^ 1], a.wt[l], a.b[l], w, lr, lq);
    __syncthreads();
    cur ^= 1;
  }
  // output layer, columns 0..15; wave w takes rows 16w .. 16w+15
  constexpr int KS = HID / 32;
  const uint16_t *wt = a.wt[a.layers];
  bf16x8 fb[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ks++)
    fb[ks] = *(const bf16x8 *)(wt + (size_t)lr * HID + ks * 32 + 2 * lq);
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    const bf16x8 fa = *(const bf16x8 *)&Hs[cur][w * 16 + lr][ks * 32 + 2 * lq];
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb[ks], acc, 0, 0, 0);
  }
  if (lr < a.dout) {
    const float bc = a.b[a.layers][lr];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = r0 + w * 16 + lq + v;
      if (row < a.rows) a.y[(size_t)row * a.dout + lr] = acc[v] + bc;
    }
  }
}

// Persistent inference for NH = 1..3 hidden layers: each wave keeps its B fragments and biases
// of every layer in registers for the whole launch, and the workgroup walks 64-row tiles; the
// next tile's input rows are requested while the current tile's layers run.  The MFMA sequence
// and the epilogue per output element are k_mlp_infer's, so the outputs are the same bit for
// bit.  Hs[2] holds the input tile (its columns din..DINP-1 stay zero).
template <int DINP, int HID, int NH>
__global__ __launch_bounds__(256) void k_mlp_infer_p(InferArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t Hs[3][64][kInferW + 8];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 15, lq = 4 * (lane >> 4);
  constexpr int KH = HID / 32, NE = DINP / 4;
  bf16x8 f0[DINP / 32][HID / 64], fh[NH][KH][HID / 64], fo[KH];
  float c0[HID / 64], ch[NH][HID / 64];
  infer_frags<DINP, HID>(f0, c0, a.wt[0], a.b[0], w, lr, lq);
#pragma unroll
  for (int l = 1; l < NH; l++) infer_frags<HID, HID>(fh[l - 1], ch[l - 1], a.wt[l], a.b[l], w, lr, lq);
#pragma unroll
  for (int ks = 0; ks < KH; ks++)
    fo[ks] = *(const bf16x8 *)(a.wt[NH] + (size_t)lr * HID + ks * 32 + 2 * lq);
  const float co = lr < a.dout ? a.b[NH][lr] : 0.f;
  for (int i = tid; i < 64 * (kInferW + 8); i += 256) (&Hs[2][0][0])[i] = 0;
  // this thread's input slots: row tid / 4, columns tid % 4 + 4 e
  const int xrow = tid >> 2, xc0 = tid & 3;
  const int ntiles = (a.rows + 63) / 64;
  float xr[NE];
  auto fetch = [&](int tile) {
    const int row = tile * 64 + xrow;
#pragma unroll
    for (int e = 0; e < NE; e++) {
      const int c = xc0 + 4 * e;
      xr[e] = (tile < ntiles && row < a.rows && c < a.din) ? a.x[(size_t)row * a.din + c] : 0.f;
    }
  };
  int tile = blockIdx.x;
  fetch(tile);
  __syncthreads();
  for (; tile < ntiles; tile += gridDim.x) {
#pragma unroll
    for (int e = 0; e < NE; e++) {
      const int c = xc0 + 4 * e;
      if (c < a.din) Hs[2][xrow][c] = f2bf(xr[e]);
    }
    fetch(tile + gridDim.x);
    __syncthreads();
    infer_layer_w<DINP, HID>(Hs[2], Hs[0], f0, c0, w, lr, lq);
    __syncthreads();
    int cur = 0;
#pragma unroll
    for (int l = 1; l < NH; l++) {
      infer_layer_w<HID, HID>(Hs[cur], Hs[cur ^ 1], fh[l - 1], ch[l - 1], w, lr, lq);
      __syncthreads();
      cur ^= 1;
    }
    // output layer, columns 0..15; wave w takes rows 16w .. 16w+15
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KH; ks++) {
      const bf16x8 fa = *(const bf16x8 *)&Hs[cur][w * 16 + lr][ks * 32 + 2 * lq];
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fo[ks], acc, 0, 0, 0);
    }
    if (lr < a.dout) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = tile * 64 + w * 16 + lq + v;
        if (row < a.rows) a.y[(size_t)row * a.dout + lr] = acc[v] + co;
      }
    }
    __syncthreads();  // every read of this tile's buffers precedes the next tile's writes
  }
}

}  // namespace nb

// ---------------------------------------------------------------------------------------------
// Host side: the MLP object (also used by the surrogate).
// ---------------------------------------------------------------------------------------------
using namespace nb;

struct nb_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int din = 0, dinp = 0, hid = 0, layers = 0, dout = 0, doutp = 0;
  // per linear layer l (0..layers): W_l [in_l x out_l] (FP32 master + BF16), b_l [out_l]
  std::vector<int> in_dim, out_dim;           // padded
  std::vector<float *> W, Wg, Wm1, Wm2, bvec, bg, bm1, bm2;
  std::vector<uint16_t *> Wb;
  int64_t rows_cap = 0;
  // activations: Xb (input, BF16), per hidden layer: Z (FP32 pre-act), Hb (BF16 act)
  uint16_t *Xb = nullptr;
  std::vector<float *> Z;
  std::vector<uint16_t *> Hb;
  float *Y = nullptr;       // output [rows x doutp]
  float *T = nullptr;       // targets [rows x dout]
  float *dY = nullptr, *dZ = nullptr, *dH = nullptr, *lrow = nullptr, *splitbuf = nullptr;
  uint16_t *dYb = nullptr, *dZb = nullptr;
  float *xin = nullptr;     // FP32 input staging [rows x din]
  int64_t adam_t = 0;
  int64_t splitbuf_len = 0;
  // transposed BF16 weights for the fused inference kernel, rebuilt when wver moves
  std::vector<uint16_t *> WbT;
  int64_t wver = 0, wt_ver = -1;
  int infer_wgs = 0;  // resident workgroups of the persistent inference kernel (0: not yet known)
};

namespace {

bool mhip(nb_mlp *m, hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  m->err = buf;
  return false;
}
#define NB_MHIP(m, call)                               \
  do {                                                 \
    if (!mhip((m), (call), #call)) return NB_E_HIP;    \
  } while (0)

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// C[M x N] = op(A) op(B) with optional split-K (partials summed in fixed order).
int gemm(nb_mlp *m, bool ta, bool tb, const uint16_t *A, const uint16_t *B, float *C, int M,
         int N, int K, int lda, int ldb, int epi, const float *bias, uint16_t *H, int splits) {
  if (M % 64 || N % 64 || K % 16) {
    m->err = "gemm: dimensions not padded";
    return NB_E_INVALID;
  }
  if (splits < 1) splits = 1;
  int kps = ((K / 16 + splits - 1) / splits) * 16;
  splits = (K + kps - 1) / kps;
  float *out = C;
  if (splits > 1) {
    if ((int64_t)splits * M * N > m->splitbuf_len) {
      if (m->splitbuf) (void)hipFree(m->splitbuf);
      m->splitbuf_len = (int64_t)splits * M * N;
      NB_MHIP(m, hipMalloc((void **)&m->splitbuf, (size_t)m->splitbuf_len * sizeof(float)));
    }
    out = m->splitbuf;
    epi = EPI_STORE;
  }
  dim3 grid((unsigned)(N / 64), (unsigned)(M / 64), (unsigned)splits);
  if (!ta && !tb && K % 32 == 0 && kps % 32 == 0)  // every forward GEMM (padded K, no split)
    hipLaunchKernelGGL((k_gemm_bf16<false, false, 32>), grid, dim3(256), 0, m->stream, A, B, out,
                       M, N, K, kps, lda, ldb, epi, bias, H);
  else if (!ta && !tb)
    hipLaunchKernelGGL((k_gemm_bf16<false, false>), grid, dim3(256), 0, m->stream, A, B, out, M,
                       N, K, kps, lda, ldb, epi, bias, H);
  else if (ta && !tb)
    hipLaunchKernelGGL((k_gemm_bf16<true, false>), grid, dim3(256), 0, m->stream, A, B, out, M,
                       N, K, kps, lda, ldb, epi, bias, H);
  else if (!ta && tb)
    hipLaunchKernelGGL((k_gemm_bf16<false, true>), grid, dim3(256), 0, m->stream, A, B, out, M,
                       N, K, kps, lda, ldb, epi, bias, H);
  else
    return NB_E_INVALID;
  NB_MHIP(m, hipGetLastError());
  if (splits > 1) {
    const int64_t cnt = (int64_t)M * N;
    hipLaunchKernelGGL(k_splitk_sum, dim3(nblk(cnt)), dim3(256), 0, m->stream, m->splitbuf, C,
                       cnt, splits);
    NB_MHIP(m, hipGetLastError());
  }
  return NB_OK;
}

void free_rows(nb_mlp *m) {
  void *p[] = {m->Xb, m->Y, m->T, m->dY, m->dZ, m->dH, m->lrow, m->dYb, m->dZb, m->xin};
  for (void *q : p)
    if (q) (void)hipFree(q);
  for (float *q : m->Z) (void)hipFree(q);
  for (uint16_t *q : m->Hb) (void)hipFree(q);
  m->Z.clear();
  m->Hb.clear();
  m->Xb = nullptr;
  m->Y = m->T = m->dY = m->dZ = m->dH = m->lrow = m->xin = nullptr;
  m->dYb = m->dZb = nullptr;
  m->rows_cap = 0;
}

int ensure_rows(nb_mlp *m, int64_t rows) {
  if (rows <= m->rows_cap) return NB_OK;
  free_rows(m);
  const int64_t R = round_up(std::max<int64_t>(rows, 64), 64);
  const int hp = m->hid;
  NB_MHIP(m, hipMalloc((void **)&m->Xb, (size_t)R * m->dinp * 2));
  NB_MHIP(m, hipMalloc((void **)&m->xin, (size_t)R * m->din * 4));
  for (int l = 0; l < m->layers; l++) {
    float *z;
    uint16_t *h;
    NB_MHIP(m, hipMalloc((void **)&z, (size_t)R * hp * 4));
    NB_MHIP(m, hipMalloc((void **)&h, (size_t)R * hp * 2));
    m->Z.push_back(z);
    m->Hb.push_back(h);
  }
  const int wmax = std::max(hp, m->doutp);
  NB_MHIP(m, hipMalloc((void **)&m->Y, (size_t)R * m->doutp * 4));
  NB_MHIP(m, hipMalloc((void **)&m->T, (size_t)R * m->dout * 4));
  NB_MHIP(m, hipMalloc((void **)&m->dY, (size_t)R * m->doutp * 4));
  NB_MHIP(m, hipMalloc((void **)&m->dYb, (size_t)R * wmax * 2));
  NB_MHIP(m, hipMalloc((void **)&m->dZ, (size_t)R * wmax * 4));
  NB_MHIP(m, hipMalloc((void **)&m->dZb, (size_t)R * wmax * 2));
  NB_MHIP(m, hipMalloc((void **)&m->dH, (size_t)R * std::max(wmax, m->dinp) * 4));
  NB_MHIP(m, hipMalloc((void **)&m->lrow, (size_t)R * 4));
  m->rows_cap = R;
  return NB_OK;
}

// Forward pass over `rows` (multiple of 64) already in Xb; output FP32 in Y [rows x doutp].
int forward(nb_mlp *m, int64_t rows) {
  const uint16_t *in = m->Xb;
  int ld = m->dinp;
  for (int l = 0; l < m->layers; l++) {
    int rc = gemm(m, false, false, in, m->Wb[l], m->Z[l], (int)rows, m->hid, ld, ld, m->hid,
                  EPI_BIAS_GELU, m->bvec[l], m->Hb[l], 1);
    if (rc) return rc;
    in = m->Hb[l];
    ld = m->hid;
  }
  return gemm(m, false, false, in, m->Wb[m->layers], m->Y, (int)rows, m->doutp, ld, ld,
              m->doutp, EPI_BIAS, m->bvec[m->layers], nullptr, 1);
}

// Backward pass of the MSE loss for `rows`; gradients into Wg / bg.  Targets in T.
int backward(nb_mlp *m, int64_t rows) {
  const int R = (int)rows;
  hipLaunchKernelGGL(k_mse, dim3(nblk(R)), dim3(256), 0, m->stream, m->Y, m->doutp, m->T, R,
                     m->dout, m->dY, m->dYb, m->lrow);
  NB_MHIP(m, hipGetLastError());
  const int splits = std::max(1, std::min(64, R / 256));
  // output layer
  const int L = m->layers;
  const uint16_t *inL = L > 0 ? m->Hb[L - 1] : m->Xb;
  const int ldL = L > 0 ? m->hid : m->dinp;
  int rc = gemm(m, true, false, inL, m->dYb, m->Wg[L], ldL, m->doutp, R, ldL, m->doutp,
                EPI_STORE, nullptr, nullptr, splits);
  if (rc) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(nblk(m->doutp, 64)), dim3(1024), 0, m->stream, m->dY, R,
                     m->doutp, m->bg[L]);
  NB_MHIP(m, hipGetLastError());
  const uint16_t *gradb = m->dYb;  // BF16 gradient w.r.t. this layer's output
  int gw = m->doutp;
  for (int l = L - 1; l >= 0; l--) {
    // dH_l = grad * W_{l+1}^T   [R x hid]
    rc = gemm(m, false, true, gradb, m->Wb[l + 1], m->dH, R, m->hid, gw, gw, gw, EPI_STORE,
              nullptr, nullptr, 1);
    if (rc) return rc;
    const int64_t cnt = (int64_t)R * m->hid;
    hipLaunchKernelGGL(k_gelu_back, dim3(nblk(cnt)), dim3(256), 0, m->stream, m->dH, m->Z[l],
                       m->dZ, m->dZb, cnt);
    NB_MHIP(m, hipGetLastError());
    const uint16_t *in = l > 0 ? m->Hb[l - 1] : m->Xb;
    const int ld = l > 0 ? m->hid : m->dinp;
    rc = gemm(m, true, false, in, m->dZb, m->Wg[l], ld, m->hid, R, ld, m->hid, EPI_STORE,
              nullptr, nullptr, splits);
    if (rc) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(m->hid, 64)), dim3(1024), 0, m->stream, m->dZ, R,
                       m->hid, m->bg[l]);
    NB_MHIP(m, hipGetLastError());
    gradb = m->dZb;
    gw = m->hid;
  }
  return NB_OK;
}

int adam_step(nb_mlp *m, float lr) {
  m->adam_t++;
  m->wver++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)m->adam_t), c2 = 1.f - powf(b2, (float)m->adam_t);
  for (int l = 0; l <= m->layers; l++) {
    const int64_t nw = (int64_t)m->in_dim[l] * m->out_dim[l];
    hipLaunchKernelGGL(k_adam, dim3(nblk(nw)), dim3(256), 0, m->stream, m->W[l], m->Wg[l],
                       m->Wm1[l], m->Wm2[l], m->Wb[l], nw, lr, b1, b2, 1e-8f, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(nblk(m->out_dim[l])), dim3(256), 0, m->stream, m->bvec[l],
                       m->bg[l], m->bm1[l], m->bm2[l], (uint16_t *)nullptr,
                       (int64_t)m->out_dim[l], lr, b1, b2, 1e-8f, c1, c2);
    NB_MHIP(m, hipGetLastError());
  }
  return NB_OK;
}

// Upload FP32 host rows [rows x din] and convert/pad to Xb.
int load_inputs(nb_mlp *m, const float *x, int64_t rows, int64_t rows_pad) {
  NB_MHIP(m, hipMemsetAsync(m->xin, 0, (size_t)rows_pad * m->din * 4, m->stream));
  NB_MHIP(m, hipMemcpyAsync(m->xin, x, (size_t)rows * m->din * 4, hipMemcpyHostToDevice,
                            m->stream));
  hipLaunchKernelGGL(k_pad_rows_bf16, dim3(nblk(rows_pad * m->dinp)), dim3(256), 0, m->stream,
                     m->xin, (int)rows_pad, m->din, m->din, m->Xb, m->dinp);
  NB_MHIP(m, hipGetLastError());
  return NB_OK;
}

}  // namespace

// Exposed to nb_sur.hip (same library): run the forward pass on device-resident FP32 features.
namespace nb {
int mlp_forward_device(nb_mlp *m, const float *x_dev, int64_t rows, const float **y_dev,
                       int *ldy, hipStream_t st) {
  hipStream_t keep = m->stream;
  m->stream = st;
  int rc = ensure_rows(m, rows);
  const int64_t rp = round_up(std::max<int64_t>(rows, 64), 64);
  if (!rc) {
    hipLaunchKernelGGL(k_pad_rows_bf16, dim3(nblk(rp * m->dinp)), dim3(256), 0, st, x_dev,
                       (int)rows, m->din, m->din, m->Xb, m->dinp);
    rc = forward(m, rp);
  }
  m->stream = keep;
  *y_dev = m->Y;
  *ldy = m->doutp;
  return rc;
}

// One Adam step on device-resident rows (features FP32 [rows x din], targets FP32 [rows x dout]).
int mlp_train_step_device(nb_mlp *m, const float *x_dev, const float *t_dev, int64_t rows,
                          float lr, hipStream_t st) {
  hipStream_t keep = m->stream;
  m->stream = st;
  int rc = ensure_rows(m, rows);
  if (!rc && rows % 64) rc = NB_E_INVALID;
  if (!rc) {
    hipLaunchKernelGGL(k_pad_rows_bf16, dim3(nblk(rows * m->dinp)), dim3(256), 0, st, x_dev,
                       (int)rows, m->din, m->din, m->Xb, m->dinp);
    if (hipMemcpyAsync(m->T, t_dev, (size_t)rows * m->dout * 4, hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      rc = NB_E_HIP;
  }
  if (!rc) rc = forward(m, rows);
  if (!rc) rc = backward(m, rows);
  if (!rc) rc = adam_step(m, lr);
  m->stream = keep;
  return rc;
}
// Fused inference on device-resident FP32 rows [rows x din] into y [rows x dout].
int mlp_infer_device(nb_mlp *m, const float *x_dev, int64_t rows, float *y_dev,
                     hipStream_t st) {
  if (m->dinp > kInferW || m->hid > kInferW || m->hid % 64 || m->dinp % 64 || m->dout > 16 ||
      m->layers > kInferMaxLayers || m->layers < 1) {
    m->err = "fused inference: shape beyond its limits";
    return NB_E_UNSUPPORTED;
  }
  if (rows <= 0) return NB_OK;
  if (m->wt_ver != m->wver) {
    if (m->WbT.empty())
      for (int l = 0; l <= m->layers; l++) {
        uint16_t *q = nullptr;
        NB_MHIP(m, hipMalloc((void **)&q, (size_t)m->in_dim[l] * m->out_dim[l] * 2));
        m->WbT.push_back(q);
      }
    for (int l = 0; l <= m->layers; l++) {
      const int K = m->in_dim[l], N = m->out_dim[l];
      hipLaunchKernelGGL(k_transpose_bf16, dim3(nblk((int64_t)K * N)), dim3(256), 0, st,
                         m->Wb[l], K, N, m->WbT[l]);
    }
    NB_MHIP(m, hipGetLastError());
    m->wt_ver = m->wver;
  }
  InferArgs a;
  a.x = x_dev;
  a.rows = (int)rows;
  a.din = m->din;
  a.dinp = m->dinp;
  a.hid = m->hid;
  a.layers = m->layers;
  a.dout = m->dout;
  for (int l = 0; l <= kInferMaxLayers; l++) {
    a.wt[l] = l <= m->layers ? m->WbT[l] : nullptr;
    a.b[l] = l <= m->layers ? m->bvec[l] : nullptr;
  }
  a.y = y_dev;
  const int key = (m->dinp == 64 ? 0 : 1) + (m->hid == 64 ? 0 : 2);
  if (m->layers <= 3) {
    // persistent: as many workgroups as stay resident, each walking tiles
    const void *fn = nullptr;
#define NB_INFER_P(D, H)                                                         \
  fn = m->layers == 1   ? (const void *)k_mlp_infer_p<D, H, 1>                    \
       : m->layers == 2 ? (const void *)k_mlp_infer_p<D, H, 2>                    \
                        : (const void *)k_mlp_infer_p<D, H, 3>
    switch (key) {
      case 0: NB_INFER_P(64, 64); break;
      case 1: NB_INFER_P(128, 64); break;
      case 2: NB_INFER_P(64, 128); break;
      default: NB_INFER_P(128, 128); break;
    }
#undef NB_INFER_P
    if (m->infer_wgs == 0) {
      int cus = 0, per = 0;
      NB_MHIP(m, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m->device));
      NB_MHIP(m, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 256, 0));
      m->infer_wgs = std::max(1, cus * std::max(1, per));
    }
    const int64_t tiles = (rows + 63) / 64;
    const dim3 grid((unsigned)std::min<int64_t>(tiles, m->infer_wgs));
    void *args[] = {&a};
    NB_MHIP(m, hipLaunchKernel(fn, grid, dim3(256), args, 0, st));
    return NB_OK;
  }
  const dim3 grid((unsigned)((rows + 63) / 64));
  switch (key) {
    case 0: hipLaunchKernelGGL((k_mlp_infer<64, 64>), grid, dim3(256), 0, st, a); break;
    case 1: hipLaunchKernelGGL((k_mlp_infer<128, 64>), grid, dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_mlp_infer<64, 128>), grid, dim3(256), 0, st, a); break;
    default: hipLaunchKernelGGL((k_mlp_infer<128, 128>), grid, dim3(256), 0, st, a); break;
  }
  NB_MHIP(m, hipGetLastError());
  return NB_OK;
}

int mlp_dims(const nb_mlp *m, int *din, int *dout) {
  *din = m->din;
  *dout = m->dout;
  return NB_OK;
}
}  // namespace nb

extern "C" {

int nb_mlp_create(int32_t device, int32_t din, int32_t hidden, int32_t layers, int32_t dout,
                  uint64_t seed, nb_mlp **out) {
  if (!out) return NB_E_INVALID;
  *out = nullptr;
  if (din < 1 || din > 1024 || hidden < 64 || hidden % 64 || hidden > 1024 || layers < 1 ||
      layers > 8 || dout < 1 || dout > 64)
    return NB_E_INVALID;
  int ndev = 0;
  // a HIP failure is reported with its error string through nb_last_error(NULL)
  auto hip_fail = [](const char *what, hipError_t e) {
    static thread_local std::string msg;
    msg = std::string("nb_mlp_create: ") + what + ": " + hipGetErrorString(e);
    set_global_error(msg.c_str());
    return NB_E_HIP;
  };
  {
    const hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return hip_fail("hipGetDeviceCount", e);
    if (device < 0 || device >= ndev) return hip_fail("device", hipErrorInvalidDevice);
  }
  nb_mlp *m = new (std::nothrow) nb_mlp();
  if (!m) return NB_E_NOMEM;
  m->device = device;
  m->din = din;
  m->dinp = (int)round_up(din, 64);  // the layer-0 weight gradient has M = dinp
  m->hid = hidden;
  m->layers = layers;
  m->dout = dout;
  m->doutp = 64;
  {
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete m;
      return hip_fail("stream", e);
    }
  }
  std::mt19937_64 rng(seed);
  for (int l = 0; l <= layers; l++) {
    const int in = l == 0 ? m->dinp : hidden;
    const int o = l == layers ? m->doutp : hidden;
    const int in_real = l == 0 ? din : hidden;
    const int o_real = l == layers ? dout : hidden;
    m->in_dim.push_back(in);
    m->out_dim.push_back(o);
    // Glorot-uniform on the real fan-in/out; padded rows/columns stay zero
    std::vector<float> w((size_t)in * o, 0.f), z((size_t)std::max(in * o, o), 0.f);
    const double lim = std::sqrt(6.0 / (in_real + o_real));
    std::uniform_real_distribution<double> U(-lim, lim);
    for (int i = 0; i < in_real; i++)
      for (int j = 0; j < o_real; j++) w[(size_t)i * o + j] = (float)U(rng);
    float *W, *Wg, *M1, *M2, *b, *bg, *bm1, *bm2;
    uint16_t *Wb;
    const size_t nw = (size_t)in * o;
    bool ok = hipMalloc((void **)&W, nw * 4) == hipSuccess &&
              hipMalloc((void **)&Wg, nw * 4) == hipSuccess &&
              hipMalloc((void **)&M1, nw * 4) == hipSuccess &&
              hipMalloc((void **)&M2, nw * 4) == hipSuccess &&
              hipMalloc((void **)&Wb, nw * 2) == hipSuccess &&
              hipMalloc((void **)&b, o * 4) == hipSuccess &&
              hipMalloc((void **)&bg, o * 4) == hipSuccess &&
              hipMalloc((void **)&bm1, o * 4) == hipSuccess &&
              hipMalloc((void **)&bm2, o * 4) == hipSuccess;
    if (!ok) {
      nb_mlp_destroy(m);
      return NB_E_NOMEM;
    }
    m->W.push_back(W);
    m->Wg.push_back(Wg);
    m->Wm1.push_back(M1);
    m->Wm2.push_back(M2);
    m->Wb.push_back(Wb);
    m->bvec.push_back(b);
    m->bg.push_back(bg);
    m->bm1.push_back(bm1);
    m->bm2.push_back(bm2);
    (void)hipMemcpy(W, w.data(), nw * 4, hipMemcpyHostToDevice);
    (void)hipMemset(M1, 0, nw * 4);
    (void)hipMemset(M2, 0, nw * 4);
    (void)hipMemset(Wg, 0, nw * 4);
    (void)hipMemset(b, 0, o * 4);
    (void)hipMemset(bg, 0, o * 4);
    (void)hipMemset(bm1, 0, o * 4);
    (void)hipMemset(bm2, 0, o * 4);
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk((int64_t)nw)), dim3(256), 0, m->stream, W, Wb,
                       (int64_t)nw);
  }
  {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
      nb_mlp_destroy(m);
      return hip_fail("weight conversion", e);
    }
  }
  *out = m;
  return NB_OK;
}

void nb_mlp_destroy(nb_mlp *m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  free_rows(m);
  for (auto *v : {&m->W, &m->Wg, &m->Wm1, &m->Wm2, &m->bvec, &m->bg, &m->bm1, &m->bm2})
    for (float *q : *v) (void)hipFree(q);
  for (uint16_t *q : m->Wb) (void)hipFree(q);
  for (uint16_t *q : m->WbT) (void)hipFree(q);
  if (m->splitbuf) (void)hipFree(m->splitbuf);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

const char *nb_mlp_last_error(const nb_mlp *m) { return m ? m->err.c_str() : "NULL mlp"; }

int nb_mlp_infer(nb_mlp *m, const float *x, int64_t rows, float *y) {
  if (!m || (rows > 0 && (!x || !y)) || rows < 0) return NB_E_INVALID;
  if (rows == 0) return NB_OK;
  NB_MHIP(m, hipSetDevice(m->device));
  float *dx = nullptr, *dy = nullptr;
  NB_MHIP(m, hipMalloc((void **)&dx, (size_t)rows * m->din * 4));
  hipError_t e = hipMalloc((void **)&dy, (size_t)rows * m->dout * 4);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dx, x, (size_t)rows * m->din * 4, hipMemcpyHostToDevice, m->stream);
  int rc = e == hipSuccess ? mlp_infer_device(m, dx, rows, dy, m->stream) : NB_E_HIP;
  if (!rc) {
    e = hipMemcpyAsync(y, dy, (size_t)rows * m->dout * 4, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) rc = NB_E_HIP;
  }
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (rc == NB_E_HIP && m->err.empty()) m->err = "nb_mlp_infer: HIP error";
  return rc;
}

// Parameters in the unpadded layout: for each layer W [in_real x out_real] row-major, then b.
int nb_mlp_param_count(const nb_mlp *m, int64_t *count) {
  if (!m || !count) return NB_E_INVALID;
  int64_t c = 0;
  for (int l = 0; l <= m->layers; l++) {
    const int in = l == 0 ? m->din : m->hid, o = l == m->layers ? m->dout : m->hid;
    c += (int64_t)in * o + o;
  }
  *count = c;
  return NB_OK;
}

static int xfer_params(nb_mlp *m, float *host, int64_t count, bool to_dev, bool grads) {
  int64_t need = 0;
  nb_mlp_param_count(m, &need);
  if (count != need) {
    m->err = "parameter count mismatch";
    return NB_E_INVALID;
  }
  NB_MHIP(m, hipSetDevice(m->device));
  NB_MHIP(m, hipStreamSynchronize(m->stream));
  int64_t off = 0;
  for (int l = 0; l <= m->layers; l++) {
    const int in = l == 0 ? m->din : m->hid, o = l == m->layers ? m->dout : m->hid;
    const int op = m->out_dim[l];
    float *W = grads ? m->Wg[l] : m->W[l];
    float *b = grads ? m->bg[l] : m->bvec[l];
    std::vector<float> full((size_t)m->in_dim[l] * op, 0.f), bf((size_t)op, 0.f);
    if (to_dev) {
      for (int i = 0; i < in; i++)
        for (int j = 0; j < o; j++) full[(size_t)i * op + j] = host[off + (int64_t)i * o + j];
      for (int j = 0; j < o; j++) bf[j] = host[off + (int64_t)in * o + j];
      NB_MHIP(m, hipMemcpy(W, full.data(), full.size() * 4, hipMemcpyHostToDevice));
      NB_MHIP(m, hipMemcpy(b, bf.data(), bf.size() * 4, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk((int64_t)full.size())), dim3(256), 0,
                         m->stream, W, m->Wb[l], (int64_t)full.size());
      NB_MHIP(m, hipStreamSynchronize(m->stream));
    } else {
      NB_MHIP(m, hipMemcpy(full.data(), W, full.size() * 4, hipMemcpyDeviceToHost));
      NB_MHIP(m, hipMemcpy(bf.data(), b, bf.size() * 4, hipMemcpyDeviceToHost));
      for (int i = 0; i < in; i++)
        for (int j = 0; j < o; j++) host[off + (int64_t)i * o + j] = full[(size_t)i * op + j];
      for (int j = 0; j < o; j++) host[off + (int64_t)in * o + j] = bf[j];
    }
    off += (int64_t)in * o + o;
  }
  return NB_OK;
}

int nb_mlp_set_params(nb_mlp *m, const float *p, int64_t count) {
  if (!m || !p) return NB_E_INVALID;
  int rc = xfer_params(m, const_cast<float *>(p), count, true, false);
  m->adam_t = 0;
  m->wver++;
  return rc;
}

int nb_mlp_get_params(nb_mlp *m, float *p, int64_t count) {
  if (!m || !p) return NB_E_INVALID;
  return xfer_params(m, p, count, false, false);
}

int nb_mlp_forward(nb_mlp *m, const float *x, int64_t rows, float *y) {
  if (!m || (rows > 0 && (!x || !y)) || rows < 0) return NB_E_INVALID;
  if (rows == 0) return NB_OK;
  NB_MHIP(m, hipSetDevice(m->device));
  const int64_t rp = round_up(rows, 64);
  int rc = ensure_rows(m, rp);
  if (!rc) rc = load_inputs(m, x, rows, rp);
  if (!rc) rc = forward(m, rp);
  if (rc) return rc;
  std::vector<float> Y((size_t)rp * m->doutp);
  NB_MHIP(m, hipMemcpyAsync(Y.data(), m->Y, Y.size() * 4, hipMemcpyDeviceToHost, m->stream));
  NB_MHIP(m, hipStreamSynchronize(m->stream));
  for (int64_t r = 0; r < rows; r++)
    for (int c = 0; c < m->dout; c++) y[r * m->dout + c] = Y[(size_t)r * m->doutp + c];
  return NB_OK;
}

// Gradient of the MSE loss over `rows` (a multiple of 64) w.r.t. the parameters (unpadded layout).
int nb_mlp_grad(nb_mlp *m, const float *x, const float *t, int64_t rows, float *grad,
                float *loss) {
  if (!m || !x || !t || !grad || rows <= 0 || rows % 64) return NB_E_INVALID;
  NB_MHIP(m, hipSetDevice(m->device));
  int rc = ensure_rows(m, rows);
  if (!rc) rc = load_inputs(m, x, rows, rows);
  if (rc) return rc;
  NB_MHIP(m, hipMemcpyAsync(m->T, t, (size_t)rows * m->dout * 4, hipMemcpyHostToDevice,
                            m->stream));
  rc = forward(m, rows);
  if (!rc) rc = backward(m, rows);
  if (rc) return rc;
  if (loss) {
    std::vector<float> l((size_t)rows);
    NB_MHIP(m, hipMemcpyAsync(l.data(), m->lrow, l.size() * 4, hipMemcpyDeviceToHost,
                              m->stream));
    NB_MHIP(m, hipStreamSynchronize(m->stream));
    double s = 0;
    for (float v : l) s += v;
    *loss = (float)(s / ((double)rows * m->dout));
  }
  int64_t cnt = 0;
  nb_mlp_param_count(m, &cnt);
  return xfer_params(m, grad, cnt, false, true);
}

// Adam training on host-provided rows: `steps` minibatches of `batch` rows (multiple of 64) drawn
// in order (wrapping) from the data.  Returns the mean loss of the last 10% of steps.
int nb_mlp_train(nb_mlp *m, const float *x, const float *t, int64_t rows, int32_t steps,
                 int32_t batch, float lr, float *loss_out) {
  if (!m || !x || !t || rows < batch || batch <= 0 || batch % 64 || steps < 0)
    return NB_E_INVALID;
  NB_MHIP(m, hipSetDevice(m->device));
  float *dx = nullptr, *dt = nullptr;
  NB_MHIP(m, hipMalloc((void **)&dx, (size_t)rows * m->din * 4));
  if (hipMalloc((void **)&dt, (size_t)rows * m->dout * 4) != hipSuccess) {
    (void)hipFree(dx);
    return NB_E_NOMEM;
  }
  int rc = NB_OK;
  if (hipMemcpy(dx, x, (size_t)rows * m->din * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dt, t, (size_t)rows * m->dout * 4, hipMemcpyHostToDevice) != hipSuccess)
    rc = NB_E_HIP;
  double lsum = 0;
  int lcount = 0;
  const int64_t nb = rows / batch;
  std::vector<float> l((size_t)batch);
  for (int s = 0; s < steps && !rc; s++) {
    const int64_t b0 = (s % nb) * batch;
    rc = nb::mlp_train_step_device(m, dx + b0 * m->din, dt + b0 * m->dout, batch, lr, m->stream);
    if (!rc && s >= steps - std::max(1, steps / 10)) {
      if (hipMemcpyAsync(l.data(), m->lrow, l.size() * 4, hipMemcpyDeviceToHost, m->stream) !=
              hipSuccess ||
          hipStreamSynchronize(m->stream) != hipSuccess) {
        rc = NB_E_HIP;
        break;
      }
      double q = 0;
      for (float v : l) q += v;
      lsum += q / ((double)batch * m->dout);
      lcount++;
    }
  }
  (void)hipStreamSynchronize(m->stream);
  (void)hipFree(dx);
  (void)hipFree(dt);
  if (loss_out) *loss_out = lcount ? (float)(lsum / lcount) : 0.f;
  return rc;
}

}  // extern "C"
