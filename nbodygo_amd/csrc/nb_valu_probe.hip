// nb_valu_probe.hip -- FP32 VALU issue rates on gfx950, the measurements behind the force
// kernel's ceiling: per-op throughput in lane-ops per clock per CU for independent chains whose
// operands are all VGPRs (or, for "sgpr", one SGPR operand), and for the mixes the pairwise
// kernel issues, then the matrix-core (MFMA) rates alone and beside VALU work.  Every kernel
// reads two 64-float inputs and writes one float per thread: all accesses are in bounds, and the
// loop trip count is fixed.
// Usage: nb_valu_probe
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                                      \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);        \
      return 1;                                                                    \
    }                                                                              \
  } while (0)

constexpr int kChains = 16;  // independent dependency chains per lane
constexpr int kIters = 2048;

enum Op {
  FMAC,       // x = fma(y, z, x)
  FMA_VOP3,   // x = fma(x, y, z)
  MUL,        // x = x * y
  ADD,        // x = x + y
  FMAC_MUL,   // 1:1
  FMAC_ADD,   // 1:1
  FMAC_RSQ,   // 3:1, the transcendental beside plain ops
  MUL_ADD,    // 1:1
  RSQ,        // v_rsq_f32 only
  SUB_FMA_MUL,// 1:1:1
  SGPR_FMAC,  // fma with a wave-uniform (SGPR) multiplier
  NOPS
};
static const char *kName[NOPS] = {"fmac",      "fma (vop3)", "mul",        "add",
                                  "fmac+mul",  "fmac+add",   "3 fmac+rsq", "mul+add",
                                  "rsq",       "sub+fmac+mul", "fmac, sgpr operand"};

template <int OP>
__global__ __launch_bounds__(256) void k_probe(float *out, const float *in, float s_uniform) {
  float x[kChains];
  const float y = in[threadIdx.x & 63], z = in[64 + (threadIdx.x & 63)];
#pragma unroll
  for (int c = 0; c < kChains; c++) x[c] = threadIdx.x * 1e-3f + c + 1.0f;
  for (int i = 0; i < kIters; i++) {
#pragma unroll
    for (int c = 0; c < kChains; c++) {
      if (OP == FMAC) x[c] = __builtin_fmaf(y, z, x[c]);
      if (OP == FMA_VOP3) x[c] = __builtin_fmaf(x[c], y, z);
      if (OP == MUL) x[c] = x[c] * y;
      if (OP == ADD) x[c] = x[c] + y;
      if (OP == FMAC_MUL) x[c] = (c & 1) ? x[c] * y : __builtin_fmaf(y, z, x[c]);
      if (OP == FMAC_ADD) x[c] = (c & 1) ? x[c] + y : __builtin_fmaf(y, z, x[c]);
      if (OP == FMAC_RSQ)
        x[c] = (c & 3) == 3 ? __builtin_amdgcn_rsqf(x[c]) : __builtin_fmaf(y, z, x[c]);
      if (OP == MUL_ADD) x[c] = (c & 1) ? x[c] + y : x[c] * z;
      if (OP == RSQ) x[c] = __builtin_amdgcn_rsqf(x[c]);
      if (OP == SUB_FMA_MUL)
        x[c] = (c % 3 == 0) ? x[c] - y : (c % 3 == 1) ? __builtin_fmaf(y, z, x[c]) : x[c] * z;
      if (OP == SGPR_FMAC) x[c] = __builtin_fmaf(s_uniform, z, x[c]);
    }
  }
  float s = 0;
#pragma unroll
  for (int c = 0; c < kChains; c++) s += x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Matrix-core rates and their overlap with the VALU: 8 independent accumulator chains per wave
// of v_mfma_f32_4x4x1_16b_f32 (256 MACs each) or v_mfma_f32_16x16x4_f32 (1024 MACs), alone and
// interleaved with k independent v_fmac_f32 per MFMA.  Whether the two overlap decides whether
// moving part of the pairwise arithmetic onto the matrix cores can pay.
enum MOp { M4, M16, M4_V2, M4_V4, M4_V8, V_ONLY8, M32, M32_V16, M32_V32, V_ONLY32, NMOPS };
static const char *kMName[NMOPS] = {"mfma 4x4x1 (16 blocks)", "mfma 16x16x4",
                                    "mfma 4x4x1 + 2 fmac", "mfma 4x4x1 + 4 fmac",
                                    "mfma 4x4x1 + 8 fmac", "8 fmac (same loop)",
                                    "mfma 32x32x2",        "mfma 32x32x2 + 16 fmac",
                                    "mfma 32x32x2 + 32 fmac", "32 fmac (same loop)"};
// v_mfma_f32_32x32x2_f32 (2048 MACs): 2 accumulator chains of 16 VGPRs, and kV / 2 independent
// v_fmac_f32 beside each MFMA.
template <int OP>
__global__ __launch_bounds__(256) void k_mprobe32(float *out, const float *in) {
  typedef float f16p __attribute__((ext_vector_type(16)));
  constexpr int kV = OP == M32_V16 ? 16 : (OP == M32_V32 || OP == V_ONLY32) ? 32 : 0;
  f16p acc[2];
  float x[kV > 0 ? kV : 1];
  const float y = in[threadIdx.x & 63], z = in[64 + (threadIdx.x & 63)];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[c][e] = (float)(c + e);
#pragma unroll
  for (int v = 0; v < (kV > 0 ? kV : 1); v++) x[v] = threadIdx.x * 1e-3f + v;
  for (int i = 0; i < kIters / 2; i++) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (OP != V_ONLY32) acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(y, z, acc[c], 0, 0, 0);
#pragma unroll
      for (int v = c * kV / 2; v < (c + 1) * kV / 2; v++) x[v] = __builtin_fmaf(y, z, x[v]);
    }
  }
  float s = 0;
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int e = 0; e < 16; e++) s += acc[c][e];
#pragma unroll
  for (int v = 0; v < kV; v++) s += x[v];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}
typedef float f4p __attribute__((ext_vector_type(4)));

template <int OP>
__global__ __launch_bounds__(256) void k_mprobe(float *out, const float *in) {
  constexpr int kV = OP == M4_V2 ? 2 : OP == M4_V4 ? 4 : (OP == M4_V8 || OP == V_ONLY8) ? 8 : 0;
  f4p acc[8];
  float x[8][kV > 0 ? kV : 1];
  const float y = in[threadIdx.x & 63], z = in[64 + (threadIdx.x & 63)];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    acc[c] = f4p{0.f, 0.f, 0.f, (float)c};
#pragma unroll
    for (int v = 0; v < (kV > 0 ? kV : 1); v++) x[c][v] = threadIdx.x * 1e-3f + c + v;
  }
  for (int i = 0; i < kIters / 8; i++) {
#pragma unroll
    for (int c = 0; c < 8; c++) {
      if (OP == M16) acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(y, z, acc[c], 0, 0, 0);
      else if (OP != V_ONLY8) acc[c] = __builtin_amdgcn_mfma_f32_4x4x1f32(y, z, acc[c], 0, 0, 0);
#pragma unroll
      for (int v = 0; v < kV; v++) x[c][v] = __builtin_fmaf(y, z, x[c][v]);
    }
  }
  float s = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
#pragma unroll
    for (int v = 0; v < kV; v++) s += x[c][v];
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <int OP>
static int mrun(float *d, const float *in, int cu, double clk_mhz) {
  const int blocks = cu * 32, threads = 256;
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  float ms = 0;
  for (int rep = 0; rep < 3; rep++) {
    CK(hipEventRecord(e0));
    if (OP >= M32)
      hipLaunchKernelGGL(k_mprobe32<OP>, dim3(blocks), dim3(threads), 0, 0, d, in);
    else
      hipLaunchKernelGGL(k_mprobe<OP>, dim3(blocks), dim3(threads), 0, 0, d, in);
    CK(hipGetLastError());
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
  }
  const double waves = (double)blocks * threads / 64;
  // MFMA instructions per wave: kIters (8 chains x kIters/8, or 2 chains x kIters/2)
  const double mfma = (OP == V_ONLY8 || OP == V_ONLY32) ? 0 : waves * kIters;
  const double macs = mfma * (OP == M16 ? 1024 : OP >= M32 ? 2048 : 256);
  const int kV = OP == M4_V2 ? 2 : OP == M4_V4 ? 4 : (OP == M4_V8 || OP == V_ONLY8) ? 8
               : OP == M32_V16 ? 16 : (OP == M32_V32 || OP == V_ONLY32) ? 32 : 0;
  // fmac lane-ops per wave: kV per MFMA slot for the 4x4x1 loops, kV/2 for the 32x32x2 ones
  const double valu = waves * kIters * (OP >= M32 ? kV / 2 : kV) * 64;
  const double clk = (ms * 1e-3) * cu * clk_mhz * 1e6;
  printf("%-24s %8.3f ms  %6.1f MACs/clk/CU  %6.1f VALU lane-ops/clk/CU\n", kMName[OP], ms,
         macs / clk, valu / clk);
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  return 0;
}

template <int OP>
static int run(float *d, const float *in, int cu, double clk_mhz) {
  const int blocks = cu * 32, threads = 256;
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  float ms = 0;
  for (int rep = 0; rep < 3; rep++) {  // the last of three timed launches
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(k_probe<OP>, dim3(blocks), dim3(threads), 0, 0, d, in, 0.9999f);
    CK(hipGetLastError());
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
  }
  const double ops = (double)blocks * threads * kIters * kChains;
  printf("%-22s %8.3f ms  %6.1f lane-ops/clk/CU\n", kName[OP], ms,
         ops / (ms * 1e-3) / (cu * clk_mhz * 1e6));
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  return 0;
}

int main() {
  hipDeviceProp_t p;
  CK(hipGetDeviceProperties(&p, 0));
  const int cu = p.multiProcessorCount;
  const double clk = p.clockRate / 1e3;  // MHz
  printf("CUs %d, clock %.0f MHz (peak FP32 = %d CU x %.0f MHz x 128 lane-ops x 2 flop)\n", cu,
         clk, cu, clk);
  float *d = nullptr, *in = nullptr;
  CK(hipMalloc(&d, (size_t)cu * 32 * 256 * sizeof(float)));
  float h[128];
  for (int i = 0; i < 128; i++) h[i] = 0.999f + 1e-6f * i;
  CK(hipMalloc(&in, sizeof h));
  CK(hipMemcpy(in, h, sizeof h, hipMemcpyHostToDevice));
  if (run<FMAC>(d, in, cu, clk) || run<FMA_VOP3>(d, in, cu, clk) || run<MUL>(d, in, cu, clk) ||
      run<ADD>(d, in, cu, clk) || run<FMAC_MUL>(d, in, cu, clk) ||
      run<FMAC_ADD>(d, in, cu, clk) || run<FMAC_RSQ>(d, in, cu, clk) ||
      run<MUL_ADD>(d, in, cu, clk) || run<RSQ>(d, in, cu, clk) ||
      run<SUB_FMA_MUL>(d, in, cu, clk) || run<SGPR_FMAC>(d, in, cu, clk))
    return 1;
  if (mrun<M4>(d, in, cu, clk) || mrun<M16>(d, in, cu, clk) || mrun<M4_V2>(d, in, cu, clk) ||
      mrun<M4_V4>(d, in, cu, clk) || mrun<M4_V8>(d, in, cu, clk) ||
      mrun<V_ONLY8>(d, in, cu, clk) || mrun<M32>(d, in, cu, clk) ||
      mrun<M32_V16>(d, in, cu, clk) || mrun<M32_V32>(d, in, cu, clk) ||
      mrun<V_ONLY32>(d, in, cu, clk))
    return 1;
  CK(hipFree(d));
  CK(hipFree(in));
  return 0;
}
