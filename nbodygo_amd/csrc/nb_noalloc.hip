// nb_noalloc.hip -- the softened FP32 force phase for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// nb_forces_noalloc computes what nb_compute_forces computes on a context with the same G and
// eps, no overlap cutoff, no collisions and the shape {ipl 1, tile_j 512, slices min(4, tiles),
// block 256, LDS-staged}: the FP32 source stream, the per-slice Kahan-tiled accelerations of
// k_force_lds and their FP64 fold into forces m_i * (G * a_i), with the same kernels and the same
// operation order, so the forces are bit-identical.  Inputs and outputs are the caller's device
// arrays; the two intermediate buffers (the source stream and the per-slice partials) live in one
// file-scope device array of fixed size, so the call allocates nothing.
//
// Each device holds its own copy of that array from module load until the process ends.  Its
// address on a device is resolved once, on the first call made with that device current
// (hipGetSymbolAddress), and kept per device; kernels only ever receive that address.  Every call
// on a device shares the array: calls that overlap on one device (two streams, two threads)
// race.  Every byte a call reads from it is written earlier in the same call, so nothing needs
// clearing between calls.
#include <cmath>
#include <cstdio>
#include <mutex>

#include "nb_internal.hpp"

namespace nb {

constexpr int64_t kNaMaxBodies = NB_NOALLOC_MAX_BODIES;  // padded body count limit
constexpr int kNaSplit = 4;                               // source-range slices (at most)
constexpr int kNaTile = 512;                              // sources per Kahan tile
constexpr int kNaBlock = 256;                             // threads per workgroup
constexpr size_t kNaStreamBytes = (size_t)kNaMaxBodies * sizeof(float4);
constexpr size_t kNaAccBytes = (size_t)kNaSplit * 3 * kNaMaxBodies * sizeof(double);
constexpr size_t kNaScratchBytes = kNaStreamBytes + kNaAccBytes;  // 28 MiB
static_assert(kNaMaxBodies % kPad == 0, "the body limit is a multiple of the padding");
static_assert(kNaStreamBytes % 256 == 0, "the partials start 256-byte aligned");

// The scratch: [0, kNaStreamBytes) the float4 source stream, then [split][3][npad] doubles.
static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mu;
static unsigned char *g_na_ptr[kNaMaxDevices];  // scratch's device address, per device

// Address of the scratch on `device` (the current device), resolved on first use.
static hipError_t na_scratch(int device, unsigned char **out) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_na_ptr[device] = static_cast<unsigned char *>(p);
  }
  *out = g_na_ptr[device];
  return hipSuccess;
}

// The FP32 source stream exactly as the context path writes it: (x, y, z, m) of active bodies,
// massless sentinels for inactive and padding entries.  flags == nullptr: every body is active.
__global__ __launch_bounds__(256) void k_na_stream(const double *x, const double *y,
                                                   const double *z, const double *m,
                                                   const uint8_t *flags, float4 *js, int n,
                                                   int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  bool act = i < n;
  if (act && flags) act = (flags[i] & NB_F_EXISTS) && !(flags[i] & NB_F_FRAGMENTING);
  js[i] = act ? make_float4((float)x[i], (float)y[i], (float)z[i], (float)m[i])
              : make_float4(kSentinel, kSentinel, kSentinel, 0.f);
}

// Slices summed in order, then m_i * (G * a): the context path's fold (k_combine_force), written
// to three separate outputs.
__global__ __launch_bounds__(256) void k_na_fold(const double *acc, int split, int npad, int n,
                                                 const double *m, double G, double *fx,
                                                 double *fy, double *fz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[3] = {0, 0, 0};
  for (int s = 0; s < split; s++)
    for (int c = 0; c < 3; c++) f[c] += acc[((size_t)s * 3 + c) * npad + i];
  const double mi = m[i];
  fx[i] = mi * (G * f[0]);
  fy[i] = mi * (G * f[1]);
  fz[i] = mi * (G * f[2]);
}

static int na_fail(int code, const char *msg) {
  set_global_error(msg);
  return code;
}

static int na_hip(hipError_t e, const char *what) {
  char buf[256];
  snprintf(buf, sizeof buf, "nb_forces_noalloc: HIP error %d (%s) in %s", (int)e,
           hipGetErrorString(e), what);
  return na_fail(NB_E_HIP, buf);
}

}  // namespace nb

using namespace nb;

extern "C" int nb_forces_noalloc(int32_t device, int64_t n, const double *x, const double *y,
                                 const double *z, const double *m, const uint8_t *flags,
                                 double G, double eps, double *fx, double *fy, double *fz,
                                 void *stream) {
  if (n < 0 || n > kNaMaxBodies)
    return na_fail(NB_E_INVALID, "nb_forces_noalloc: n outside [0, NB_NOALLOC_MAX_BODIES]");
  if (!(eps > 0.0) || !std::isfinite(eps) || !std::isfinite(G))
    return na_fail(NB_E_INVALID, "nb_forces_noalloc: eps must be finite and > 0, G finite");
  if ((float)(eps * eps) < 1e-30f)
    return na_fail(NB_E_INVALID, "nb_forces_noalloc: eps^2 underflows FP32");
  if (n > 0 && (!x || !y || !z || !m || !fx || !fy || !fz))
    return na_fail(NB_E_INVALID, "nb_forces_noalloc: NULL array");
  if (device < 0 || device >= kNaMaxDevices)
    return na_fail(NB_E_INVALID, "nb_forces_noalloc: device ordinal out of range");
  if (n == 0) return NB_OK;
  int ndev = 0, prev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess) return na_hip(e, "hipGetDeviceCount");
  if (device >= ndev) return na_fail(NB_E_INVALID, "nb_forces_noalloc: no such HIP device");
  e = hipGetDevice(&prev);
  if (e != hipSuccess) return na_hip(e, "hipGetDevice");
  if (prev != device && (e = hipSetDevice(device)) != hipSuccess) return na_hip(e, "hipSetDevice");
  int rc = NB_OK;
  unsigned char *base = nullptr;
  e = na_scratch(device, &base);
  if (e != hipSuccess) {
    rc = na_hip(e, "hipGetSymbolAddress(scratch)");
  } else {
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t npad = round_up(n, kPad);
    float4 *js = reinterpret_cast<float4 *>(base);
    double *acc = reinterpret_cast<double *>(base + kNaStreamBytes);
    ForceShape s;
    s.ipl = 1;
    s.tile_j = kNaTile;
    s.block = kNaBlock;
    s.variant = 2;
    const int tiles = (int)(npad / kNaTile);
    s.jsplit = tiles < kNaSplit ? tiles : kNaSplit;
    ForceArgs a{};
    a.js = js;
    a.acc = acc;
    a.npad = (int)npad;
    a.n = (int)n;
    a.tiles = tiles;
    a.tiles_per_slice = (tiles + s.jsplit - 1) / s.jsplit;
    a.tile_j = kNaTile;
    a.eps2 = (float)(eps * eps);
    a.detect = 0;
    a.drop_flagged = 0;
    hipLaunchKernelGGL(k_na_stream, dim3((unsigned)(npad / 256)), dim3(256), 0, st, x, y, z, m,
                       flags, js, (int)n, (int)npad);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_force_f32(s, a, st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_na_fold, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, acc,
                         s.jsplit, (int)npad, (int)n, m, G, fx, fy, fz);
      e = hipGetLastError();
    }
    if (e != hipSuccess) rc = na_hip(e, "kernel launch");
  }
  if (prev != device) (void)hipSetDevice(prev);
  return rc;
}
