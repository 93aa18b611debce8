// nb_sur.hip -- learned far-field surrogate for the softened force (no counterpart in the
// reference, whose own far-field design, the Barnes-Hut octree of cmd/body/octree.go, is never
// built).  It replaces the O(N^2) force phase with
//
//   exact near field   FP32 pair sums over the 27 cells adjacent to a body's cell
//   multipole shell    per-body monopole + quadrupole of every cell at Chebyshev distance 2..K
//   far expansion      per-cell field and its first three derivatives at the cell's centre of mass
//                      from all cells beyond distance K (monopole + quadrupole field, monopole
//                      derivatives), applied to each body by a third-order Taylor step
//   learned residual   an MLP (BF16 MFMA, nb_mlp.hip) maps scale-free features of the body's cell
//                      and position to the remaining far-field error, trained on direct-sum data
//   fallback           every check_every steps the force of check_samples bodies is recomputed by
//                      direct summation; if the sampled relative error exceeds tol, the step's
//                      force is redone with the plain kernel (a device flag, no host round trip)
//
// Grid: 32^3 cells in stretched coordinates u = d / (|d| + r0), d = x - centre of mass, which map
// all of space into the unit ball, so no body falls outside the grid; cells are about r0/16 wide
// at the centre and grow outward where the density falls.  Bodies are sorted by cell (stable
// radix sort), so the three cells of a grid row are one contiguous run of sources.
//
// All reductions run in a fixed order; the result is deterministic.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "nb_internal.hpp"

namespace nb {

int mlp_infer_device(nb_mlp *m, const float *x_dev, int64_t rows, float *y_dev,
                     hipStream_t st);
int mlp_train_step_device(nb_mlp *m, const float *x_dev, const float *t_dev, int64_t rows,
                          float lr, hipStream_t st);

constexpr int kCellRec = 16;  // COM(3), M, Qxx, Qyy, Qzz, Qxy, Qxz, Qyz, radius, ix, iy, iz, n, 0
constexpr int kFarRec = 36;   // a(3), J(6: xx yy zz xy xz yz), H(10), T(15), 0, 0
constexpr int kFarUsed = 34;  // words of the record the far-field kernel writes
constexpr int kFeat = 30;     // MLP input width
constexpr int kMaxGrid = 64;  // cell keys 0..64^3 (the last for inactive bodies): 19 bits
// Bits of the cell keys 0..ncell sorted by the radix sort (16 up to 40^3 cells).
// Stable radix sort by onesweep passes at every size (rocPRIM's default sorts inputs below a
// million items by block sort plus merge passes: 17 launches for 262,144 keys).
using SortCfg = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                           rocprim::default_config, 0>;

static unsigned key_bits(int ncell) { return 32u - (unsigned)__builtin_clz((unsigned)ncell); }

typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) f4v *cf4p;  // scalar-cache (uniform) loads
typedef const __attribute__((address_space(4))) float *cfp;
typedef const __attribute__((address_space(4))) int32_t *cip;

constexpr int kNearSlots = 64;

struct Sur {
  nb_sur_params p{};
  int G = 32, ncell = 0, K = 2;
  double r0 = 0;
  int64_t cap = 0;
  uint32_t *keys = nullptr, *keys2 = nullptr, *perm_in = nullptr, *perm = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_bytes = 0;
  int32_t *cstart = nullptr, *ccount = nullptr, *nelist = nullptr, *necount = nullptr;
  float4 *sp = nullptr;
  float *crec = nullptr, *crec_ne = nullptr, *fexp = nullptr, *part = nullptr, *feat = nullptr;
  int2 *items = nullptr, *btot = nullptr;
  int32_t *nitems = nullptr;
  int64_t items_cap = 0;
  float4 *base = nullptr;
  double *center = nullptr, *red = nullptr;
  float *mu = nullptr, *sd = nullptr;
  nb_mlp *mlp = nullptr;
  bool trained = false;
  float out_scale = 1.f;
  float *tx = nullptr, *ty = nullptr;
  int64_t trows = 0, tcap = 0;
  int32_t *sidx = nullptr;
  int S = 0;
  double *sexact = nullptr, *serr = nullptr, *spart = nullptr;
  uint32_t *sbody = nullptr;  // per-sample relative error (float bits) of the last check
  uint64_t draws = 0;         // sample draws so far (rotates the samples)
  int64_t spart_len = 0;
  float *ymlp = nullptr;  // MLP output [cap x 3]
  uint32_t *flag = nullptr;
  int64_t steps = 0, fallbacks = 0, checks = 0;
  double last_err = 0;
  float train_loss = 0;
  float *mu0 = nullptr, *sd1 = nullptr;  // identity normalisation (raw features)
  // the item count of the last cycle whose scan has run, written by k_sur_scan into pinned host
  // memory: the near kernel's grid (its workgroups loop over items, so any size is correct)
  int32_t *nitems_host = nullptr, *nitems_host_dev = nullptr;
  double *near_pairs = nullptr;          // running count of exact near-field pairs, in
                                         // kNearSlots slots (workgroup w adds to slot w mod
                                         // kNearSlots: one address for every workgroup of a
                                         // launch serialises the atomics)
};

__device__ __forceinline__ bool act(uint8_t f) {
  return (f & NB_F_EXISTS) && !(f & NB_F_FRAGMENTING);
}

// ---- centre of mass and largest radius (two-stage, fixed order) ------------------------------
// zero_ctr (may be null): the cycle's pair and record counts are reset here, the first kernel of
// the force phase, instead of by a launch of their own.
__global__ __launch_bounds__(256) void k_sur_com_part(const double *x, const double *y,
                                                      const double *z, const double *m,
                                                      const double *r, const uint8_t *flags,
                                                      int n, double *part, DevCounters *zero_ctr) {
  __shared__ double sh[5][256];
  if (zero_ctr && blockIdx.x == 0 && threadIdx.x == 0) {
    zero_ctr->pair_count = 0;
    zero_ctr->rec_count = 0;
  }
  double s[5] = {0, 0, 0, 0, 0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    if (act(flags[i])) {
      s[0] += m[i];
      s[1] += m[i] * x[i];
      s[2] += m[i] * y[i];
      s[3] += m[i] * z[i];
      s[4] = fmax(s[4], r[i]);
    }
  for (int c = 0; c < 5; c++) sh[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int c = 0; c < 4; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
      sh[4][threadIdx.x] = fmax(sh[4][threadIdx.x], sh[4][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 5; c++) part[blockIdx.x * 5 + c] = sh[c][0];
}

// center = (COM x, y, z, total mass, largest radius of an active body)
__global__ __launch_bounds__(256) void k_sur_com_final(const double *part, int nb,
                                                       double *center) {
  __shared__ double sh[5][256];
  double s[5] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += 256) {
    for (int c = 0; c < 4; c++) s[c] += part[b * 5 + c];
    s[4] = fmax(s[4], part[b * 5 + 4]);
  }
  for (int c = 0; c < 5; c++) sh[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int c = 0; c < 4; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
      sh[4][threadIdx.x] = fmax(sh[4][threadIdx.x], sh[4][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double M = sh[0][0] > 0 ? sh[0][0] : 1.0;
    center[0] = sh[1][0] / M;
    center[1] = sh[2][0] / M;
    center[2] = sh[3][0] / M;
    center[3] = sh[0][0];
    center[4] = sh[4][0];
  }
}

// ---- binning ---------------------------------------------------------------------------------
// Stretched grid: cells uniform in u = d / (|d| + r0) around the centre of mass.  The map is
// 1/r0-Lipschitz (its Jacobian's singular values are 1/(|d|+r0) tangentially and
// r0/(|d|+r0)^2 radially), so two bodies closer than 2 r0 / G fall in the same or adjacent cells
// on every axis: with every radius <= r0 / G, each overlapping pair is a near-field pair.
__device__ __forceinline__ int cell_of(double dx, double dy, double dz, double r0, int G) {
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const double s = 1.0 / (r + r0);
  const double h = 0.5 * G;
  int ix = (int)floor((dx * s + 1.0) * h), iy = (int)floor((dy * s + 1.0) * h),
      iz = (int)floor((dz * s + 1.0) * h);
  ix = min(max(ix, 0), G - 1);
  iy = min(max(iy, 0), G - 1);
  iz = min(max(iz, 0), G - 1);
  return ix + G * (iy + G * iz);
}

__global__ void k_sur_keys(const double *x, const double *y, const double *z,
                           const uint8_t *flags, int n, const double *center, double r0, int G,
                           uint32_t *keys, uint32_t *perm_in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t key = (uint32_t)(G * G * G);
  if (act(flags[i]))
    key = (uint32_t)cell_of(x[i] - center[0], y[i] - center[1], z[i] - center[2], r0, G);
  keys[i] = key;
  perm_in[i] = (uint32_t)i;
}

// cstart[c] = first sorted position whose key is >= c, for c = 0..ncell (inactive bodies carry
// key ncell and sort last, so cstart[ncell] = active bodies).  Binary search, no atomics.  Each
// 1024-cell block also totals its nonempty cells and near-field work items (64-body chunks).
__device__ __forceinline__ int lower_bound_key(const uint32_t *keys, int n, int c) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int)keys[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

__global__ __launch_bounds__(1024) void k_sur_bounds(const uint32_t *keys2, int n, int ncell,
                                                     int32_t *cstart, int2 *btot) {
  __shared__ int32_t wa[16], wb[16];
  const int c = blockIdx.x * 1024 + threadIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int a = 0, b = 0;
  if (c <= ncell) {
    const int lo = lower_bound_key(keys2, n, c);
    cstart[c] = lo;
    if (c < ncell) {
      const int cnt = lower_bound_key(keys2, n, c + 1) - lo;
      a = cnt > 0;
      b = (cnt + 63) >> 6;
    }
  }
  a = wave_incl_scan(a, lane);
  b = wave_incl_scan(b, lane);
  if (lane == 63) {
    wa[w] = a;
    wb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int sa = 0, sb = 0;
    for (int q = 0; q < 16; q++) {
      sa += wa[q];
      sb += wb[q];
    }
    btot[blockIdx.x] = make_int2(sa, sb);
  }
}

// Per-cell counts, the list of nonempty cells and the near-field work items (cell, 64-body
// chunk): 1024 cells per workgroup, block offsets from the block totals, in-block wave prefix
// sums -- integer scans, so the order is fixed.
__global__ __launch_bounds__(1024) void k_sur_scan(const int32_t *cstart, int ncell,
                                                   const int2 *btot, int32_t *ccount,
                                                   int32_t *nelist, int32_t *necount, int2 *items,
                                                   int32_t *nitems, int32_t *nitems_host) {
  __shared__ int32_t wa[16], wb[16], base[2];
  const int c = blockIdx.x * 1024 + threadIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    int sa = 0, sb = 0;
    for (int q = 0; q < (int)blockIdx.x; q++) {
      sa += btot[q].x;
      sb += btot[q].y;
    }
    base[0] = sa;
    base[1] = sb;
    if (blockIdx.x == gridDim.x - 1) {
      *necount = sa + btot[blockIdx.x].x;
      *nitems = sb + btot[blockIdx.x].y;
      if (nitems_host) *nitems_host = sb + btot[blockIdx.x].y;  // the next cycle's grid hint
    }
  }
  const int cnt = c < ncell ? cstart[c + 1] - cstart[c] : 0;
  const int a = cnt > 0, b = (cnt + 63) >> 6;
  const int ia = wave_incl_scan(a, lane), ib = wave_incl_scan(b, lane);
  if (lane == 63) {
    wa[w] = ia;
    wb[w] = ib;
  }
  __syncthreads();
  if (w == 0) {
    const int va = wave_incl_scan(lane < 16 ? wa[lane] : 0, lane);
    const int vb = wave_incl_scan(lane < 16 ? wb[lane] : 0, lane);
    if (lane < 16) {
      wa[lane] = va;
      wb[lane] = vb;
    }
  }
  __syncthreads();
  if (c >= ncell) return;
  const int y = base[0] + (w > 0 ? wa[w - 1] : 0) + ia - a;
  const int q = base[1] + (w > 0 ? wb[w - 1] : 0) + ib - b;
  ccount[c] = cnt;
  if (a) nelist[y] = c;
  for (int j = 0; j < b; j++) items[q + j] = make_int2(c, j);
}

__global__ void k_sur_gather(const uint32_t *perm, const uint32_t *keys2, int n, int ncell,
                             const double *x, const double *y, const double *z, const double *m,
                             const double *center, float4 *sp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || (int)keys2[k] >= ncell) return;
  const uint32_t i = perm[k];
  sp[k] = make_float4((float)(x[i] - center[0]), (float)(y[i] - center[1]),
                      (float)(z[i] - center[2]), (float)m[i]);
}

// Butterfly sum over the 64 lanes; every lane ends with the same bits (each step adds the two
// operands in either order, and FP addition commutes).
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}

// Orders this wave's writes of its private LDS tile before its later reads (and the reads before
// the next overwrite); the waitcnt of a workgroup-scope fence covers the LDS queue.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Cell moments about the cell's centre of mass (FP64 accumulation), one wave per nonempty cell:
// M, COM, traceless quadrupole Q_ab = sum m (3 d_a d_b - |d|^2 delta_ab), radius = max |d|.
// Written by cell index (crec, for the per-body shell) and by nonempty rank (crec_ne, the
// contiguous source stream of the far expansion).
__global__ __launch_bounds__(64) void k_sur_moments(const float4 *sp, const int32_t *cstart,
                                                    const int32_t *nelist,
                                                    const int32_t *necount, int G, float *crec,
                                                    float *crec_ne) {
  const int t = blockIdx.x;
  if (t >= *necount) return;
  const int c = nelist[t];
  const int s = cstart[c], e = cstart[c + 1], lane = threadIdx.x;
  double M = 0, X = 0, Y = 0, Z = 0, gx = 0, gy = 0, gz = 0;
  for (int k = s + lane; k < e; k += 64) {
    const float4 b = sp[k];
    M += b.w;
    X += (double)b.w * b.x;
    Y += (double)b.w * b.y;
    Z += (double)b.w * b.z;
    gx += b.x;
    gy += b.y;
    gz += b.z;
  }
  M = wsum(M);
  if (M > 0) {
    X = wsum(X) / M;
    Y = wsum(Y) / M;
    Z = wsum(Z) / M;
  } else {  // massless bodies only: geometric centre
    X = wsum(gx) / (e - s);
    Y = wsum(gy) / (e - s);
    Z = wsum(gz) / (e - s);
  }
  double q[6] = {0, 0, 0, 0, 0, 0}, rad = 0;
  for (int k = s + lane; k < e; k += 64) {
    const float4 b = sp[k];
    const double dx = b.x - X, dy = b.y - Y, dz = b.z - Z, d2 = dx * dx + dy * dy + dz * dz;
    q[0] += b.w * (3 * dx * dx - d2);
    q[1] += b.w * (3 * dy * dy - d2);
    q[2] += b.w * (3 * dz * dz - d2);
    q[3] += b.w * 3 * dx * dy;
    q[4] += b.w * 3 * dx * dz;
    q[5] += b.w * 3 * dy * dz;
    rad = fmax(rad, sqrt(d2));
  }
  for (int j = 0; j < 6; j++) q[j] = wsum(q[j]);
  rad = wmax(rad);
  if (lane < 2) {
    float *r = (lane == 0 ? crec + (size_t)c * kCellRec : crec_ne + (size_t)t * kCellRec);
    r[0] = (float)X;
    r[1] = (float)Y;
    r[2] = (float)Z;
    r[3] = (float)M;
    for (int j = 0; j < 6; j++) r[4 + j] = (float)q[j];
    r[10] = (float)rad;
    r[11] = (float)(c % G);
    r[12] = (float)((c / G) % G);
    r[13] = (float)(c / (G * G));
    r[14] = (float)(e - s);
    r[15] = 0.f;
  }
}

// ---- far expansion: per nonempty target cell, sum over cells beyond distance K ---------------
// Softened monopole field a = -M r / rs^3 (rs^2 = r^2 + eps^2), its Jacobian
// J_ab = M (3 r_a r_b - rs^2 d_ab) / rs^5 and Hessian
// H_abc = 3 M [(d_ab r_c + d_ac r_b + d_bc r_a) / rs^5 - 5 r_a r_b r_c / rs^7], plus the
// quadrupole field Q r / rs^5 - 5/2 (r.Q.r) r / rs^7; r = X_target - X_source.
//
// Work unit = (64 target cells, one lane each) x (one slice of the source cells); the four waves
// of a workgroup take a quarter of the slice each, staging 64 source records at a time in a
// private LDS tile (one record per lane) and reading them back as wave-uniform broadcasts, so
// every VALU operand is a VGPR; the waves are summed in LDS in wave order.  The slice count is
// chosen from the nonempty-cell count so that there are about kM2LUnits work units; partials per
// slice go to `part` ([slice][nonempty cell][kFarRec]), summed by k_sur_m2l_sum.  `part` then
// never exceeds (kM2LUnits * 64 + ncell) * kFarRec floats.
constexpr int kM2LUnits = 4096;
struct M2LSlices {
  int ntg, nsl, width;  // target groups, slices, source cells per slice (a multiple of 256)
};
__host__ __device__ inline M2LSlices m2l_slices(int nec) {
  M2LSlices r;
  r.ntg = (nec + 63) / 64;
  int nsl = r.ntg > 0 ? (kM2LUnits + r.ntg - 1) / r.ntg : 1;
  const int most = (nec + 255) / 256;
  nsl = nsl < most ? nsl : most;
  nsl = nsl > 1 ? nsl : 1;
  r.width = ((nec + nsl - 1) / nsl + 255) / 256 * 256;
  r.nsl = r.width > 0 ? (nec + r.width - 1) / r.width : 1;
  return r;
}
inline size_t m2l_part_floats(int ncell) { return ((size_t)kM2LUnits * 64 + ncell) * kFarRec; }
__global__ __launch_bounds__(256) void k_sur_m2l(const float *crec_ne, const int32_t *necount,
                                                 int K, float eps2_s, int ncell, float *part) {
  __shared__ float red[3][kFarUsed][64];
  __shared__ float4 recs[4][64][4];
  const int nec = *necount;
  const M2LSlices ms = m2l_slices(nec);
  const int ntg = ms.ntg, nsl = ms.nsl;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const float eps2 = to_vgpr(eps2_s);
  const float Kf = (float)K + 0.5f;
  const float4 *src4 = (const float4 *)crec_ne;
  float4 (*rl)[4] = recs[wave];
  for (int u = blockIdx.x; u < ntg * nsl; u += gridDim.x) {
    const int tg = u % ntg, sl = u / ntg;
    const int t = tg * 64 + lane;
    const bool valid = t < nec;
    const float *tr = crec_ne + (size_t)(valid ? t : nec - 1) * kCellRec;
    const float X = tr[0], Y = tr[1], Z = tr[2], ix = tr[11], iy = tr[12], iz = tr[13];
    // Per source the monopole's derivative tensors are
    //   D1_a = -m3 r_a,  D2_ab = 3 m5 r_a r_b - d_ab m3,
    //   D3_abc = 3 m5 (d_ab r_c + d_ac r_b + d_bc r_a) - 15 m7 r_a r_b r_c,
    //   D4_abcd = 105 m9 r_a r_b r_c r_d - 15 m7 (six d r r) + 3 m5 (three d d),
    // with m_k = M / |r|^k.  They are linear in the sums below, so the loop accumulates those
    // and the Kronecker terms are formed once per target after it.  Index orders: pairs
    // (xx yy zz xy xz yz), triples (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz), quadruples
    // (xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz xyzz xzzz yyyy yyyz yyzz yzzz zzzz).
    float qf[3] = {0.f, 0.f, 0.f};   // quadrupole field
    float a1[3] = {0.f, 0.f, 0.f};   // sum m3 r
    float a0 = 0.f, b0 = 0.f;        // sum m3, sum m5
    float b1[3] = {0.f, 0.f, 0.f};   // sum m5 r
    float b2[6], c2[6], c3[10], e4[15];  // sum m5 rr, m7 rr, m7 rrr, m9 rrrr
#pragma unroll
    for (int q = 0; q < 6; q++) b2[q] = c2[q] = 0.f;
#pragma unroll
    for (int q = 0; q < 10; q++) c3[q] = 0.f;
#pragma unroll
    for (int q = 0; q < 15; q++) e4[q] = 0.f;
    // The target group's rows: nonempty cells are in cell order (x fastest, then y, then z), so
    // z runs from the first lane's to the last valid lane's, and so does y when z is constant.
    const int tlast = min(63, nec - 1 - tg * 64);
    auto lane_f = [](float v, int l) {
      return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
    };
    const float tza = lane_f(iz, 0), tzb = lane_f(iz, tlast);
    const float tya = tza == tzb ? lane_f(iy, 0) : -1e30f;
    const float tyb = tza == tzb ? lane_f(iy, tlast) : 1e30f;
    // One source cell: its monopole's sums and its quadrupole's field.  With TEST the cell is
    // skipped when it lies within Chebyshev distance K of the target (the per-body shell has it);
    // without, the caller has shown that every pair of the chunk lies beyond.
    auto pair = [&](const int j, auto test) {
        const float4 r0 = rl[j][0], r1 = rl[j][1], r2v = rl[j][2], r3 = rl[j][3];
        // r0 = COM, M; r1 = Qxx Qyy Qzz Qxy; r2 = Qxz Qyz radius ix; r3 = iy iz n 0
        const float rx = X - r0.x, ry = Y - r0.y, rz = Z - r0.z;
        float rr = __builtin_fmaf(rx, rx, eps2);
        rr = __builtin_fmaf(ry, ry, rr);
        rr = __builtin_fmaf(rz, rz, rr);
        float inv = __builtin_amdgcn_rsqf(rr);
        if constexpr (decltype(test)::value) {
          const float cheb =
              fmaxf(fabsf(r2v.w - ix), fmaxf(fabsf(r3.x - iy), fabsf(r3.y - iz)));
          inv = cheb > Kf ? inv : 0.f;
        }
        const float inv2 = inv * inv, M = r0.w;
        const float m3 = M * inv2 * inv, m5 = m3 * inv2, m7 = m5 * inv2, m9 = m7 * inv2;
        // quadrupole field
        // (the accumulations are written as fused multiply-adds: the build keeps a * b + c
        // unfused unless asked)
        const float qx = __builtin_fmaf(r2v.x, rz, __builtin_fmaf(r1.w, ry, r1.x * rx));
        const float qy = __builtin_fmaf(r2v.y, rz, __builtin_fmaf(r1.y, ry, r1.w * rx));
        const float qz = __builtin_fmaf(r1.z, rz, __builtin_fmaf(r2v.y, ry, r2v.x * rx));
        const float rqr = __builtin_fmaf(qz, rz, __builtin_fmaf(qy, ry, qx * rx));
        const float i5 = inv2 * inv2 * inv, i7 = i5 * inv2;
        const float c7 = 2.5f * rqr * i7;
        qf[0] = __builtin_fmaf(qx, i5, __builtin_fmaf(-c7, rx, qf[0]));
        qf[1] = __builtin_fmaf(qy, i5, __builtin_fmaf(-c7, ry, qf[1]));
        qf[2] = __builtin_fmaf(qz, i5, __builtin_fmaf(-c7, rz, qf[2]));
        // monopole sums
        a1[0] = __builtin_fmaf(m3, rx, a1[0]);
        a1[1] = __builtin_fmaf(m3, ry, a1[1]);
        a1[2] = __builtin_fmaf(m3, rz, a1[2]);
        a0 += m3;
        b0 += m5;
        const float m5x = m5 * rx, m5y = m5 * ry, m5z = m5 * rz;
        b1[0] += m5x;
        b1[1] += m5y;
        b1[2] += m5z;
        b2[0] = __builtin_fmaf(m5x, rx, b2[0]);
        b2[1] = __builtin_fmaf(m5y, ry, b2[1]);
        b2[2] = __builtin_fmaf(m5z, rz, b2[2]);
        b2[3] = __builtin_fmaf(m5x, ry, b2[3]);
        b2[4] = __builtin_fmaf(m5x, rz, b2[4]);
        b2[5] = __builtin_fmaf(m5y, rz, b2[5]);
        const float xx = rx * rx, yy = ry * ry, zz = rz * rz;
        const float xy = rx * ry, xz = rx * rz, yz = ry * rz;
        const float cxx = m7 * xx, cyy = m7 * yy, czz = m7 * zz;
        const float cxy = m7 * xy, cxz = m7 * xz, cyz = m7 * yz;
        c2[0] += cxx;
        c2[1] += cyy;
        c2[2] += czz;
        c2[3] += cxy;
        c2[4] += cxz;
        c2[5] += cyz;
        c3[0] = __builtin_fmaf(cxx, rx, c3[0]);
        c3[1] = __builtin_fmaf(cxx, ry, c3[1]);
        c3[2] = __builtin_fmaf(cxx, rz, c3[2]);
        c3[3] = __builtin_fmaf(cyy, rx, c3[3]);
        c3[4] = __builtin_fmaf(cxy, rz, c3[4]);
        c3[5] = __builtin_fmaf(czz, rx, c3[5]);
        c3[6] = __builtin_fmaf(cyy, ry, c3[6]);
        c3[7] = __builtin_fmaf(cyy, rz, c3[7]);
        c3[8] = __builtin_fmaf(czz, ry, c3[8]);
        c3[9] = __builtin_fmaf(czz, rz, c3[9]);
        const float exx = m9 * xx, eyy = m9 * yy, ezz = m9 * zz;
        e4[0] = __builtin_fmaf(exx, xx, e4[0]);
        e4[1] = __builtin_fmaf(exx, xy, e4[1]);
        e4[2] = __builtin_fmaf(exx, xz, e4[2]);
        e4[3] = __builtin_fmaf(exx, yy, e4[3]);
        e4[4] = __builtin_fmaf(exx, yz, e4[4]);
        e4[5] = __builtin_fmaf(exx, zz, e4[5]);
        e4[6] = __builtin_fmaf(eyy, xy, e4[6]);
        e4[7] = __builtin_fmaf(eyy, xz, e4[7]);
        e4[8] = __builtin_fmaf(ezz, xy, e4[8]);
        e4[9] = __builtin_fmaf(ezz, xz, e4[9]);
        e4[10] = __builtin_fmaf(eyy, yy, e4[10]);
        e4[11] = __builtin_fmaf(eyy, yz, e4[11]);
        e4[12] = __builtin_fmaf(eyy, zz, e4[12]);
        e4[13] = __builtin_fmaf(ezz, yz, e4[13]);
        e4[14] = __builtin_fmaf(ezz, zz, e4[14]);
    };
    const int per = ms.width / 4;
    const int s0 = sl * ms.width + wave * per, s1 = min(s0 + per, nec);
    for (int sb = s0; sb < s1; sb += 64) {
      const int cnt = min(64, s1 - sb);
      // padding records are all zero: no mass, no quadrupole, no contribution
      for (int q = 0; q < 4; q++)
        rl[lane][q] = lane < cnt ? src4[(size_t)(sb + lane) * 4 + q]
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
      wave_lds_sync();
      // the chunk's rows, as for the targets; a gap above K + 1/2 in z or y puts every pair
      // beyond the shell (Chebyshev distance is at least the gap on any one axis)
      const float4 fa = rl[0][3], fb = rl[cnt - 1][3];
      const float sya = fa.y == fb.y ? fa.x : -1e30f, syb = fa.y == fb.y ? fb.x : 1e30f;
      const float gap = fmaxf(fmaxf(fa.y - tzb, tza - fb.y), fmaxf(sya - tyb, tya - syb));
      if (__builtin_amdgcn_readfirstlane(gap > Kf)) {
#pragma unroll 2
        for (int j = 0; j < cnt; j++) pair(j, std::false_type{});
      } else {
#pragma unroll 2
        for (int j = 0; j < cnt; j++) pair(j, std::true_type{});
      }
      wave_lds_sync();
    }
    // the tensors from the sums: field, Jacobian, Hessian, third derivatives
    float o[kFarUsed];
    o[0] = qf[0] - a1[0];
    o[1] = qf[1] - a1[1];
    o[2] = qf[2] - a1[2];
    o[3] = 3.f * b2[0] - a0;
    o[4] = 3.f * b2[1] - a0;
    o[5] = 3.f * b2[2] - a0;
    o[6] = 3.f * b2[3];
    o[7] = 3.f * b2[4];
    o[8] = 3.f * b2[5];
    o[9] = 9.f * b1[0] - 15.f * c3[0];
    o[10] = 3.f * b1[1] - 15.f * c3[1];
    o[11] = 3.f * b1[2] - 15.f * c3[2];
    o[12] = 3.f * b1[0] - 15.f * c3[3];
    o[13] = -15.f * c3[4];
    o[14] = 3.f * b1[0] - 15.f * c3[5];
    o[15] = 9.f * b1[1] - 15.f * c3[6];
    o[16] = 3.f * b1[2] - 15.f * c3[7];
    o[17] = 3.f * b1[1] - 15.f * c3[8];
    o[18] = 9.f * b1[2] - 15.f * c3[9];
    o[19] = 105.f * e4[0] - 90.f * c2[0] + 9.f * b0;
    o[20] = 105.f * e4[1] - 45.f * c2[3];
    o[21] = 105.f * e4[2] - 45.f * c2[4];
    o[22] = 105.f * e4[3] - 15.f * (c2[0] + c2[1]) + 3.f * b0;
    o[23] = 105.f * e4[4] - 15.f * c2[5];
    o[24] = 105.f * e4[5] - 15.f * (c2[0] + c2[2]) + 3.f * b0;
    o[25] = 105.f * e4[6] - 45.f * c2[3];
    o[26] = 105.f * e4[7] - 15.f * c2[4];
    o[27] = 105.f * e4[8] - 15.f * c2[3];
    o[28] = 105.f * e4[9] - 45.f * c2[4];
    o[29] = 105.f * e4[10] - 90.f * c2[1] + 9.f * b0;
    o[30] = 105.f * e4[11] - 45.f * c2[5];
    o[31] = 105.f * e4[12] - 15.f * (c2[1] + c2[2]) + 3.f * b0;
    o[32] = 105.f * e4[13] - 45.f * c2[5];
    o[33] = 105.f * e4[14] - 90.f * c2[2] + 9.f * b0;
    if (wave > 0)
      for (int q = 0; q < kFarUsed; q++) red[wave - 1][q][lane] = o[q];
    __syncthreads();
    if (wave == 0 && valid) {
      float *dst = part + ((size_t)sl * nec + t) * kFarRec;
      for (int q = 0; q < kFarUsed; q++)
        dst[q] = ((o[q] + red[0][q][lane]) + red[1][q][lane]) + red[2][q][lane];
      for (int q = kFarUsed; q < kFarRec; q++) dst[q] = 0.f;
    }
    __syncthreads();
  }
}

// fexp[cell] = sum of the slice partials in slice order (one thread per record word).
__global__ void k_sur_m2l_sum(const float *part, const int32_t *nelist, const int32_t *necount,
                              int ncell, float *fexp) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int nec = *necount;
  if (g >= nec * kFarRec) return;
  const int t = g / kFarRec, q = g % kFarRec;
  const int nsl = m2l_slices(nec).nsl;
  float s = 0.f;
  for (int sl = 0; sl < nsl; sl++) s += part[(size_t)sl * nec * kFarRec + g];
  fexp[(size_t)nelist[t] * kFarRec + q] = s;
}

// ---- near field, multipole shell, Taylor step and features per body ------------------------
__device__ __forceinline__ void quad_field(f4v r0, f4v r1, f4v r2v, float px, float py,
                                           float pz, float eps2, float &ax, float &ay,
                                           float &az) {
  const float rx = px - r0.x, ry = py - r0.y, rz = pz - r0.z;
  float r2 = __builtin_fmaf(rx, rx, eps2);
  r2 = __builtin_fmaf(ry, ry, r2);
  r2 = __builtin_fmaf(rz, rz, r2);
  const float inv = __builtin_amdgcn_rsqf(r2);
  const float inv2 = inv * inv, m3 = r0.w * inv2 * inv;
  const float qx = __builtin_fmaf(r2v.x, rz, __builtin_fmaf(r1.w, ry, r1.x * rx));
  const float qy = __builtin_fmaf(r2v.y, rz, __builtin_fmaf(r1.y, ry, r1.w * rx));
  const float qz = __builtin_fmaf(r1.z, rz, __builtin_fmaf(r2v.y, ry, r2v.x * rx));
  const float rqr = __builtin_fmaf(qz, rz, __builtin_fmaf(qy, ry, qx * rx));
  const float i5 = inv2 * inv2 * inv, c7 = 2.5f * rqr * i5 * inv2;
  const float k = -(m3 + c7);  // monopole and quadrupole radial terms together
  ax = __builtin_fmaf(qx, i5, __builtin_fmaf(k, rx, ax));
  ay = __builtin_fmaf(qy, i5, __builtin_fmaf(k, ry, ay));
  az = __builtin_fmaf(qz, i5, __builtin_fmaf(k, rz, az));
}

__device__ __forceinline__ void near_pair(f4v b, float4 me, float eps2, float &px, float &py,
                                          float &pz) {
  const float dx = b.x - me.x, dy = b.y - me.y, dz = b.z - me.z;
  float r2 = __builtin_fmaf(dx, dx, eps2);
  r2 = __builtin_fmaf(dy, dy, r2);
  r2 = __builtin_fmaf(dz, dz, r2);
  const float inv = __builtin_amdgcn_rsqf(r2);
  const float s = b.w * inv * inv * inv;
  px = __builtin_fmaf(dx, s, px);
  py = __builtin_fmaf(dy, s, py);
  pz = __builtin_fmaf(dz, s, pz);
}

__device__ __forceinline__ void quad_field(float4 r0, float4 r1, float4 r2v, float px, float py,
                                           float pz, float eps2, float &ax, float &ay,
                                           float &az) {
  quad_field(f4v{r0.x, r0.y, r0.z, r0.w}, f4v{r1.x, r1.y, r1.z, r1.w},
             f4v{r2v.x, r2v.y, r2v.z, r2v.w}, px, py, pz, eps2, ax, ay, az);
}

__device__ __forceinline__ void near_pair(float4 b, float4 me, float eps2, float &px, float &py,
                                          float &pz) {
  near_pair(f4v{b.x, b.y, b.z, b.w}, me, eps2, px, py, pz);
}

// near_pair that also keeps the smallest r^2 = d^2 + eps^2 seen (collision candidates).
__device__ __forceinline__ void near_pair_min(float4 b, float4 me, float eps2, float &px,
                                              float &py, float &pz, float &r2min) {
  const float dx = b.x - me.x, dy = b.y - me.y, dz = b.z - me.z;
  float r2 = __builtin_fmaf(dx, dx, eps2);
  r2 = __builtin_fmaf(dy, dy, r2);
  r2 = __builtin_fmaf(dz, dz, r2);
  r2min = fminf(r2min, r2);
  const float inv = __builtin_amdgcn_rsqf(r2);
  const float s = b.w * inv * inv * inv;
  px = __builtin_fmaf(dx, s, px);
  py = __builtin_fmaf(dy, s, py);
  pz = __builtin_fmaf(dz, s, pz);
}

struct NearArgs {
  const float4 *sp;
  const int32_t *cstart, *ccount;
  const int2 *items;
  const int32_t *nitems;
  const float *crec, *fexp;
  const float *mu, *sd;
  float *feat;
  float4 *base;
  int G, K, write_feat;
  float eps2, r0;
  double *near_pairs;
  // collision detection (near field only; the grid keeps every overlapping pair adjacent)
  const uint32_t *perm;                // sorted position -> body index
  const double *x, *y, *z, *r;         // FP64 state
  const int32_t *beh;
  uint64_t *pairs;
  DevCounters *ctr;
  int64_t pair_cap;
  int events;
};

__device__ __forceinline__ float round_up_f(double d) {  // smallest float >= d (d >= 0)
  float f = (float)d;
  if ((double)f < d) f = __uint_as_float(__float_as_uint(f) + 1u);
  return f;
}

__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
  return v;
}

__device__ __forceinline__ void sur_emit_pair(const NearArgs &a, uint32_t i, uint32_t j) {
  const uint32_t q = atomicAdd(&a.ctr->pair_count, 1u);
  if ((int64_t)q < a.pair_cap)
    a.pairs[q] = ((uint64_t)i << 32) | j;
  else
    atomicOr(&a.ctr->pair_overflow, 1u);
}

// The reference's overlap test, Collided (body.go:192-208), in FP64 with its operation order:
// sqrt(dx*dx + dy*dy + dz*dz) <= Ri + Rj, for the sources of one chunk against target i; each
// unordered pair is reported once, by its lower body index.
__device__ void exact_chunk_pairs(const NearArgs &a, uint32_t i, int s0, int cnt) {
  const double xi = a.x[i], yi = a.y[i], zi = a.z[i], ri = a.r[i];
  const int32_t bi = a.beh[i];
  if (!nb_may_pair(bi, a.events)) return;
  for (int jj = 0; jj < cnt; jj++) {
    const uint32_t j = a.perm[s0 + jj];
    if (j <= i) continue;
    const double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
    const double d = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
    const double rj = a.r[j];
    if (!(d > ri + rj) && nb_pair_event(bi, a.beh[j], ri, rj, d, a.events))
      sur_emit_pair(a, i, j);
  }
}

// One workgroup (4 waves) per work item = up to 64 bodies of one cell, one lane each.  The waves
// split the cell's neighbourhood: four equal contiguous shares of the nine rows of adjacent cells,
// walked in chunks of 64 sources (128 with collisions) staged in the wave's LDS tile, and the
// shell cells (round robin, staged by ballot); partial sums meet in LDS in wave order, and wave 0
// adds the Taylor step of the far expansion and writes the base acceleration and the MLP features.
// COLL: each lane also keeps the minimum r^2 of every chunk and compares it against a bound that
// no overlapping pair can exceed in FP32 (the chunk's largest radius and coordinate, rounded up,
// plus the FP32 rounding of positions); a flagged (body, chunk) is re-tested exactly in FP64.
// LDS: one 3 KB region per wave, used in turn for the wave's source tile (64 float4, 128 with
// collisions), its staged shell records (64 x 3 float4) and, in waves 1-3, its partial sums for
// wave 0 (3 x 64 doubles): 12 KB per workgroup.  Without collisions the kernel is held to 64
// VGPRs, so 8 waves fit per SIMD; with collisions it is held to 80 VGPRs (6 waves per SIMD, no
// spills; 94 VGPRs and 5 waves unconstrained).
template <bool COLL>
__device__ __forceinline__ void near_item(const NearArgs &a, const int item,
                                          float4 (*wlds)[64 * 3]) {
  const int G = a.G;
  const int2 it = a.items[item];
  const int c = it.x;
  const int cx = c % G, cy = (c / G) % G, cz = c / (G * G);
  cip cs = (cip)a.cstart;
  const int c_begin = cs[c], c_end = cs[c + 1];
  const int t_begin = c_begin + it.y * 64, t_end = min(c_end, t_begin + 64);
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  // A wave holds `copies` copies of the item's targets (the most, up to 8, that fit in 64
  // lanes); copy cp takes a contiguous share of each tile and every copies-th shell cell, and
  // the copies' partial sums are combined in a fixed order at the end.
  const int tcount = t_end - t_begin;
  const int copies = tcount > 32 ? 1 : tcount > 16 ? 2 : tcount > 8 ? 4 : 8;
  const int cwid = 64 / copies;  // lanes per copy
  const int cp = lane / cwid;
  const int k = t_begin + lane % cwid;
  const bool tv = k < t_end;
  const float4 me = a.sp[tv ? k : t_begin];
  const float eps2 = to_vgpr(a.eps2);
  float4 *tl = wlds[wave];
  float4 (*rl)[3] = reinterpret_cast<float4 (*)[3]>(wlds[wave]);
  double ax = 0, ay = 0, az = 0;
  int nsrc = 0;
  uint32_t i_body = 0;
  float ri = 0.f, me_inf = 0.f;
  if (COLL && tv) {
    i_body = a.perm[k];
    ri = round_up_f(a.r[i_body]);
    me_inf = fmaxf(fabsf(me.x), fmaxf(fabsf(me.y), fabsf(me.z)));
  }
  // The sources are the bodies of the 3 x 3 rows of three adjacent cells around the cell, each
  // row contiguous in the sorted order.  Taken in row order they are dealt to the four waves as
  // four equal contiguous shares (so a wave's work does not depend on how the rows split into
  // chunks), and each share is walked in chunks of up to 64 that stay within one row.
  const int xl = max(cx - 1, 0), xh = min(cx + 1, G - 1);
  auto row_range = [&](int q, int &b, int &l) {
    const int y = cy - 1 + q % 3, z = cz - 1 + q / 3;
    b = l = 0;
    if (y >= 0 && y < G && z >= 0 && z < G) {
      const int row = G * (y + G * z);
      b = cs[xl + row];
      l = cs[xh + row + 1] - b;
    }
  };
  for (int q = 0; q < 9; q++) {
    int b, l;
    row_range(q, b, l);
    nsrc += l;
  }
  // multipole shell positions (below): the first round's cell counts are requested here, so that
  // their latency overlaps the first source chunk's
  const int K = a.K, D = 2 * K + 1, box = D * D * D;
  const float invD = 1.f / (float)D;
  auto shell_pos = [&](int p, int &cell) -> bool {
    if (p >= box) return false;
    const int pq = (int)(((float)p + 0.5f) * invD);     // p / D
    const int pr = (int)(((float)pq + 0.5f) * invD);    // p / D^2
    const int dx = p - pq * D - K, dy = pq - pr * D - K, dz = pr - K;
    const int x = cx + dx, y = cy + dy, z = cz + dz;
    if (max(abs(dx), max(abs(dy), abs(dz))) < 2 || x < 0 || x >= G || y < 0 || y >= G || z < 0 ||
        z >= G)
      return false;
    cell = x + G * (y + G * z);
    return true;
  };
  int cell0 = 0;
  const bool in0 = shell_pos(4 * lane + wave, cell0);
  const int cnt0 = in0 ? a.ccount[cell0] : 0;
  // sources per chunk: a chunk's fixed costs (staging, the collision bound's reductions, the
  // FP64 accumulation) are spread over twice the pairs with collisions on
  constexpr int CH = COLL ? 128 : 64;
  const int share = (nsrc + 3) / 4;
  const int w0 = wave * share, w1 = min(nsrc, w0 + share);
  int base = 0;
  for (int q = 0; q < 9; q++) {
    int rb, rlen;
    row_range(q, rb, rlen);
    const int lo = max(w0, base), hi = min(w1, base + rlen);
    for (int s = lo; s < hi; s += CH) {
      {
        const int s0 = rb + (s - base);
        const int cnt = min(CH, hi - s);
        // the wave's private tile: CH / 64 sources per lane, zero-mass padding to a multiple of 4
        float rmax_c = 0.f, xmax_c = 0.f;
#pragma unroll
        for (int h = 0; h < CH; h += 64) {
          const int e = h + lane;
          const float4 src = e < cnt ? a.sp[s0 + e] : make_float4(0.f, 0.f, 0.f, 0.f);
          tl[e] = src;
          if (COLL && e < cnt) {
            rmax_c = fmaxf(rmax_c, round_up_f(a.r[a.perm[s0 + e]]));
            xmax_c = fmaxf(xmax_c, fmaxf(fabsf(src.x), fmaxf(fabsf(src.y), fabsf(src.z))));
          }
        }
        if (COLL) {
          rmax_c = wave_maxf(rmax_c);
          xmax_c = wave_maxf(xmax_c);
        }
        wave_lds_sync();
        float px = 0, py = 0, pz = 0, r2min = FLT_MAX;
        // the tile padded with zero-mass entries to a multiple of 4 copies (at most CH, the
        // tile's size); copy cp takes the contiguous share [cp len, (cp + 1) len), four sources
        // per step, with the same trip count on every lane
        const int q4 = 4 * copies;
        const int cntp = (cnt + q4 - 1) / q4 * q4;
        const int len = cntp / copies;
        const float4 *tc = tl + cp * len;
        if (COLL) {
          // the body's own entry (r^2 = eps^2) must not count as a candidate
          if (s0 < t_end && s0 + cnt > t_begin) {
            const int self = k - s0;
            const int j1 = min(cp * len + len, cnt);
            for (int j = cp * len; j < j1; j++) {
              const float keep = r2min;
              near_pair_min(tl[j], me, eps2, px, py, pz, r2min);
              if (j == self) r2min = keep;
            }
          } else {
            // zero-mass padding may only add candidates (an exact pass over the real sources)
            for (int j = 0; j < len; j += 4) {
              const float4 b0 = tc[j], b1 = tc[j + 1], b2 = tc[j + 2], b3 = tc[j + 3];
              near_pair_min(b0, me, eps2, px, py, pz, r2min);
              near_pair_min(b1, me, eps2, px, py, pz, r2min);
              near_pair_min(b2, me, eps2, px, py, pz, r2min);
              near_pair_min(b3, me, eps2, px, py, pz, r2min);
            }
          }
          // the chunk's minimum over all copies, for the target's one decision below
          for (int o = cwid; o < 64; o <<= 1) r2min = fminf(r2min, __shfl_xor(r2min, o));
        } else {
          for (int j = 0; j < len; j += 4) {
            const float4 b0 = tc[j], b1 = tc[j + 1], b2 = tc[j + 2], b3 = tc[j + 3];
            near_pair(b0, me, eps2, px, py, pz);
            near_pair(b1, me, eps2, px, py, pz);
            near_pair(b2, me, eps2, px, py, pz);
            near_pair(b3, me, eps2, px, py, pz);
          }
        }
        wave_lds_sync();
        if (COLL && tv && cp == 0) {
          // |d_fp32 - d| <= e for every pair of this chunk; 2^-20 covers the rounding of both
          // positions to FP32 and of their differences with room to spare
          const float e = 9.5367431640625e-7f * (me_inf + xmax_c + ri + rmax_c);
          const float thr = ri + rmax_c + e;
          if (r2min <= (thr * thr + eps2) * 1.000002f) {
            atomicAdd((unsigned long long *)&a.ctr->fixup_tiles, 1ull);
            exact_chunk_pairs(a, i_body, s0, cnt);
          }
        }
        ax += px;
        ay += py;
        az += pz;
      }
    }
    base += rlen;
  }
  // multipole shell, Chebyshev distance 2..K: wave w takes the box positions w, w + 4, ... of
  // the (2K+1)^3 box, one per lane and 64 per round; the nonempty ones are staged in LDS (one per
  // lane) by ballot and read back as broadcasts.  Positions are decoded with a float reciprocal:
  // for p < (2K+1)^3 <= 2197 the product (p + 0.5) / D lies at least 0.5 / D from an integer,
  // far beyond its rounding, so the truncation is the exact quotient.
  {
    float sx = 0, sy = 0, sz = 0;
    for (int p0 = 0; p0 < box; p0 += 256) {
      int cell = cell0;
      bool ok = cnt0 > 0;
      if (p0 > 0) {
        ok = false;
        if (shell_pos(p0 + 4 * lane + wave, cell)) ok = a.ccount[cell] > 0;
      }
      const uint64_t okm = __ballot(ok);
      const int slot = __builtin_amdgcn_mbcnt_hi((uint32_t)(okm >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)okm, 0u));
      if (ok) {
        const float4 *r = (const float4 *)(a.crec + (size_t)cell * kCellRec);
        rl[slot][0] = r[0];
        rl[slot][1] = r[1];
        rl[slot][2] = r[2];
      }
      wave_lds_sync();
      const int nsel = __builtin_popcountll(okm);
      for (int j = cp; j < nsel; j += copies)
        quad_field(rl[j][0], rl[j][1], rl[j][2], me.x, me.y, me.z, eps2, sx, sy, sz);
      wave_lds_sync();
    }
    ax += sx;
    ay += sy;
    az += sz;
  }
  if (wave > 0) {
    // the wave's own region: its shell loop ended with a wave barrier after its last read
    double (*rw)[64] = reinterpret_cast<double (*)[64]>(wlds[wave]);
    rw[0][lane] = ax;
    rw[1][lane] = ay;
    rw[2][lane] = az;
  }
  __syncthreads();
  if (wave != 0) return;
  if (lane == 0)
    atomicAdd(a.near_pairs + (blockIdx.x & (kNearSlots - 1)),
              (double)(t_end - t_begin) * (double)nsrc);
  {
    double (*r1)[64] = reinterpret_cast<double (*)[64]>(wlds[1]);
    double (*r2)[64] = reinterpret_cast<double (*)[64]>(wlds[2]);
    double (*r3)[64] = reinterpret_cast<double (*)[64]>(wlds[3]);
    ax = ((ax + r1[0][lane]) + r2[0][lane]) + r3[0][lane];
    ay = ((ay + r1[1][lane]) + r2[1][lane]) + r3[1][lane];
    az = ((az + r1[2][lane]) + r2[2][lane]) + r3[2][lane];
  }
  for (int o = cwid; o < 64; o <<= 1) {  // the copies, pairwise
    ax += __shfl_xor(ax, o);
    ay += __shfl_xor(ay, o);
    az += __shfl_xor(az, o);
  }
  if (!tv || cp != 0) return;
  // third-order Taylor step of the far expansion from the cell's centre of mass
  cfp cr = (cfp)(a.crec + (size_t)c * kCellRec);
  cfp fe = (cfp)(a.fexp + (size_t)c * kFarRec);
  const float dx = me.x - cr[0], dy = me.y - cr[1], dz = me.z - cr[2];
  cfp J = fe + 3;
  cfp H = fe + 9;
  float tx = fe[0] + J[0] * dx + J[3] * dy + J[4] * dz;
  float ty = fe[1] + J[3] * dx + J[1] * dy + J[5] * dz;
  float tz = fe[2] + J[4] * dx + J[5] * dy + J[2] * dz;
  // 1/2 H_abc d_b d_c with H index (xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz, zzz)
  tx += 0.5f * (H[0] * dx * dx + H[3] * dy * dy + H[5] * dz * dz +
                2.f * (H[1] * dx * dy + H[2] * dx * dz + H[4] * dy * dz));
  ty += 0.5f * (H[1] * dx * dx + H[6] * dy * dy + H[8] * dz * dz +
                2.f * (H[3] * dx * dy + H[4] * dx * dz + H[7] * dy * dz));
  tz += 0.5f * (H[2] * dx * dx + H[7] * dy * dy + H[9] * dz * dz +
                2.f * (H[4] * dx * dy + H[5] * dx * dz + H[8] * dy * dz));
  // 1/6 T_abcd d_b d_c d_d with T index (xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz xyzz xzzz yyyy
  // yyyz yyzz yzzz zzzz); for component a the sum over (b, c, d) counts each index multiset by
  // its number of orderings
  cfp T = fe + 19;
  {
    const float x2 = dx * dx, y2 = dy * dy, z2 = dz * dz;
    const float x3 = x2 * dx, y3 = y2 * dy, z3 = z2 * dz;
    const float x2y = x2 * dy, x2z = x2 * dz, xy2 = dx * y2, xz2 = dx * z2, y2z = y2 * dz,
                yz2 = dy * z2, xyz = dx * dy * dz;
    const float c6 = 1.f / 6.f;
    // component x: T_x... = (xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz xyzz xzzz)
    tx += c6 * (T[0] * x3 + 3.f * (T[1] * x2y + T[2] * x2z + T[3] * xy2 + T[5] * xz2 +
                                   T[7] * y2z + T[8] * yz2) +
                6.f * T[4] * xyz + T[6] * y3 + T[9] * z3);
    // component y: T_y... = (xxxy xxyy xxyz xyyy xyyz xyzz yyyy yyyz yyzz yzzz)
    ty += c6 * (T[1] * x3 + 3.f * (T[3] * x2y + T[4] * x2z + T[6] * xy2 + T[8] * xz2 +
                                   T[11] * y2z + T[12] * yz2) +
                6.f * T[7] * xyz + T[10] * y3 + T[13] * z3);
    // component z: T_z... = (xxxz xxyz xxzz xyyz xyzz xzzz yyyz yyzz yzzz zzzz)
    tz += c6 * (T[2] * x3 + 3.f * (T[4] * x2y + T[5] * x2z + T[7] * xy2 + T[9] * xz2 +
                                   T[12] * y2z + T[13] * yz2) +
                6.f * T[8] * xyz + T[11] * y3 + T[14] * z3);
  }
  ax += tx;
  ay += ty;
  az += tz;
  // scale-free features
  const float sc = fmaxf(cr[10], 1e-6f * a.r0);
  const float na = sqrtf(fe[0] * fe[0] + fe[1] * fe[1] + fe[2] * fe[2]);
  float nj = 0, nh = 0;
  for (int q = 0; q < 6; q++) nj += J[q] * J[q] * (q < 3 ? 1.f : 2.f);
  const float hw[10] = {1, 3, 3, 3, 6, 3, 1, 3, 3, 1};
  for (int q = 0; q < 10; q++) nh += hw[q] * H[q] * H[q];
  float nt = 0;
  const float tw[15] = {1, 4, 4, 6, 12, 6, 4, 12, 12, 4, 1, 4, 6, 4, 1};
  for (int q = 0; q < 15; q++) nt += tw[q] * T[q] * T[q];
  const float A = na + sqrtf(nj) * sc + sqrtf(nh) * sc * sc + sqrtf(nt) * sc * sc * sc + 1e-30f;
  a.base[k] = make_float4((float)ax, (float)ay, (float)az, A);
  if (a.write_feat) {
    float f[kFeat];
    const float ux = dx / sc, uy = dy / sc, uz = dz / sc;
    f[0] = ux; f[1] = uy; f[2] = uz;
    f[3] = ux * ux; f[4] = uy * uy; f[5] = uz * uz; f[6] = ux * uy; f[7] = ux * uz;
    f[8] = uy * uz;
    const float iA = 1.f / A;
    f[9] = fe[0] * iA; f[10] = fe[1] * iA; f[11] = fe[2] * iA;
    for (int q = 0; q < 6; q++) f[12 + q] = J[q] * sc * iA;
    for (int q = 0; q < 10; q++) f[18 + q] = H[q] * sc * sc * iA;
    f[28] = logf(sc / a.r0);
    f[29] = logf((float)(c_end - c_begin));
    float *o = a.feat + (size_t)k * kFeat;
    for (int q = 0; q < kFeat; q++) o[q] = (f[q] - a.mu[q]) / a.sd[q];
  }
}

// One workgroup per item (the launch may be smaller than the item count: workgroups then take
// further items; the barrier between two items keeps wave 0's reads of the partial sums ahead of
// the other waves' next writes).
template <bool COLL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(COLL ? 6 : 8)))
void k_sur_near(NearArgs a) {
  __shared__ float4 wlds[4][64 * 3];
  const int nitems = *a.nitems;
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    if (item != (int)blockIdx.x) __syncthreads();
    near_item<COLL>(a, item, wlds);
  }
}

// Scatter the surrogate's accelerations (base + learned residual) to the accumulator slice 0.
__global__ void k_sur_apply(const uint32_t *perm, const uint32_t *keys2, int n, int ncell,
                            const float4 *base, const float *Y, int ldy, float out_scale,
                            int use_mlp, double *acc, int npad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || (int)keys2[k] >= ncell) return;
  const uint32_t i = perm[k];
  const float4 b = base[k];
  double ax = b.x, ay = b.y, az = b.z;
  if (use_mlp) {
    const double s = (double)b.w * out_scale;
    ax += s * Y[(size_t)k * ldy + 0];
    ay += s * Y[(size_t)k * ldy + 1];
    az += s * Y[(size_t)k * ldy + 2];
  }
  acc[i] = ax;
  acc[npad + i] = ay;
  acc[2 * (size_t)npad + i] = az;
}

// ---- sampled exact check -------------------------------------------------------------------
// FP32 direct sum for the sampled (sorted) bodies: one wave per (64 samples, 256-source slice),
// a sample per lane, sources streamed through the scalar cache; FP64 slice partials.
constexpr int kExactSlice = 256;

__global__ __launch_bounds__(64) void k_sur_exact(const float4 *sp, const int32_t *cstart,
                                                  int ncell, const int32_t *sidx, int S,
                                                  float eps2_s, double *part) {
  __shared__ float4 tl[64];
  cip cs = (cip)cstart;
  const int nact = cs[ncell];
  const int sl = blockIdx.y, j0 = sl * kExactSlice;
  if (j0 >= nact) return;  // uniform
  const int j1 = min(j0 + kExactSlice, nact);
  const int s = blockIdx.x * 64 + threadIdx.x;
  const bool sv = s < S;
  const int k = sv ? min(sidx[s], nact - 1) : 0;
  const float4 me = sp[k];
  const float eps2 = to_vgpr(eps2_s);
  const int lane = threadIdx.x;
  float px = 0, py = 0, pz = 0;
  for (int jb = j0; jb < j1; jb += 64) {
    const int cnt = min(64, j1 - jb);
    tl[lane] = lane < cnt ? sp[jb + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
    wave_lds_sync();
    const int cnt4 = (cnt + 3) & ~3;
    for (int j = 0; j < cnt4; j += 4) {
      const float4 b0 = tl[j], b1 = tl[j + 1], b2 = tl[j + 2], b3 = tl[j + 3];
      near_pair(b0, me, eps2, px, py, pz);
      near_pair(b1, me, eps2, px, py, pz);
      near_pair(b2, me, eps2, px, py, pz);
      near_pair(b3, me, eps2, px, py, pz);
    }
    wave_lds_sync();
  }
  if (sv) {
    double *o = part + ((size_t)sl * S + s) * 3;
    o[0] = px;
    o[1] = py;
    o[2] = pz;
  }
}

// exact[s][q] = sum of the slice partials: one wave per (s, q), lanes over slices (lane l takes
// slices l, l+64, ... in order), then the butterfly -- a fixed order.
__global__ __launch_bounds__(256) void k_sur_exact_sum(const double *part, int S,
                                                       const int32_t *cstart, int ncell,
                                                       double *exact) {
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (g >= S * 3) return;  // uniform per wave
  const int nsl = (cstart[ncell] + kExactSlice - 1) / kExactSlice;
  double v = 0;
  for (int sl = lane; sl < nsl; sl += 64) v += part[(size_t)sl * S * 3 + g];
  v = wsum(v);
  if (lane == 0) exact[g] = v;
}

// Relative error of the surrogate over the samples: the global ||da|| / ||a|| (fixed-order tree
// reduction) and each sample's own |da_s| / |a_s|, of which the k-th largest is found by a radix
// select on the float bits (exact; k = floor((1 - quantile) S) + 1, so at most a fraction
// 1 - quantile of the samples lies above it).  The fallback flag is raised when the global error
// exceeds tol or that per-body value exceeds tol_body (tol_body > 0).  err[0] = global, err[1] =
// the per-body value.
__global__ __launch_bounds__(256) void k_sur_check(const int32_t *sidx, int S,
                                                   const uint32_t *perm, const int32_t *cstart,
                                                   int ncell, const double *exact,
                                                   const double *acc, int npad, double tol,
                                                   double tol_body, double quantile,
                                                   uint32_t *body, uint32_t *flag, double *err) {
  __shared__ double sn[256], sd[256];
  __shared__ int hist[256], suf[257];
  __shared__ uint32_t s_prefix, s_digit;
  __shared__ int s_k;
  const int nact = cstart[ncell];
  double num = 0, den = 0;
  for (int s = threadIdx.x; s < S; s += 256) {
    const int k = sidx[s];
    double bn = 0, bd = 0;
    if (k < nact) {
      const uint32_t i = perm[k];
      for (int q = 0; q < 3; q++) {
        const double e = exact[(size_t)s * 3 + q];
        const double d = acc[(size_t)q * npad + i] - e;
        bn += d * d;
        bd += e * e;
      }
    }
    num += bn;
    den += bd;
    const float r = bd > 0 ? (float)sqrt(bn / bd) : (bn > 0 ? INFINITY : 0.f);
    body[s] = r != r ? 0x7fc00000u : __float_as_uint(r);  // NaN ranks above every number
  }
  sn[threadIdx.x] = num;
  sd[threadIdx.x] = den;
  if (threadIdx.x == 0) {
    s_prefix = 0;
    const double above = floor((1.0 - quantile) * (double)S);
    s_k = (int)(above < 0 ? 0 : (above > S - 1 ? S - 1 : above)) + 1;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sn[threadIdx.x] += sn[threadIdx.x + w];
      sd[threadIdx.x] += sd[threadIdx.x + w];
    }
    __syncthreads();
  }
  uint32_t mask = 0;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t prefix = s_prefix;
    for (int s = threadIdx.x; s < S; s += 256) {
      const uint32_t u = body[s];
      if ((u & mask) == prefix) atomicAdd(&hist[(u >> shift) & 255u], 1);
    }
    // the digit of the k-th largest among the values with this prefix: the largest b whose
    // suffix count suf[b] = sum over b' >= b of hist[b'] reaches k (b = 0 when none above does)
    suf[threadIdx.x] = hist[threadIdx.x];
    if (threadIdx.x == 0) {
      suf[256] = 0;
      s_digit = 0;
    }
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int add = threadIdx.x + off < 256 ? suf[threadIdx.x + off] : 0;
      __syncthreads();
      suf[threadIdx.x] += add;
      __syncthreads();
    }
    const int k = s_k;
    const int b = (int)threadIdx.x;
    if (b > 0 && suf[b] >= k && suf[b + 1] < k) s_digit = (uint32_t)b;
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t d = s_digit;
      s_k = k - suf[d + 1];
      s_prefix = prefix | (d << shift);
    }
    mask |= 255u << shift;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double r = sd[0] > 0 ? sqrt(sn[0] / sd[0]) : 0.0;
    const double rb = (double)__uint_as_float(s_prefix);
    err[0] = r;
    err[1] = rb;
    *flag = (r > tol || (tol_body > 0 && !(rb <= tol_body))) ? 1u : 0u;
  }
}

// After a fallback force launch: fold the plain kernel's slices into slice 0 (flag set only).
// After a fallback (flag up): the direct sum's slices folded into slice 0, and the fallback
// counted once.
__global__ void k_sur_fold(double *acc, int split, int npad, int n, const uint32_t *flag,
                           DevCounters *ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!*flag) return;
  if (i == 0) atomicAdd((unsigned long long *)&ctr->sur_fallbacks, 1ull);
  if (i >= n) return;
  for (int c = 0; c < 3; c++) {
    double s = 0;
    for (int q = 0; q < split; q++) s += acc[((size_t)q * 3 + c) * npad + i];
    acc[(size_t)c * npad + i] = s;
  }
}

__global__ void k_sur_setflag(uint32_t *flag, uint32_t v) {
  if (threadIdx.x == 0) *flag = v;
}

// Training rows: features of sampled bodies and their normalised residual
// (a_exact - base) / A, with a_exact folded from the plain kernel's slices.
__global__ void k_sur_label(const int32_t *sidx, int S, const uint32_t *perm,
                            const uint32_t *keys2, int ncell, const float4 *base,
                            const float *feat, const double *acc, int split, int npad,
                            float *tx, float *ty) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int k = sidx[s];
  const uint32_t i = perm[k];
  const float4 b = base[k];
  double ex[3] = {0, 0, 0};
  for (int q = 0; q < split; q++)
    for (int c = 0; c < 3; c++) ex[c] += acc[((size_t)q * 3 + c) * npad + i];
  ty[(size_t)s * 3 + 0] = (float)((ex[0] - b.x) / b.w);
  ty[(size_t)s * 3 + 1] = (float)((ex[1] - b.y) / b.w);
  ty[(size_t)s * 3 + 2] = (float)((ex[2] - b.z) / b.w);
  for (int q = 0; q < kFeat; q++) tx[(size_t)s * kFeat + q] = feat[(size_t)k * kFeat + q];
  (void)keys2;
  (void)ncell;
}

__global__ void k_sur_gather_rows(const float *src, const int32_t *rows, int nrows, int width,
                                  float *dst) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)nrows * width) return;
  const int r = (int)(q / width), c = (int)(q % width);
  dst[q] = src[(size_t)rows[r] * width + c];
}

inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace nb

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
namespace nb {

// Sample s of a check: rank floor((s + u_s) nact / S) of the cell-sorted order, with u_s in [0, 1)
// from a counter-based hash of (seed, draw, s): one body from each of S equal strata of the
// sorted order (every region of space in proportion to its bodies), a different one at each
// draw.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ void k_sur_samples(int32_t *sidx, int S, const int32_t *cstart, int ncell,
                              uint64_t seed, uint64_t draw) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int nact = cstart[ncell];
  const uint64_t h = mix64(seed ^ mix64(draw * 0x2545f4914f6cdd1dull + (uint64_t)s));
  const double u = (double)(h >> 11) * 0x1.0p-53;
  int64_t k = (int64_t)(((double)s + u) * (double)nact / (double)S);
  if (k >= nact) k = nact - 1;
  sidx[s] = nact > 0 ? (int)k : 0;
}

template <class T>
static hipError_t salloc(T **p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  return hipMalloc((void **)p, (count ? count : 1) * sizeof(T));
}

void sur_free(Sur *s) {
  if (!s) return;
  if (s->nitems_host) (void)hipHostFree(s->nitems_host);
  void *ptrs[] = {s->keys, s->keys2, s->perm_in, s->perm, s->sort_tmp, s->cstart, s->ccount,
                  s->nelist, s->necount, s->sp, s->crec, s->crec_ne, s->fexp, s->part,
                  s->items, s->nitems, s->btot, s->feat, s->base, s->center,
                  s->red, s->mu, s->sd, s->tx, s->ty, s->sidx, s->sexact, s->serr, s->flag,
                  s->sbody,
                  s->spart, s->ymlp,
                  s->mu0, s->sd1, s->near_pairs};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (s->mlp) nb_mlp_destroy(s->mlp);
  delete s;
}

static int sur_ensure_bodies(Ctx *c) {
  Sur *s = c->sur;
  const int64_t need_items = (int64_t)s->ncell + c->cap / 64 + 2;
  if (s->items_cap < need_items) {
    NB_HIP(c, salloc(&s->items, (size_t)need_items));
    s->items_cap = need_items;
  }
  if (s->cap >= c->cap) return NB_OK;
  const size_t n = (size_t)c->cap;
  NB_HIP(c, salloc(&s->keys, n));
  NB_HIP(c, salloc(&s->keys2, n));
  NB_HIP(c, salloc(&s->perm_in, n));
  NB_HIP(c, salloc(&s->perm, n));
  NB_HIP(c, salloc(&s->sp, n));
  NB_HIP(c, salloc(&s->base, n));
  NB_HIP(c, salloc(&s->ymlp, n * 3));
  NB_HIP(c, salloc(&s->feat, n * kFeat));
  size_t bytes = 0;
  NB_HIP(c, rocprim::radix_sort_pairs<SortCfg>(
                nullptr, bytes, s->keys, s->keys2, s->perm_in, s->perm, (size_t)n, 0u,
                key_bits(kMaxGrid * kMaxGrid * kMaxGrid), c->stream));
  void *tmp = nullptr;
  NB_HIP(c, hipMalloc(&tmp, bytes ? bytes : 1));
  if (s->sort_tmp) (void)hipFree(s->sort_tmp);
  s->sort_tmp = tmp;
  s->sort_bytes = bytes;
  s->cap = c->cap;
  return NB_OK;
}

static int sur_ensure_samples(Ctx *c, int S) {
  Sur *s = c->sur;
  if (S > s->S) {
    NB_HIP(c, salloc(&s->sidx, (size_t)S));
    NB_HIP(c, salloc(&s->sexact, (size_t)S * 3));
    NB_HIP(c, salloc(&s->sbody, (size_t)S));
    s->S = S;
  }
  const int64_t need = (c->cap + kExactSlice - 1) / kExactSlice * (int64_t)s->S * 3;
  if (need > s->spart_len) {
    NB_HIP(c, salloc(&s->spart, (size_t)need));
    s->spart_len = need;
  }
  return NB_OK;
}

// Sampled error of the accelerations in acc slice 0 against the FP32 direct sum: S bodies, one
// per stratum of the sorted order, drawn afresh at every check (rotate) or the same fixed draw
// (a measurement that callers compare across settings); the global and per-body errors go to
// serr and, above tol or tol_body, the fallback flag is set.
static int sur_check_launch(Ctx *c, int S, double tol, double tol_body, bool rotate) {
  Sur *s = c->sur;
  hipStream_t st = c->stream;
  int rc = sur_ensure_samples(c, S);
  if (rc) return rc;
  const uint64_t draw = rotate ? (uint64_t)s->draws++ : ~0ull;
  hipLaunchKernelGGL(k_sur_samples, dim3(nb256(S)), dim3(256), 0, st, s->sidx, S, s->cstart,
                     s->ncell, (uint64_t)s->p.seed, draw);
  const unsigned nsl = (unsigned)((c->n + kExactSlice - 1) / kExactSlice);
  hipLaunchKernelGGL(k_sur_exact, dim3((unsigned)((S + 63) / 64), nsl), dim3(64), 0, st, s->sp,
                     s->cstart, s->ncell, s->sidx, S, (float)(c->p.eps * c->p.eps), s->spart);
  hipLaunchKernelGGL(k_sur_exact_sum, dim3((unsigned)((S * 3 + 3) / 4)), dim3(256), 0, st,
                     s->spart, S, s->cstart, s->ncell, s->sexact);
  hipLaunchKernelGGL(k_sur_check, dim3(1), dim3(256), 0, st, s->sidx, S, s->perm, s->cstart,
                     s->ncell, s->sexact, c->acc, (int)c->npad, tol, tol_body,
                     s->p.body_quantile, s->sbody, s->flag, s->serr);
  NB_HIP(c, hipGetLastError());
  return NB_OK;
}

// Binning, moments, far expansion, near field (+ features).  apply: scatter into acc slice 0.
// detect: collision pairs from the near field into the context's pair list.
static int sur_pipeline(Ctx *c, bool feat_raw, bool feat_norm, bool apply, bool use_mlp,
                        bool detect = false, bool zero_counters = false) {
  Sur *s = c->sur;
  hipStream_t st = c->stream;
  const int n = (int)c->n, G = s->G, ncell = s->ncell;
  if (n == 0) return NB_OK;
  int rc = sur_ensure_bodies(c);
  if (rc) return rc;
  const int nbk = (int)std::min<int64_t>(1024, (n + 255) / 256);
  hipLaunchKernelGGL(k_sur_com_part, dim3(nbk), dim3(256), 0, st, c->x, c->y, c->z, c->m, c->r,
                     c->flags, n, s->red, zero_counters ? c->ctr : nullptr);
  hipLaunchKernelGGL(k_sur_com_final, dim3(1), dim3(256), 0, st, s->red, nbk, s->center);
  hipLaunchKernelGGL(k_sur_keys, dim3(nb256(n)), dim3(256), 0, st, c->x, c->y, c->z, c->flags,
                     n, s->center, s->r0, G, s->keys, s->perm_in);
  NB_HIP(c, hipGetLastError());
  size_t bytes = s->sort_bytes;
  NB_HIP(c, rocprim::radix_sort_pairs<SortCfg>(s->sort_tmp, bytes, s->keys, s->keys2, s->perm_in,
                                               s->perm, (size_t)n, 0u, key_bits(ncell), st));
  hipLaunchKernelGGL(k_sur_bounds, dim3((unsigned)(ncell / 1024 + 1)), dim3(1024), 0, st,
                     s->keys2, n, ncell, s->cstart, s->btot);
  hipLaunchKernelGGL(k_sur_scan, dim3((unsigned)((ncell + 1023) / 1024)), dim3(1024), 0, st,
                     s->cstart, ncell, s->btot, s->ccount, s->nelist, s->necount, s->items,
                     s->nitems, s->nitems_host_dev);
  hipLaunchKernelGGL(k_sur_gather, dim3(nb256(n)), dim3(256), 0, st, s->perm, s->keys2, n, ncell,
                     c->x, c->y, c->z, c->m, s->center, s->sp);
  hipLaunchKernelGGL(k_sur_moments, dim3((unsigned)ncell), dim3(64), 0, st, s->sp, s->cstart,
                     s->nelist, s->necount, G, s->crec, s->crec_ne);
  const float eps2 = (float)(c->p.eps * c->p.eps);
  hipLaunchKernelGGL(k_sur_m2l, dim3(2048), dim3(256), 0, st, s->crec_ne, s->necount, s->K, eps2,
                     ncell, s->part);
  hipLaunchKernelGGL(k_sur_m2l_sum, dim3(nb256((int64_t)ncell * kFarRec)), dim3(256), 0, st,
                     s->part, s->nelist,
                     s->necount, ncell, s->fexp);
  NB_HIP(c, hipGetLastError());
  NearArgs na;
  na.sp = s->sp;
  na.cstart = s->cstart;
  na.ccount = s->ccount;
  na.items = s->items;
  na.nitems = s->nitems;
  na.crec = s->crec;
  na.fexp = s->fexp;
  na.mu = feat_norm ? s->mu : s->mu0;
  na.sd = feat_norm ? s->sd : s->sd1;
  na.feat = s->feat;
  na.base = s->base;
  na.G = G;
  na.K = s->K;
  na.write_feat = (feat_raw || feat_norm) ? 1 : 0;
  na.eps2 = eps2;
  na.r0 = (float)s->r0;
  na.near_pairs = s->near_pairs;
  na.perm = s->perm;
  na.x = c->x;
  na.y = c->y;
  na.z = c->z;
  na.r = c->r;
  na.beh = c->beh;
  na.pairs = c->pairs;
  na.ctr = c->ctr;
  na.pair_cap = c->pair_cap;
  na.events = c->p.events;
  // one workgroup per item: the bound on items, or a quarter more than the last count seen
  // (workgroups loop over items, so a count that has grown since only costs a second trip)
  const int64_t item_ub = (int64_t)ncell + (n + 63) / 64 + 1;
  const int32_t seen = *(volatile int32_t *)s->nitems_host;
  const unsigned grid =
      (unsigned)(seen > 0 ? std::min<int64_t>(item_ub, (int64_t)seen + seen / 4 + 256) : item_ub);
  if (detect)
    hipLaunchKernelGGL(k_sur_near<true>, dim3(grid), dim3(256), 0, st, na);
  else
    hipLaunchKernelGGL(k_sur_near<false>, dim3(grid), dim3(256), 0, st, na);
  NB_HIP(c, hipGetLastError());
  if (!apply) return NB_OK;
  const bool mlp_on = use_mlp && s->trained && s->mlp;
  if (mlp_on) {
    rc = mlp_infer_device(s->mlp, s->feat, n, s->ymlp, st);
    if (rc) return fail_msg(c, rc, "surrogate MLP inference failed");
  }
  hipLaunchKernelGGL(k_sur_apply, dim3(nb256(n)), dim3(256), 0, st, s->perm, s->keys2, n, ncell,
                     s->base, s->ymlp, 3, s->out_scale, mlp_on ? 1 : 0, c->acc, (int)c->npad);
  NB_HIP(c, hipGetLastError());
  return NB_OK;
}

// The force phase of one cycle when the surrogate is on.
int sur_force_phase(Ctx *c, bool timed, int ev_slot) {
  Sur *s = c->sur;
  hipStream_t st = c->stream;
  if (c->p.collisions && !(c->rmax <= s->r0 / s->G * (1.0 - 1e-6)))  // a later nb_load
    return fail_msg(c, NB_E_UNSUPPORTED,
                    "collisions with the surrogate need every radius <= r0 / grid");
  // the pipeline's first kernel resets the cycle's counters (it launches nothing without bodies)
  if (c->n == 0) NB_HIP(c, launch_zero_counters(c->ctr, 0, st));
  if (timed) NB_HIP(c, hipEventRecord(c->ev[2 * ev_slot], st));
  const bool detect = c->p.collisions != 0;
  int rc = sur_pipeline(c, false, s->p.use_mlp && s->trained, true, s->p.use_mlp != 0, detect,
                        true);
  if (rc) return rc;
  if (timed) NB_HIP(c, hipEventRecord(c->ev[2 * ev_slot + 1], st));
  if (s->p.check_every > 0 && s->steps % s->p.check_every == 0 && c->n > 0) {
    rc = sur_check_launch(c, s->p.check_samples, s->p.tol, s->p.tol_body, true);
    if (rc) return rc;
    // direct-sum redo of this cycle's force when the flag is up (the kernel exits otherwise)
    ForceArgs fa = make_force_args(c);
    fa.run_flag = s->flag;
    fa.detect = 0;  // the near field already listed every colliding pair
    fa.drop_flagged = 0;
    NB_HIP(c, launch_force_f32(c->shape, fa, st));
    hipLaunchKernelGGL(k_sur_fold, dim3(nb256(c->n)), dim3(256), 0, st, c->acc, c->shape.jsplit,
                       (int)c->npad, (int)c->n, s->flag, c->ctr);
    NB_HIP(c, hipGetLastError());
    s->checks++;
  }
  if (detect) {
    const int rcs = sort_pairs(c);
    if (rcs) return rcs;
  }
  s->steps++;
  return NB_OK;
}

}  // namespace nb

using namespace nb;

extern "C" {

int nb_sur_params_init(nb_sur_params *p) {
  if (!p) return NB_E_INVALID;
  memset(p, 0, sizeof *p);
  p->grid = 32;
  p->shell = 2;
  p->r0 = 0.0;
  p->use_mlp = 1;
  p->check_every = 1;
  p->check_samples = 256;
  p->tol = 1e-3;
  p->seed = 12345;
  p->tol_body = 0.0;
  p->body_quantile = 0.99;
  return NB_OK;
}

int nb_sur_enable(nb_ctx *c, const nb_sur_params *p) {
  if (!c || !p) return NB_E_INVALID;
  if (!c->loaded) return fail_msg(c, NB_E_STATE, "nb_sur_enable: load bodies first");
  if (c->p.precision != NB_PREC_FP32 || !(c->p.eps > 0.0))
    return fail_msg(c, NB_E_UNSUPPORTED, "the surrogate needs the FP32 path and eps > 0");
  if (c->p.overlap_cutoff)
    return fail_msg(c, NB_E_UNSUPPORTED, "the surrogate path has no overlap cutoff");
  if (p->grid < 4 || p->grid > kMaxGrid || p->shell < 1 || p->shell > 6 || p->check_every < 0 ||
      (p->check_every > 0 && (p->check_samples < 1 || p->check_samples > 65536)) ||
      !(p->tol > 0) || p->r0 < 0 || !(p->tol_body >= 0) ||
      !(p->body_quantile > 0 && p->body_quantile <= 1))
    return fail_msg(c, NB_E_INVALID, "nb_sur_enable: bad parameters");
  NB_HIP(c, hipSetDevice(c->device));
  Sur *s = c->sur;
  const bool fresh = !s;
  if (fresh) {
    s = new (std::nothrow) Sur();
    if (!s) return fail_msg(c, NB_E_NOMEM, "out of host memory");
    c->sur = s;
    NB_HIP(c, hipHostMalloc((void **)&s->nitems_host, sizeof(int32_t), hipHostMallocMapped));
    *s->nitems_host = -1;
    NB_HIP(c, hipHostGetDevicePointer((void **)&s->nitems_host_dev, s->nitems_host, 0));
  }
  const bool grid_changed = fresh || s->G != p->grid;
  s->p = *p;
  s->G = p->grid;
  s->ncell = p->grid * p->grid * p->grid;
  s->K = p->shell;
  if (grid_changed) {
    const size_t nc = (size_t)s->ncell;
    NB_HIP(c, salloc(&s->cstart, nc + 1));
    NB_HIP(c, salloc(&s->ccount, nc));
    NB_HIP(c, salloc(&s->nelist, nc));
    NB_HIP(c, salloc(&s->necount, 1));
    NB_HIP(c, salloc(&s->crec, nc * kCellRec));
    NB_HIP(c, salloc(&s->crec_ne, nc * kCellRec));
    NB_HIP(c, salloc(&s->btot, nc / 1024 + 2));
    NB_HIP(c, salloc(&s->fexp, nc * kFarRec));
    NB_HIP(c, salloc(&s->part, m2l_part_floats((int)nc)));
    s->items_cap = 0;  // re-sized for the new grid by sur_ensure_bodies
  }
  if (fresh) {
    NB_HIP(c, salloc(&s->center, 8));
    NB_HIP(c, salloc(&s->red, 1024 * 5));
    NB_HIP(c, salloc(&s->mu, kFeat));
    NB_HIP(c, salloc(&s->sd, kFeat));
    NB_HIP(c, salloc(&s->mu0, kFeat));
    NB_HIP(c, salloc(&s->sd1, kFeat));
    NB_HIP(c, salloc(&s->serr, 2));
    NB_HIP(c, salloc(&s->flag, 1));
    NB_HIP(c, salloc(&s->near_pairs, kNearSlots));
    NB_HIP(c, salloc(&s->nitems, 1));
    std::vector<float> z(kFeat, 0.f), o(kFeat, 1.f);
    NB_HIP(c, hipMemcpy(s->mu0, z.data(), kFeat * 4, hipMemcpyHostToDevice));
    NB_HIP(c, hipMemcpy(s->sd1, o.data(), kFeat * 4, hipMemcpyHostToDevice));
    NB_HIP(c, hipMemcpy(s->mu, z.data(), kFeat * 4, hipMemcpyHostToDevice));
    NB_HIP(c, hipMemcpy(s->sd, o.data(), kFeat * 4, hipMemcpyHostToDevice));
    NB_HIP(c, hipMemset(s->serr, 0, 2 * sizeof(double)));
    NB_HIP(c, hipMemset(s->flag, 0, sizeof(uint32_t)));
    NB_HIP(c, hipMemset(s->near_pairs, 0, kNearSlots * sizeof(double)));
  }
  // stretch radius: 1.5 x the half-mass radius about the centre of mass
  if (p->r0 > 0) {
    s->r0 = p->r0;
  } else {
    const size_t n = (size_t)c->n;
    std::vector<double> hx(n), hy(n), hz(n), hm(n);
    std::vector<uint8_t> hf(n);
    NB_HIP(c, hipMemcpy(hx.data(), c->x, n * 8, hipMemcpyDeviceToHost));
    NB_HIP(c, hipMemcpy(hy.data(), c->y, n * 8, hipMemcpyDeviceToHost));
    NB_HIP(c, hipMemcpy(hz.data(), c->z, n * 8, hipMemcpyDeviceToHost));
    NB_HIP(c, hipMemcpy(hm.data(), c->m, n * 8, hipMemcpyDeviceToHost));
    NB_HIP(c, hipMemcpy(hf.data(), c->flags, n, hipMemcpyDeviceToHost));
    double M = 0, X = 0, Y = 0, Z = 0;
    for (size_t i = 0; i < n; i++)
      if ((hf[i] & NB_F_EXISTS) && !(hf[i] & NB_F_FRAGMENTING)) {
        M += hm[i];
        X += hm[i] * hx[i];
        Y += hm[i] * hy[i];
        Z += hm[i] * hz[i];
      }
    if (M > 0) {
      X /= M;
      Y /= M;
      Z /= M;
    }
    std::vector<std::pair<double, double>> rm;
    for (size_t i = 0; i < n; i++)
      if ((hf[i] & NB_F_EXISTS) && !(hf[i] & NB_F_FRAGMENTING))
        rm.push_back({std::sqrt((hx[i] - X) * (hx[i] - X) + (hy[i] - Y) * (hy[i] - Y) +
                                (hz[i] - Z) * (hz[i] - Z)),
                      hm[i]});
    std::sort(rm.begin(), rm.end());
    double acc = 0, rh = 1.0;
    for (auto &q : rm) {
      acc += q.second;
      if (acc >= 0.5 * M) {
        rh = q.first;
        break;
      }
    }
    s->r0 = 1.5 * std::max(rh, 1e-30);
  }
  if (c->p.collisions && !(c->rmax <= s->r0 / s->G * (1.0 - 1e-6))) {
    c->sur_on = false;
    return fail_msg(c, NB_E_UNSUPPORTED,
                    "collisions with the surrogate need every radius <= r0 / grid (overlapping "
                    "pairs must be near-field pairs): use a larger r0 or a coarser grid");
  }
  c->sur_on = true;
  return sur_ensure_bodies(c);
}

int nb_sur_disable(nb_ctx *c) {
  if (!c) return NB_E_INVALID;
  c->sur_on = false;
  return NB_OK;
}

// Append training rows (features and normalised far-field residual) for `samples` bodies of the
// loaded state, labelled by the plain FP32 direct sum.
int nb_sur_collect(nb_ctx *c, int64_t samples, uint64_t seed) {
  if (!c) return NB_E_INVALID;
  if (!c->sur_on) return fail_msg(c, NB_E_STATE, "nb_sur_collect: surrogate not enabled");
  if (samples < 1 || samples > (1 << 24))
    return fail_msg(c, NB_E_INVALID, "nb_sur_collect: samples out of range");
  Sur *s = c->sur;
  NB_HIP(c, hipSetDevice(c->device));
  int rc = sur_pipeline(c, true, false, false, false);
  if (rc) return rc;
  ForceArgs fa = make_force_args(c);
  fa.run_flag = nullptr;
  fa.detect = 0;
  fa.drop_flagged = 0;
  NB_HIP(c, launch_force_f32(c->shape, fa, c->stream));
  int32_t nact = 0;
  NB_HIP(c, hipMemcpyAsync(&nact, s->cstart + s->ncell, 4, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  if (nact == 0) return NB_OK;
  const int64_t S = std::min<int64_t>(samples, nact);
  std::vector<int32_t> pick((size_t)nact);
  for (int32_t q = 0; q < nact; q++) pick[q] = q;
  std::mt19937_64 rng(seed);
  for (int64_t q = 0; q < S; q++) {
    std::uniform_int_distribution<int64_t> U(q, nact - 1);
    std::swap(pick[(size_t)q], pick[(size_t)U(rng)]);
  }
  pick.resize((size_t)S);
  std::sort(pick.begin(), pick.end());
  if (s->trows + S > s->tcap) {
    const int64_t ncap = std::max<int64_t>(2 * s->tcap, s->trows + S);
    float *nx = nullptr, *ny = nullptr;
    NB_HIP(c, hipMalloc((void **)&nx, (size_t)ncap * kFeat * 4));
    NB_HIP(c, hipMalloc((void **)&ny, (size_t)ncap * 3 * 4));
    if (s->trows) {
      NB_HIP(c, hipMemcpy(nx, s->tx, (size_t)s->trows * kFeat * 4, hipMemcpyDeviceToDevice));
      NB_HIP(c, hipMemcpy(ny, s->ty, (size_t)s->trows * 3 * 4, hipMemcpyDeviceToDevice));
    }
    if (s->tx) (void)hipFree(s->tx);
    if (s->ty) (void)hipFree(s->ty);
    s->tx = nx;
    s->ty = ny;
    s->tcap = ncap;
  }
  int32_t *dpick = nullptr;
  NB_HIP(c, hipMalloc((void **)&dpick, (size_t)S * 4));
  hipError_t e = hipMemcpyAsync(dpick, pick.data(), (size_t)S * 4, hipMemcpyHostToDevice,
                                c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sur_label, dim3(nb256(S)), dim3(256), 0, c->stream, dpick, (int)S,
                       s->perm, s->keys2, s->ncell, s->base, s->feat, c->acc, c->shape.jsplit,
                       (int)c->npad, s->tx + (size_t)s->trows * kFeat, s->ty + (size_t)s->trows * 3);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dpick);
  NB_HIP(c, e);
  s->trows += S;
  return NB_OK;
}

int nb_sur_fit(nb_ctx *c, int32_t steps, int32_t batch, float lr, float *loss) {
  if (!c) return NB_E_INVALID;
  if (!c->sur_on || !c->sur) return fail_msg(c, NB_E_STATE, "nb_sur_fit: surrogate not enabled");
  Sur *s = c->sur;
  if (s->trows < batch || batch < 64 || batch % 64 || steps < 1)
    return fail_msg(c, NB_E_INVALID, "nb_sur_fit: need at least one batch of collected rows");
  NB_HIP(c, hipSetDevice(c->device));
  const size_t R = (size_t)s->trows;
  std::vector<float> X(R * kFeat), Y(R * 3);
  NB_HIP(c, hipMemcpy(X.data(), s->tx, X.size() * 4, hipMemcpyDeviceToHost));
  NB_HIP(c, hipMemcpy(Y.data(), s->ty, Y.size() * 4, hipMemcpyDeviceToHost));
  std::vector<double> mu(kFeat, 0.0), var(kFeat, 0.0);
  for (size_t r = 0; r < R; r++)
    for (int q = 0; q < kFeat; q++) mu[q] += X[r * kFeat + q];
  for (int q = 0; q < kFeat; q++) mu[q] /= (double)R;
  for (size_t r = 0; r < R; r++)
    for (int q = 0; q < kFeat; q++) {
      const double d = X[r * kFeat + q] - mu[q];
      var[q] += d * d;
    }
  std::vector<float> fmu(kFeat), fsd(kFeat);
  for (int q = 0; q < kFeat; q++) {
    fmu[q] = (float)mu[q];
    fsd[q] = (float)std::max(std::sqrt(var[q] / (double)R), 1e-12);
  }
  double ys = 0;
  for (float v : Y) ys += (double)v * v;
  const float oscale = (float)std::max(std::sqrt(ys / (double)Y.size()), 1e-30);
  // shuffle rows once (seeded), normalise inputs and targets
  std::vector<size_t> ord(R);
  for (size_t r = 0; r < R; r++) ord[r] = r;
  std::mt19937_64 rng(s->p.seed);
  std::shuffle(ord.begin(), ord.end(), rng);
  std::vector<float> Xn(R * kFeat), Yn(R * 3);
  for (size_t r = 0; r < R; r++) {
    const size_t o = ord[r];
    for (int q = 0; q < kFeat; q++) Xn[r * kFeat + q] = (X[o * kFeat + q] - fmu[q]) / fsd[q];
    for (int q = 0; q < 3; q++) Yn[r * 3 + q] = Y[o * 3 + q] / oscale;
  }
  if (!s->mlp) {
    int rc = nb_mlp_create(c->device, kFeat, 128, 3, 3, s->p.seed, &s->mlp);
    if (rc) return fail_msg(c, rc, "nb_sur_fit: could not create the MLP");
  }
  const int64_t rows = (int64_t)(R / batch) * batch;
  float l = 0;
  int rc = nb_mlp_train(s->mlp, Xn.data(), Yn.data(), rows, steps, batch, lr, &l);
  if (rc) return fail_msg(c, rc, nb_mlp_last_error(s->mlp));
  NB_HIP(c, hipMemcpy(s->mu, fmu.data(), kFeat * 4, hipMemcpyHostToDevice));
  NB_HIP(c, hipMemcpy(s->sd, fsd.data(), kFeat * 4, hipMemcpyHostToDevice));
  s->out_scale = oscale;
  s->trained = true;
  s->train_loss = l;
  if (loss) *loss = l;
  return NB_OK;
}

// Sampled relative force error of the surrogate on the current state (no update): `samples`
// bodies spread over the sorted order against the FP32 direct sum.  use_mlp: 0 off, 1 on.
int nb_sur_error(nb_ctx *c, int32_t samples, int32_t use_mlp, double *err) {
  if (!c || !err) return NB_E_INVALID;
  if (!c->sur_on) return fail_msg(c, NB_E_STATE, "nb_sur_error: surrogate not enabled");
  if (samples < 1 || samples > 65536) return fail_msg(c, NB_E_INVALID, "bad sample count");
  Sur *s = c->sur;
  NB_HIP(c, hipSetDevice(c->device));
  NB_HIP(c, launch_zero_counters(c->ctr, 0, c->stream));
  const bool mlp = use_mlp && s->trained;
  int rc = sur_pipeline(c, false, mlp, true, mlp);
  if (rc) return rc;
  rc = sur_check_launch(c, samples, 1e300, 0.0, false);
  if (rc) return rc;
  NB_HIP(c, hipMemcpyAsync(err, s->serr, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  return NB_OK;
}

int nb_sur_get_stats(nb_ctx *c, nb_sur_stats *out) {
  if (!c || !out) return NB_E_INVALID;
  memset(out, 0, sizeof *out);
  if (!c->sur) return NB_OK;
  Sur *s = c->sur;
  NB_HIP(c, hipSetDevice(c->device));
  DevCounters ctr;
  double err[2] = {0, 0}, slots[kNearSlots];
  int32_t nec = 0;
  NB_HIP(c, hipMemcpyAsync(&ctr, c->ctr, sizeof ctr, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipMemcpyAsync(err, s->serr, sizeof err, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipMemcpyAsync(slots, s->near_pairs, sizeof slots, hipMemcpyDeviceToHost,
                           c->stream));
  NB_HIP(c, hipMemcpyAsync(&nec, s->necount, 4, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  out->steps = s->steps;
  out->checks = s->checks;
  out->fallbacks = (int64_t)ctr.sur_fallbacks;
  out->last_error = err[0];
  out->last_body_error = err[1];
  double pairs = 0;  // integer-valued terms below 2^53: the sum is exact in any order
  for (int q = 0; q < kNearSlots; q++) pairs += slots[q];
  out->near_pairs_total = pairs;
  out->nonempty_cells = nec;
  out->train_rows = s->trows;
  out->train_loss = s->train_loss;
  out->r0 = s->r0;
  out->trained = s->trained ? 1 : 0;
  return NB_OK;
}

int nb_sur_save(nb_ctx *c, const char *path) {
  if (!c || !path) return NB_E_INVALID;
  Sur *s = c->sur;
  if (!s || !s->trained) return fail_msg(c, NB_E_STATE, "nb_sur_save: no trained model");
  int64_t cnt = 0;
  nb_mlp_param_count(s->mlp, &cnt);
  std::vector<float> prm((size_t)cnt), mu(kFeat), sd(kFeat);
  int rc = nb_mlp_get_params(s->mlp, prm.data(), cnt);
  if (rc) return fail_msg(c, rc, "nb_sur_save: reading parameters failed");
  NB_HIP(c, hipMemcpy(mu.data(), s->mu, kFeat * 4, hipMemcpyDeviceToHost));
  NB_HIP(c, hipMemcpy(sd.data(), s->sd, kFeat * 4, hipMemcpyDeviceToHost));
  FILE *f = fopen(path, "wb");
  if (!f) return fail_msg(c, NB_E_INVALID, "nb_sur_save: cannot open file");
  const char magic[8] = {'N', 'B', 'S', 'U', 'R', '0', '0', '1'};
  const int32_t hdr[4] = {kFeat, 128, 3, 3};
  bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(hdr, 4, 4, f) == 4 &&
            fwrite(&cnt, 8, 1, f) == 1 && fwrite(mu.data(), 4, kFeat, f) == (size_t)kFeat &&
            fwrite(sd.data(), 4, kFeat, f) == (size_t)kFeat &&
            fwrite(&s->out_scale, 4, 1, f) == 1 &&
            fwrite(prm.data(), 4, (size_t)cnt, f) == (size_t)cnt;
  ok = (fclose(f) == 0) && ok;
  return ok ? NB_OK : fail_msg(c, NB_E_INVALID, "nb_sur_save: write failed");
}

int nb_sur_load(nb_ctx *c, const char *path) {
  if (!c || !path) return NB_E_INVALID;
  Sur *s = c->sur;
  if (!s) return fail_msg(c, NB_E_STATE, "nb_sur_load: enable the surrogate first");
  FILE *f = fopen(path, "rb");
  if (!f) return fail_msg(c, NB_E_INVALID, "nb_sur_load: cannot open file");
  char magic[8];
  int32_t hdr[4];
  int64_t cnt = 0;
  std::vector<float> mu(kFeat), sd(kFeat);
  float osc = 1.f;
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "NBSUR001", 8) == 0 &&
            fread(hdr, 4, 4, f) == 4 && hdr[0] == kFeat && hdr[1] == 128 && hdr[2] == 3 &&
            hdr[3] == 3 && fread(&cnt, 8, 1, f) == 1 && cnt > 0 && cnt < (1 << 24) &&
            fread(mu.data(), 4, kFeat, f) == (size_t)kFeat &&
            fread(sd.data(), 4, kFeat, f) == (size_t)kFeat && fread(&osc, 4, 1, f) == 1;
  std::vector<float> prm;
  if (ok) {
    prm.resize((size_t)cnt);
    ok = fread(prm.data(), 4, (size_t)cnt, f) == (size_t)cnt;
  }
  fclose(f);
  if (!ok) return fail_msg(c, NB_E_INVALID, "nb_sur_load: not a surrogate model file");
  if (!s->mlp) {
    int rc = nb_mlp_create(c->device, kFeat, 128, 3, 3, s->p.seed, &s->mlp);
    if (rc) return fail_msg(c, rc, "nb_sur_load: could not create the MLP");
  }
  int rc = nb_mlp_set_params(s->mlp, prm.data(), cnt);
  if (rc) return fail_msg(c, rc, "nb_sur_load: parameter count mismatch");
  NB_HIP(c, hipMemcpy(s->mu, mu.data(), kFeat * 4, hipMemcpyHostToDevice));
  NB_HIP(c, hipMemcpy(s->sd, sd.data(), kFeat * 4, hipMemcpyHostToDevice));
  s->out_scale = osc;
  s->trained = true;
  return NB_OK;
}

}  // extern "C"
