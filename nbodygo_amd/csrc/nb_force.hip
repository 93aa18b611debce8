// nb_force.hip -- FP32 pairwise gravity for gfx950 and its FP64 exact fix-up pass.
//
// Restates the O(N^2) force loop of Body.Compute / calcForceFrom (cmd/body/body.go:156-170,
// :214-225) as an acceleration sum a_i = sum_j m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^(3/2)
// (G is applied in FP64 by the update), and folds in the candidate test of the collision
// detection loop (body.go:171-186, Collided :192-208).
//
// Layout: sources are one float4 (x, y, z, m) per body in HBM, padded to a multiple of 1024 with
// far-away massless sentinels.  Each lane owns IPL targets; a workgroup covers block*IPL targets
// and one slice of the source range (grid.y = slices).  Sources are consumed in tiles of tile_j:
// within a tile each target keeps plain FP32 partial sums; at the tile end the partial is added
// to a running (sum, compensation) pair with Kahan summation.  Per interaction the fast path
// issues 12 VALU ops plus one v_rsq_f32:
//   3 sub, 3 fma (r^2 = eps^2 + dx^2 + dy^2 + dz^2), rsq, 3 mul (m * r^-3), 3 fma (accumulate).
//
// Detection (collisions on, overlap cutoff, or eps == 0): each target also keeps the tile's
// minimum r^2 (one v_min_f32).  At the tile end it is compared with a conservative bound of
// (R_i + Rmax_tile)^2 + eps^2 that covers every FP32 rounding of the positions and of r^2.  A tile
// that passes the bound is "flagged": its (slice, target, tile) is recorded and, when the overlap
// cutoff or eps == 0 needs exact per-pair decisions, its FP32 partial is dropped.  The fix-up
// kernel then redoes exactly those tiles in FP64 with the reference's distance test
// sqrt(dx*dx + dy*dy + dz*dz) <= Ri + Rj (body.go:196-202), emits collision pairs and adds the
// FP64 contributions of the pairs the gate lets through.
#include <cfloat>

#include "nb_internal.hpp"

namespace nb {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) f4v *cf4p;
typedef const __attribute__((address_space(4))) float *cfp;

// Round a non-negative double up to a float (next representable float when conversion rounded
// down).
__device__ __forceinline__ float rup(double v) {
  float f = (float)v;
  if ((double)f < v) f = __uint_as_float(__float_as_uint(f) + 1u);
  return f;
}

template <int IPL>
struct TargetState {
  float xi[IPL], yi[IPL], zi[IPL];
  float sx[IPL], sy[IPL], sz[IPL];  // Kahan sums
  float cx[IPL], cy[IPL], cz[IPL];  // Kahan compensations
  float ri[IPL], amax[IPL];
  int fc[IPL];
  int f0[IPL], f1[IPL], f2[IPL], f3[IPL];
};

template <int IPL, bool DETECT>
__device__ __forceinline__ void tile_end(TargetState<IPL> &T, const float (&px)[IPL],
                                         const float (&py)[IPL], const float (&pz)[IPL],
                                         const float (&rmin)[IPL], float rmax, float eps2,
                                         int drop, int t) {
#pragma unroll
  for (int k = 0; k < IPL; k++) {
    bool flagged = false;
    if (DETECT) {
      // |dx_fp32 - dx| <= 2^-23 (|x_i|inf + |dx|) per component, and |dx| <= Ri + Rj for any
      // overlapping pair; 2^-21 bounds the vector norm of that error.  The final factor covers
      // the three roundings of the fma chain and of this bound's own arithmetic.
      float rs = T.ri[k] + rmax;
      float e = 4.76837158e-7f * (T.amax[k] + rs);
      float a = rs + e;
      float thr = __builtin_fmaf(a, a, eps2) * 1.0000039f;
      flagged = rmin[k] <= thr;
      if (flagged) {
        int c = T.fc[k];
        if (c == 0) T.f0[k] = t;
        else if (c == 1) T.f1[k] = t;
        else if (c == 2) T.f2[k] = t;
        else if (c == 3) T.f3[k] = t;
        T.fc[k] = c + 1;
      }
    }
    if (!(flagged && drop)) {
      float y, s;
      y = px[k] - T.cx[k]; s = T.sx[k] + y; T.cx[k] = (s - T.sx[k]) - y; T.sx[k] = s;
      y = py[k] - T.cy[k]; s = T.sy[k] + y; T.cy[k] = (s - T.sy[k]) - y; T.sy[k] = s;
      y = pz[k] - T.cz[k]; s = T.sz[k] + y; T.cz[k] = (s - T.sz[k]) - y; T.sz[k] = s;
    }
  }
}

// One source against the lane's IPL targets, written phase by phase across the targets so that
// each v_rsq_f32 result is consumed several instructions later (no transcendental-use stall).
#define NB_INTERACT(bx, by, bz, bw)                                   \
  {                                                                    \
    float dx[IPL], dy[IPL], dz[IPL], r2[IPL], inv[IPL];                \
    _Pragma("unroll") for (int k = 0; k < IPL; k++) {                  \
      dx[k] = (bx) - T.xi[k];                                          \
      dy[k] = (by) - T.yi[k];                                          \
      dz[k] = (bz) - T.zi[k];                                          \
      r2[k] = __builtin_fmaf(dx[k], dx[k], eps2);                      \
      r2[k] = __builtin_fmaf(dy[k], dy[k], r2[k]);                     \
      r2[k] = __builtin_fmaf(dz[k], dz[k], r2[k]);                     \
      inv[k] = __builtin_amdgcn_rsqf(r2[k]);                           \
    }                                                                  \
    _Pragma("unroll") for (int k = 0; k < IPL; k++) {                  \
      float s = (bw) * inv[k];                                         \
      float inv2 = inv[k] * inv[k];                                    \
      s = s * inv2;                                                    \
      px[k] = __builtin_fmaf(dx[k], s, px[k]);                         \
      py[k] = __builtin_fmaf(dy[k], s, py[k]);                         \
      pz[k] = __builtin_fmaf(dz[k], s, pz[k]);                         \
      if (DETECT) rmin[k] = __builtin_fminf(rmin[k], r2[k]);           \
    }                                                                  \
  }

// With eps > 0 the self pair contributes exactly zero force but its r^2 = eps^2 would flag every
// target's own tile.  For that one tile the minimum is recomputed without the self pair (same
// FP32 operations, so the other pairs' r^2 are bit-identical).  A wave's targets are 64
// consecutive bodies, so the branch is wave-uniform for tile_j >= 64.
template <int IPL, bool DETECT, class Src>
__device__ __forceinline__ void self_tile_rmin(const TargetState<IPL> &T, float (&rmin)[IPL],
                                               int base, int stride, int t, int tile_j,
                                               float eps2, Src src) {
  if (!DETECT || !(eps2 > 0.f)) return;
#pragma unroll
  for (int k = 0; k < IPL; k++) {
    const int sl = base + k * stride - t * tile_j;
    if (sl >= 0 && sl < tile_j) {
      float mn = FLT_MAX;
      for (int jj = 0; jj < tile_j; jj++) {
        if (jj == sl) continue;
        const f4v b = src(jj);
        const float dx = b.x - T.xi[k], dy = b.y - T.yi[k], dz = b.z - T.zi[k];
        float r2 = __builtin_fmaf(dx, dx, eps2);
        r2 = __builtin_fmaf(dy, dy, r2);
        r2 = __builtin_fmaf(dz, dz, r2);
        mn = __builtin_fminf(mn, r2);
      }
      rmin[k] = mn;
    }
  }
}

template <int IPL, bool DETECT>
__device__ __forceinline__ void load_targets(TargetState<IPL> &T, const ForceArgs &a, int base,
                                             int stride) {
#pragma unroll
  for (int k = 0; k < IPL; k++) {
    const int i = base + k * stride;
    float4 b = a.js[i];
    T.xi[k] = b.x;
    T.yi[k] = b.y;
    T.zi[k] = b.z;
    T.sx[k] = T.sy[k] = T.sz[k] = 0.f;
    T.cx[k] = T.cy[k] = T.cz[k] = 0.f;
    T.fc[k] = 0;
    T.f0[k] = T.f1[k] = T.f2[k] = T.f3[k] = -1;
    if (DETECT) {
      T.ri[k] = i < a.n ? rup(a.r[i]) : 0.f;
      T.amax[k] = fmaxf(fabsf(b.x), fmaxf(fabsf(b.y), fabsf(b.z)));
    }
  }
}

template <int IPL, bool DETECT>
__device__ __forceinline__ void store_targets(const TargetState<IPL> &T, const ForceArgs &a,
                                              int base, int stride) {
  const int slice = blockIdx.y;
#pragma unroll
  for (int k = 0; k < IPL; k++) {
    const int i = base + k * stride;
    double *o = a.acc + (size_t)slice * 3 * a.npad;
    o[i] = (double)T.sx[k] - (double)T.cx[k];
    o[a.npad + i] = (double)T.sy[k] - (double)T.cy[k];
    o[2 * a.npad + i] = (double)T.sz[k] - (double)T.cz[k];
    if (DETECT) {
      const size_t q = (size_t)slice * a.npad + i;
      a.fcnt[q] = T.fc[k];
      int4 l = make_int4(T.f0[k], T.f1[k], T.f2[k], T.f3[k]);
      reinterpret_cast<int4 *>(a.flist)[q] = l;
      if (T.fc[k] > 0 && i < a.n) {
        const uint32_t r = atomicAdd(&a.ctr->rec_count, 1u);
        if ((int64_t)r < a.rec_cap) a.recs[r] = make_int2(slice, i);
      }
    }
  }
}

// Variant 1: the source stream is read through the scalar data cache.  Every lane of a wave
// needs the same source body, so the loads are wave-uniform s_load_dwordx4 into SGPRs and the
// VALU ops take the source operands from SGPRs; no LDS and no barriers.
template <int IPL, bool DETECT>
__global__ __launch_bounds__(256) void k_force_scalar(ForceArgs a) {
  if (a.run_flag && *a.run_flag == 0u) return;  // conditional launch (surrogate fallback)
  TargetState<IPL> T;
  const int stride = blockDim.x;
  const int base = blockIdx.x * blockDim.x * IPL + threadIdx.x;
  load_targets<IPL, DETECT>(T, a, base, stride);
  const float eps2 = to_vgpr(a.eps2);
  const int t0 = blockIdx.y * a.tiles_per_slice;
  const int t1 = min(t0 + a.tiles_per_slice, a.tiles);
  cf4p js = (cf4p)a.js;
  cfp rmt = (cfp)a.rmax_tile;
  for (int t = t0; t < t1; t++) {
    float px[IPL], py[IPL], pz[IPL], rmin[IPL];
#pragma unroll
    for (int k = 0; k < IPL; k++) px[k] = py[k] = pz[k] = 0.f, rmin[k] = FLT_MAX;
    cf4p jt = js + (size_t)t * a.tile_j;
#pragma unroll 4
    for (int jj = 0; jj < a.tile_j; jj++) {
      const f4v b = jt[jj];
      NB_INTERACT(b.x, b.y, b.z, b.w)
    }
    self_tile_rmin<IPL, DETECT>(T, rmin, base, stride, t, a.tile_j, eps2,
                                [&](int jj) -> f4v { return jt[jj]; });
    tile_end<IPL, DETECT>(T, px, py, pz, rmin, DETECT ? rmt[t] : 0.f, eps2, a.drop_flagged, t);
  }
  store_targets<IPL, DETECT>(T, a, base, stride);
}

// Variant 2: source tiles are staged in LDS by the workgroup (one coalesced 16-byte load per
// lane per row), then read back as wave-uniform broadcasts (ds_read_b128, no bank conflicts).
constexpr int kLdsTileMax = 1024;
template <int IPL, bool DETECT>
__global__ __launch_bounds__(256) void k_force_lds(ForceArgs a) {
  __shared__ float4 sh[kLdsTileMax];
  if (a.run_flag && *a.run_flag == 0u) return;  // conditional launch (surrogate fallback)
  TargetState<IPL> T;
  const int stride = blockDim.x;
  const int base = blockIdx.x * blockDim.x * IPL + threadIdx.x;
  load_targets<IPL, DETECT>(T, a, base, stride);
  const float eps2 = to_vgpr(a.eps2);
  const int t0 = blockIdx.y * a.tiles_per_slice;
  const int t1 = min(t0 + a.tiles_per_slice, a.tiles);
  cfp rmt = (cfp)a.rmax_tile;
  for (int t = t0; t < t1; t++) {
    __syncthreads();
    const float4 *src = a.js + (size_t)t * a.tile_j;
    for (int q = threadIdx.x; q < a.tile_j; q += blockDim.x) sh[q] = src[q];
    __syncthreads();
    float px[IPL], py[IPL], pz[IPL], rmin[IPL];
#pragma unroll
    for (int k = 0; k < IPL; k++) px[k] = py[k] = pz[k] = 0.f, rmin[k] = FLT_MAX;
    // tile_j is a multiple of 64.  Eight sources per step are read first (eight ds_read_b128 in
    // flight, one LDS base address with immediate offsets), then used in order.
    for (int jj = 0; jj < a.tile_j; jj += 8) {
      float4 b8[8];
#pragma unroll
      for (int u = 0; u < 8; u++) b8[u] = sh[jj + u];
#pragma unroll
      for (int u = 0; u < 8; u++) NB_INTERACT(b8[u].x, b8[u].y, b8[u].z, b8[u].w)
    }
    self_tile_rmin<IPL, DETECT>(T, rmin, base, stride, t, a.tile_j, eps2, [&](int jj) -> f4v {
      const float4 b = sh[jj];
      return f4v{b.x, b.y, b.z, b.w};
    });
    tile_end<IPL, DETECT>(T, px, py, pz, rmin, DETECT ? rmt[t] : 0.f, eps2, a.drop_flagged, t);
  }
  store_targets<IPL, DETECT>(T, a, base, stride);
}

template <int IPL, bool DETECT>
static hipError_t launch_t(const ForceShape &s, const ForceArgs &a, hipStream_t st) {
  dim3 grid((unsigned)(a.npad / (s.block * IPL)), (unsigned)s.jsplit);
  if (s.variant == 2)
    hipLaunchKernelGGL((k_force_lds<IPL, DETECT>), grid, dim3(s.block), 0, st, a);
  else
    hipLaunchKernelGGL((k_force_scalar<IPL, DETECT>), grid, dim3(s.block), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_force_f32(const ForceShape &s, const ForceArgs &a, hipStream_t st) {
  if (s.block < 64 || s.block > 256 || (s.block & 63) || a.npad % (s.block * s.ipl) ||
      a.tile_j <= 0 || a.tile_j > kLdsTileMax || a.npad % a.tile_j || s.jsplit < 1 ||
      s.jsplit > kMaxSplit)
    return hipErrorInvalidValue;
  switch (s.ipl * 2 + (a.detect ? 1 : 0)) {
    case 2: return launch_t<1, false>(s, a, st);
    case 3: return launch_t<1, true>(s, a, st);
    case 4: return launch_t<2, false>(s, a, st);
    case 5: return launch_t<2, true>(s, a, st);
    case 8: return launch_t<4, false>(s, a, st);
    case 9: return launch_t<4, true>(s, a, st);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------------------------
// FP64 fix-up of flagged tiles: one thread per (slice, target).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ bool active_body(uint8_t f) {
  return (f & NB_F_EXISTS) && !(f & NB_F_FRAGMENTING);
}
__device__ __forceinline__ bool can_collide(int32_t b) {
  return b == NB_CB_ELASTIC || b == NB_CB_FRAGMENT;
}

__device__ void emit_pair(uint64_t *pairs, DevCounters *ctr, int64_t cap, int i, int j) {
  uint32_t q = atomicAdd(&ctr->pair_count, 1u);
  if ((int64_t)q < cap)
    pairs[q] = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
  else
    atomicOr(&ctr->pair_overflow, 1u);
}

// One wavefront per record: the 64 lanes split each flagged tile's sources, and the FP64 partial
// sums are combined with a fixed butterfly, so the result does not depend on scheduling.  The
// grid is fixed; waves stride over the record count the force kernel left in device memory.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_fixup(FixupArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
  const int nrec = (int)min((int64_t)a.ctr->rec_count, a.rec_cap);
  for (int rr = wave; rr < nrec; rr += nwaves) {
    const int2 rec = a.recs[rr];
    const int slice = rec.x, i = rec.y;
    if (!active_body(a.flags[i])) continue;
    const size_t q = (size_t)slice * a.npad + i;
    const int cnt = a.fcnt[q];
    const int4 lst = reinterpret_cast<const int4 *>(a.flist)[q];
    const int t0 = slice * a.tiles_per_slice;
    const int t1 = min(t0 + a.tiles_per_slice, a.tiles);
    const bool all = cnt > kFlagSlots;
    const int ntile = all ? (t1 - t0) : cnt;
    const double xi = a.x[i], yi = a.y[i], zi = a.z[i], ri = a.r[i];
    const int32_t bi = a.beh[i];
    const bool ci = nb_may_pair(bi, a.events);
    double ax = 0, ay = 0, az = 0;
    for (int u = 0; u < ntile; u++) {
      int t;
      if (all) t = t0 + u;
      else t = u == 0 ? lst.x : u == 1 ? lst.y : u == 2 ? lst.z : lst.w;
      const int j0 = t * a.tile_j;
      const int j1 = min(j0 + a.tile_j, a.n);
      for (int j = j0 + lane; j < j1; j += 64) {
        if (j == i || !active_body(a.flags[j])) continue;
        const double dx = a.x[j] - xi;
        const double dy = a.y[j] - yi;
        const double dz = a.z[j] - zi;
        const double d2 = dx * dx + dy * dy + dz * dz;
        const double dist = __builtin_sqrt(d2);
        const double rj = a.r[j];
        const bool overlap = !(dist > ri + rj);
        if (overlap && a.collisions && i < j && ci &&
            nb_pair_event(bi, a.beh[j], ri, rj, dist, a.events))
          emit_pair(a.pairs, a.ctr, a.pair_cap, i, j);
        if (a.recompute && !(overlap && a.cutoff)) {
          const double ds = d2 + a.eps2;
          const double w = a.m[j] / (ds * __builtin_sqrt(ds));
          ax += w * dx;
          ay += w * dy;
          az += w * dz;
        }
      }
    }
    if (a.recompute) {
      ax = wave_sum(ax);
      ay = wave_sum(ay);
      az = wave_sum(az);
    }
    if (lane == 0) {
      atomicAdd((unsigned long long *)&a.ctr->fixup_tiles, (unsigned long long)ntile);
      if (a.recompute) {
        double *o = a.acc + (size_t)slice * 3 * a.npad;
        if (all) {
          o[i] = ax;
          o[a.npad + i] = ay;
          o[2 * a.npad + i] = az;
        } else {
          o[i] += ax;
          o[a.npad + i] += ay;
          o[2 * a.npad + i] += az;
        }
      }
    }
  }
}

hipError_t launch_fixup(const FixupArgs &a, hipStream_t st) {
  // 8 waves per CU on 256 CUs; the kernel strides over however many records there are.
  hipLaunchKernelGGL(k_fixup, dim3(512), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace nb
