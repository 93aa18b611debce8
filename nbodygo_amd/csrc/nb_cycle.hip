// nb_cycle.hip -- the rest of one compute cycle on gfx950: FP64 reference-order force, pair sort,
// elastic collision resolution, FP64 update, source-stream maintenance, diagnostics.
//
// Compiled with -ffp-contract=off: every FP64 expression below is evaluated as written, one
// correctly rounded IEEE operation at a time (gfx950's FP64 divide and square root are correctly
// rounded), which is what makes the FP64 path bit-identical to the reference's arithmetic.
#include <cfloat>
#include <cmath>

#include <algorithm>

#include "nb_internal.hpp"

namespace nb {

typedef const __attribute__((address_space(4))) double *cdp;
typedef const __attribute__((address_space(4))) int32_t *cip;

__device__ __forceinline__ bool active_f(uint8_t f) {
  return (f & NB_F_EXISTS) && !(f & NB_F_FRAGMENTING);
}
__device__ __forceinline__ bool collide_b(int32_t b) {
  return b == NB_CB_ELASTIC || b == NB_CB_FRAGMENT;
}

__device__ static void emit(uint64_t *pairs, DevCounters *ctr, int64_t cap, int i, int j) {
  uint32_t q = atomicAdd(&ctr->pair_count, 1u);
  if ((int64_t)q < cap)
    pairs[q] = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
  else
    atomicOr(&ctr->pair_overflow, 1u);
}

// ---------------------------------------------------------------------------------------------
// FP64 force in the reference's operation order, one thread per target, sources in index order.
//   eps == 0 && cutoff: calcForceFrom (body.go:214-225)
//       dist = sqrt(dx*dx+dy*dy+dz*dz); if dist > Ri+Rj: force = G*mi*mj/(dist*dist);
//       fx += force*dx/dist
//   otherwise: directForce (octree.go:314-325) with optional gate on the unsoftened distance
//       distSq = d2 + eps*eps; dist = sqrt(distSq); force = G*mi*mj/distSq; fx += force*dx/dist
// Pairs are emitted from the lower index (Collided, body.go:192-208).
// ---------------------------------------------------------------------------------------------
// Sources are staged in LDS 256 at a time by the workgroup and read back as broadcasts; each
// lane still sums its sources in array order, op for op as calcForceFrom.
__global__ __launch_bounds__(256) void k_force_f64(Force64Args a) {
  __shared__ double sx[256], sy[256], sz[256], sm[256], sr[256];
  __shared__ int32_t sb[256];
  __shared__ uint8_t sf[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // npad is a multiple of 256
  double fx = 0, fy = 0, fz = 0;
  const bool ai = i < a.n && active_f(a.flags[i]);
  const bool wave_on = __builtin_amdgcn_ballot_w64(ai) != 0;
  const double xi = ai ? a.x[i] : 0, yi = ai ? a.y[i] : 0, zi = ai ? a.z[i] : 0;
  const double mi = ai ? a.m[i] : 0, ri = ai ? a.r[i] : 0;
  const int32_t bi = ai ? a.beh[i] : NB_CB_NONE;
  const bool ci = ai && nb_may_pair(bi, a.events);
  const double G = a.G, eps2 = a.eps * a.eps;
  const bool ref = a.eps == 0.0 && a.cutoff;
  // the unsoftened distance is only needed for the overlap gate and for detection
  const bool need_d0 = ref || a.cutoff || a.collisions;
  for (int t0 = 0; t0 < a.n; t0 += 256) {
    __syncthreads();
    const int js = t0 + threadIdx.x;
    if (js < a.n) {
      sx[threadIdx.x] = a.x[js];
      sy[threadIdx.x] = a.y[js];
      sz[threadIdx.x] = a.z[js];
      sm[threadIdx.x] = a.m[js];
      sr[threadIdx.x] = a.r[js];
      sb[threadIdx.x] = a.beh[js];
      sf[threadIdx.x] = a.flags[js];
    }
    __syncthreads();
    if (!wave_on) continue;
    const int cnt = min(256, a.n - t0);
    for (int jj = 0; jj < cnt; jj++) {
      if (!active_f(sf[jj])) continue;
      const int j = t0 + jj;
      const double dx = sx[jj] - xi;
      const double dy = sy[jj] - yi;
      const double dz = sz[jj] - zi;
      const double d2 = dx * dx + dy * dy + dz * dz;
      const double rs = ri + sr[jj];
      if (j != i && ai) {
        const double dist0 = need_d0 ? __builtin_sqrt(d2) : 0.0;
        const bool overlap = need_d0 && !(dist0 > rs);
        if (ref) {
          if (!overlap) {
            const double force = G * mi * sm[jj] / (dist0 * dist0);
            fx += force * dx / dist0;
            fy += force * dy / dist0;
            fz += force * dz / dist0;
          }
        } else {
          if (!(a.cutoff && overlap)) {
            const double distSq = d2 + eps2;
            const double dist = __builtin_sqrt(distSq);
            const double force = G * mi * sm[jj] / distSq;
            fx += force * dx / dist;
            fy += force * dy / dist;
            fz += force * dz / dist;
          }
        }
        if (a.collisions && overlap && i < j && ci &&
            nb_pair_event(bi, sb[jj], ri, sr[jj], dist0, a.events))
          emit(a.pairs, a.ctr, a.pair_cap, i, j);
      }
    }
  }
  a.acc[i] = fx;
  a.acc[a.npad + i] = fy;
  a.acc[2 * a.npad + i] = fz;
}

hipError_t launch_force_f64(const Force64Args &a, hipStream_t st) {
  hipLaunchKernelGGL(k_force_f64, dim3((unsigned)(a.npad / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// FP64 accelerations for a sample of targets (error reference for the FP32 path).
__global__ __launch_bounds__(256) void k_accel_f64(Accel64Args a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nidx) return;
  const int64_t i = a.idx[k];
  double ax = 0, ay = 0, az = 0;
  if (i >= 0 && i < a.n && active_f(a.flags[i])) {
    const double xi = a.x[i], yi = a.y[i], zi = a.z[i], ri = a.r[i];
    const double eps2 = a.eps * a.eps;
    for (int j = 0; j < a.n; j++) {
      if (j == i || !active_f(a.flags[j])) continue;
      const double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (a.cutoff && !(__builtin_sqrt(d2) > ri + a.r[j])) continue;
      const double ds = d2 + eps2;
      const double w = a.G * a.m[j] / (ds * __builtin_sqrt(ds));
      ax += w * dx;
      ay += w * dy;
      az += w * dz;
    }
  }
  a.out[k] = ax;
  a.out[a.nidx + k] = ay;
  a.out[2 * a.nidx + k] = az;
}

hipError_t launch_accel_f64(const Accel64Args &a, hipStream_t st) {
  if (a.nidx <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_accel_f64, dim3((unsigned)((a.nidx + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Pair sort: keys (i << 32 | j) ascending = canonical order.  The count is known on the device
// only, so the host launches a fixed sequence sized by the capacity and each kernel reads the
// count: kSortLdsKeys-key chunks are sorted in LDS by one workgroup each (one workgroup sorts the
// whole list when it fits), then runs of doubling length are merged (each key finds its place by
// a binary search in the other run), one launch per run length up to the capacity.  A pass whose
// runs already span the list copies it, so the buffers alternate in a fixed order and the result
// always ends in the key buffer.
// ---------------------------------------------------------------------------------------------
// Sorts sh[0, P) ascending (P a power of two) with the block's threads.
__device__ static void bitonic_lds(uint64_t *sh, int64_t P) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int64_t k = 2; k <= P; k <<= 1) {
    for (int64_t j = k >> 1; j > 0; j >>= 1) {
      for (int64_t q = tid; q < P; q += nt) {
        const int64_t o = q ^ j;
        if (o > q) {
          const uint64_t u = sh[q], v = sh[o];
          const bool up = (q & k) == 0;
          if ((u > v) == up) {
            sh[q] = v;
            sh[o] = u;
          }
        }
      }
      __syncthreads();
    }
  }
}

// src[base, base + cnt) (cnt <= kSortLdsKeys) sorted through LDS into dst[base, base + cnt).
__device__ static void sort_chunk_lds(uint64_t *sh, const uint64_t *src, uint64_t *dst,
                                      int64_t base, int64_t cnt) {
  const int tid = threadIdx.x, nt = blockDim.x;
  int64_t P = 1;
  while (P < cnt) P <<= 1;
  for (int64_t q = tid; q < P; q += nt) sh[q] = q < cnt ? src[base + q] : ~0ull;
  __syncthreads();
  bitonic_lds(sh, P);
  for (int64_t q = tid; q < cnt; q += nt) dst[base + q] = sh[q];
}

__global__ __launch_bounds__(1024) void k_sort_chunks(const uint64_t *keys, uint64_t *dst,
                                                      int64_t cap, const DevCounters *ctr) {
  __shared__ uint64_t sh[kSortLdsKeys];
  const int64_t cnt = min((int64_t)ctr->pair_count, cap);
  const int64_t base = (int64_t)blockIdx.x * kSortLdsKeys;
  if (base >= cnt) return;
  sort_chunk_lds(sh, keys, dst, base, min((int64_t)kSortLdsKeys, cnt - base));
}

// Number of entries of the sorted run r[0, n) below key (strict) or, with upper, not above it.
__device__ __forceinline__ int64_t run_rank(const uint64_t *r, int64_t n, uint64_t key,
                                            bool upper) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    const uint64_t v = r[mid];
    if (v < key || (upper && v == key)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Runs of length L in src merged pairwise into dst (copied when one run spans the list).
__global__ __launch_bounds__(256) void k_merge_pass(const uint64_t *src, uint64_t *dst,
                                                    int64_t L, int64_t cap,
                                                    const DevCounters *ctr) {
  const int64_t cnt = min((int64_t)ctr->pair_count, cap);
  const int64_t gs = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gs) {
    const uint64_t key = src[q];
    if (cnt <= L) {
      dst[q] = key;
      continue;
    }
    const int64_t r = q / L, m0 = (r & ~(int64_t)1) * L;
    if (!(r & 1)) {  // left run: place after the right run's smaller keys
      const int64_t rb = m0 + L;
      dst[rb < cnt ? q + run_rank(src + rb, min(L, cnt - rb), key, false) : q] = key;
    } else {  // right run: after the left run's keys not above it (stable)
      dst[q - L + run_rank(src + m0, L, key, true)] = key;
    }
  }
}

hipError_t launch_sort_pairs(uint64_t *keys, uint64_t *tmp, int64_t cap, const DevCounters *ctr,
                             hipStream_t st) {
  int passes = 0;
  for (int64_t L = kSortLdsKeys; L < cap; L <<= 1) passes++;
  // the chunks go to whichever buffer makes the last pass land in keys
  uint64_t *buf[2] = {(passes & 1) ? tmp : keys, (passes & 1) ? keys : tmp};
  const unsigned chunks = (unsigned)((cap + kSortLdsKeys - 1) / kSortLdsKeys);
  hipLaunchKernelGGL(k_sort_chunks, dim3(chunks), dim3(1024), 0, st, keys, buf[0], cap, ctr);
  const unsigned g = (unsigned)std::min<int64_t>(2048, (cap + 255) / 256);
  int64_t L = kSortLdsKeys;
  for (int p = 0; p < passes; p++, L <<= 1)
    hipLaunchKernelGGL(k_merge_pass, dim3(g), dim3(256), 0, st, buf[p & 1], buf[(p + 1) & 1], L,
                       cap, ctr);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Elastic collision resolution (ProcessMods, body_collection.go:212-233), one thread walking the
// sorted pair list: each unordered pair is resolved once, as lo.ResolveCollision(hi) when lo is
// Elastic, else hi.ResolveCollision(lo) (only an Elastic first body resolves, body.go:252-253).
// ---------------------------------------------------------------------------------------------
struct Elastic {
  double v1[3], v2[3], cm[3];
};

// calcElasticCollision, cmd/body/collisioncalc.go:42-186 (coll3d).
__device__ static bool elastic(const ResolveArgs &a, int b, int o, Elastic &res) {
  double r12, m21, d, v, theta2, phi2, st, ct, sp, cp, vx1r, vy1r, vz1r, fvz1r, thetav, phiv, dr,
      alpha, beta, sbeta, cbeta, aa, dvz2, vx2r, vy2r, vz2r, x21, y21, z21, vx21, vy21, vz21,
      vx_cm, vy_cm, vz_cm;
  const double m1 = a.m[b], m2 = a.m[o];
  const double r1 = a.r[b], r2 = a.r[o];
  const double x1 = a.x[b], y1 = a.y[b], z1 = a.z[b];
  double x2 = a.x[o], y2 = a.y[o], z2 = a.z[o];
  double vx1 = a.vx[b], vy1 = a.vy[b], vz1 = a.vz[b];
  const double vx2 = a.vx[o], vy2 = a.vy[o], vz2 = a.vz[o];

  r12 = r1 + r2;
  m21 = m2 / m1;
  x21 = x2 - x1;
  y21 = y2 - y1;
  z21 = z2 - z1;
  vx21 = vx2 - vx1;
  vy21 = vy2 - vy1;
  vz21 = vz2 - vz1;

  vx_cm = (m1 * vx1 + m2 * vx2) / (m1 + m2);
  vy_cm = (m1 * vy1 + m2 * vy2) / (m1 + m2);
  vz_cm = (m1 * vz1 + m2 * vz2) / (m1 + m2);

  d = sqrt(x21 * x21 + y21 * y21 + z21 * z21);
  v = sqrt(vx21 * vx21 + vy21 * vy21 + vz21 * vz21);
  if (v == 0) return false;

  x2 = x21;
  y2 = y21;
  z2 = z21;
  vx1 = -vx21;
  vy1 = -vy21;
  vz1 = -vz21;

  theta2 = acos(z2 / d);
  if (x2 == 0 && y2 == 0)
    phi2 = 0;
  else
    phi2 = atan2(y2, x2);
  st = sin(theta2);
  ct = cos(theta2);
  sp = sin(phi2);
  cp = cos(phi2);

  vx1r = ct * cp * vx1 + ct * sp * vy1 - st * vz1;
  vy1r = cp * vy1 - sp * vx1;
  vz1r = st * cp * vx1 + st * sp * vy1 + ct * vz1;
  fvz1r = vz1r / v;
  if (fvz1r > 1)
    fvz1r = 1;
  else if (fvz1r < -1)
    fvz1r = -1;
  thetav = acos(fvz1r);
  if (vx1r == 0 && vy1r == 0)
    phiv = 0;
  else
    phiv = atan2(vy1r, vx1r);

  dr = d * sin(thetav) / r12;
  if (thetav > M_PI / 2 || fabs(dr) > 1) return false;

  alpha = asin(-dr);
  beta = phiv;
  sbeta = sin(beta);
  cbeta = cos(beta);

  aa = tan(thetav + alpha);
  dvz2 = 2 * (vz1r + aa * (cbeta * vx1r + sbeta * vy1r)) / ((1 + aa * aa) * (1 + m21));

  vz2r = dvz2;
  vx2r = aa * cbeta * dvz2;
  vy2r = aa * sbeta * dvz2;
  vz1r = vz1r - m21 * vz2r;
  vx1r = vx1r - m21 * vx2r;
  vy1r = vy1r - m21 * vy2r;

  res.v1[0] = ct * cp * vx1r - sp * vy1r + st * cp * vz1r + vx2;
  res.v1[1] = ct * sp * vx1r + cp * vy1r + st * sp * vz1r + vy2;
  res.v1[2] = ct * vz1r - st * vx1r + vz2;
  res.v2[0] = ct * cp * vx2r - sp * vy2r + st * cp * vz2r + vx2;
  res.v2[1] = ct * sp * vx2r + cp * vy2r + st * sp * vz2r + vy2;
  res.v2[2] = ct * vz2r - st * vx2r + vz2;
  res.cm[0] = vx_cm;
  res.cm[1] = vy_cm;
  res.cm[2] = vz_cm;
  return true;
}

// shouldFragment (fragcalc.go:24-49); the factors go to doFragment.
__device__ static bool should_fragment(const ResolveArgs &a, int b, int o, const Elastic &e,
                                       double *tf, double *of) {
  if (!(a.beh[b] == NB_CB_FRAGMENT || a.beh[o] == NB_CB_FRAGMENT)) return false;
  const double br = a.rest[b];
  const double vThis = a.vx[b] + a.vy[b] + a.vz[b];
  const double dvThis = fabs(a.vx[b] - ((e.v1[0] - e.cm[0]) * br + e.cm[0])) +
                        fabs(a.vy[b] - ((e.v1[1] - e.cm[1]) * br + e.cm[1])) +
                        fabs(a.vz[b] - ((e.v1[2] - e.cm[2]) * br + e.cm[2]));
  const double thisFactor = dvThis / fabs(vThis);
  const double vOther = a.vx[o] + a.vy[o] + a.vz[o];
  const double dvOther = fabs(a.vx[o] - ((e.v2[0] - e.cm[0]) * br + e.cm[0])) +
                         fabs(a.vy[o] - ((e.v2[1] - e.cm[1]) * br + e.cm[1])) +
                         fabs(a.vz[o] - ((e.v2[2] - e.cm[2]) * br + e.cm[2]));
  const double otherFactor = dvOther / fabs(vOther);
  const double fb = a.frag[b], fo = a.frag[o];
  *tf = thisFactor;
  *of = otherFactor;
  return (a.beh[b] == NB_CB_FRAGMENT && thisFactor > fb) ||
         (a.beh[o] == NB_CB_FRAGMENT && otherFactor > fo);
}

// initiateFragmentation (fragcalc.go:65-83): fragDelta, the fragment count Min(delta * FragStep,
// 2000), and fragInfo.  The body keeps this cycle's force while it fragments.
__device__ static void initiate_fragmentation(const ResolveArgs &a, int b, double factor,
                                              uint64_t *started, uint64_t *inactive) {
  const double fragDelta = a.frag[b] > 10 ? 10.0 : factor - a.frag[b];
  const double prod = fragDelta * a.fstep[b];
  const double fragments = isnan(prod) ? prod : (prod < 2000.0 ? prod : 2000.0);  // math.Min
  if (fragments <= 1) {
    a.beh[b] = NB_CB_FRAGMENT;
    return;
  }
  const uint8_t f = a.flags[b];
  if (!(f & NB_F_FRAGMENTING)) {
    (*started)++;
    if (f & NB_F_EXISTS) (*inactive)++;
  }
  a.flags[b] = (uint8_t)(f | NB_F_FRAGMENTING);
  FragState st;
  double fx = 0, fy = 0, fz = 0;
  for (int sl = 0; sl < a.split; sl++) {
    const double *o = a.acc + (size_t)sl * 3 * a.npad;
    fx += o[b];
    fy += o[a.npad + b];
    fz += o[2 * a.npad + b];
  }
  st.fx = fx;
  st.fy = fy;
  st.fz = fz;
  st.mass = a.m[b] / fragments;
  st.radius = a.r[b];
  st.px = a.x[b];
  st.py = a.y[b];
  st.pz = a.z[b];
  // int(fragments): truncation; NaN converts to the most negative int64 on amd64
  st.left = isnan(fragments) ? INT64_MIN : (int64_t)fragments;
  st.made = 0;
  st.id = (int64_t)(a.ctr->frag_seq++);
  a.fst[b] = st;
}

// Tallies of one resolution pass.
struct ResolveTally {
  uint64_t resolved = 0, fragreq = 0, subsumed = 0, started = 0, inactive = 0;
};

// Whether the pair (lo, hi) can change anything when events are off: an Elastic first body
// (body.go:252-253) and both bodies existing (:249-251).  With events off nothing in the pass
// changes behaviours or flags, so this is decided once per pass.  b receives the first body.
__device__ static bool elastic_pair(const ResolveArgs &a, int lo, int hi, int *b, int *o) {
  if (a.beh[lo] == NB_CB_ELASTIC && collide_b(a.beh[hi])) {
    *b = lo;
    *o = hi;
  } else if (a.beh[hi] == NB_CB_ELASTIC && collide_b(a.beh[lo])) {
    *b = hi;
    *o = lo;
  } else {
    return false;
  }
  return (a.flags[*b] & NB_F_EXISTS) && (a.flags[*o] & NB_F_EXISTS);
}

// One canonical pair, as ProcessMods handles its event (body_collection.go:212-233).
__device__ static void resolve_pair(const ResolveArgs &a, int lo, int hi, ResolveTally &t) {
  if (a.events && !(collide_b(a.beh[lo]) && collide_b(a.beh[hi]))) {
    // a subsume pair: ResolveSubsume (body.go:228-244) by the larger body, no existence check
    const int b = a.r[lo] > a.r[hi] ? lo : hi, o = b == lo ? hi : lo;
    a.m[b] = a.m[b] + a.m[o];
    const uint8_t fo = a.flags[o];
    if ((fo & NB_F_EXISTS) && !(fo & NB_F_FRAGMENTING)) t.inactive++;
    a.flags[o] = (uint8_t)(fo & ~NB_F_EXISTS);  // SetNotExists (body.go:93-96)
    a.m[o] = 0.0;
    t.subsumed++;
    return;
  }
  int b, o;
  if (!elastic_pair(a, lo, hi, &b, &o)) return;
  Elastic e;
  if (!elastic(a, b, o, e)) return;
  t.resolved++;
  double tf = 0, of = 0;
  if (should_fragment(a, b, o, e, &tf, &of)) {
    t.fragreq++;
    if (a.events) {  // doFragment (fragcalc.go:54-63)
      if (a.beh[b] == NB_CB_FRAGMENT && tf > a.frag[b])
        initiate_fragmentation(a, b, tf, &t.started, &t.inactive);
      if (a.beh[o] == NB_CB_FRAGMENT && of > a.frag[o])
        initiate_fragmentation(a, o, of, &t.started, &t.inactive);
    }
    return;
  }
  // doElastic (collisioncalc.go:26-35): restitution of the first body.
  const double rr = a.rest[b];
  a.vx[b] = (e.v1[0] - e.cm[0]) * rr + e.cm[0];
  a.vy[b] = (e.v1[1] - e.cm[1]) * rr + e.cm[1];
  a.vz[b] = (e.v1[2] - e.cm[2]) * rr + e.cm[2];
  a.vx[o] = (e.v2[0] - e.cm[0]) * rr + e.cm[0];
  a.vy[o] = (e.v2[1] - e.cm[1]) * rr + e.cm[1];
  a.vz[o] = (e.v2[2] - e.cm[2]) * rr + e.cm[2];
  a.collided[b] = 1;
  a.collided[o] = 1;
}

__device__ static void add_tally(DevCounters *ctr, uint32_t found, const ResolveTally &t) {
  ctr->pairs_total += found;
  ctr->pairs_last = found;
  ctr->resolved_total += t.resolved;
  ctr->fragment_requests += t.fragreq;
  ctr->subsumed_total += t.subsumed;
  ctr->frag_started += t.started;
  ctr->inactive_events += t.inactive;
}

// The pass one pair at a time, in canonical order, on one lane: the ordering reference, and the
// pass used with events on (subsumes and fragmentation ids depend on the order across pairs).
__global__ __launch_bounds__(64) void k_resolve(ResolveArgs a) {
  if (threadIdx.x != 0) return;
  const uint32_t found = a.ctr->pair_count;
  const int64_t cnt = min((int64_t)found, a.pair_cap);
  ResolveTally t;
  for (int64_t k = 0; k < cnt; k++) {
    const uint64_t key = a.pairs[k];
    resolve_pair(a, (int)(key >> 32), (int)(key & 0xffffffffu), t);
  }
  add_tally(a.ctr, found, t);
}

// The same pass with pairs that share no body resolved together (events off).  A pair reads and
// writes only its two bodies, and nothing in the pass changes which pairs act, so the state is
// bit-identical to the one-by-one pass as long as every pair sees both its bodies exactly as the
// pairs before it in canonical order left them.  The pairs touching one body b form a chain in
// canonical order: the pairs (x, b), x ascending, then b's own pairs (b, y), y ascending.  A pair
// therefore waits for at most two predecessors (the previous pair on each of its bodies) and
// releases at most two successors.  The chains come from two sorted lists: the pair keys
// (lo << 32 | hi) and the same keys swapped (hi << 32 | lo), where b's pairs (x, b) are one run.
// Then one workgroup resolves the ready pairs round by round: pairs of one round share no body.
__device__ __forceinline__ int64_t key_lower_bound(const uint64_t *k, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (k[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// keys_sw[k] = the k-th pair key with its halves swapped (sorted afterwards).
__global__ void k_pairs_swap(const uint64_t *keys, uint64_t *keys_sw, int64_t cap,
                             const DevCounters *ctr) {
  const int64_t cnt = min((int64_t)ctr->pair_count, cap);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < cnt;
       k += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t v = keys[k];
    keys_sw[k] = (v << 32) | (v >> 32);
  }
}

// For each pair: its successor on each of its two bodies (-1: none) and how many predecessors it
// waits for.  keys: sorted pair keys; sw: sorted swapped keys.
__global__ void k_pair_dag(const uint64_t *keys, const uint64_t *sw, int64_t cap,
                           const DevCounters *ctr, int32_t *succ, int32_t *pending) {
  const int64_t cnt = min((int64_t)ctr->pair_count, cap);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < cnt;
       k += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t key = keys[k];
    const uint32_t lo = (uint32_t)(key >> 32), hi = (uint32_t)key;
    int npred = 0;
    // chain of lo: k lies in lo's own run (lo, y)
    int32_t s_lo = -1;
    if (k + 1 < cnt && (uint32_t)(keys[k + 1] >> 32) == lo) s_lo = (int32_t)(k + 1);
    if (k > 0 && (uint32_t)(keys[k - 1] >> 32) == lo) {
      npred++;
    } else {  // the first of lo's own run: after the last pair (x, lo), if there is one
      const int64_t e = key_lower_bound(sw, cnt, ((uint64_t)lo + 1) << 32) - 1;
      if (e >= 0 && (uint32_t)(sw[e] >> 32) == lo) npred++;
    }
    // chain of hi: k is the pair (lo, hi) among hi's pairs (x, hi), at position p of sw
    const int64_t p = key_lower_bound(sw, cnt, ((uint64_t)hi << 32) | lo);
    int32_t s_hi = -1;
    if (p > 0 && (uint32_t)(sw[p - 1] >> 32) == hi) npred++;
    if (p + 1 < cnt && (uint32_t)(sw[p + 1] >> 32) == hi) {
      const uint64_t nk = ((uint64_t)(uint32_t)sw[p + 1] << 32) | hi;  // the pair (x', hi)
      s_hi = (int32_t)key_lower_bound(keys, cnt, nk);
    } else {  // the last pair (x, hi): next comes hi's own run (hi, y), if there is one
      const int64_t f = key_lower_bound(keys, cnt, (uint64_t)hi << 32);
      if (f < cnt && (uint32_t)(keys[f] >> 32) == hi) s_hi = (int32_t)f;
    }
    succ[2 * k] = s_lo;
    succ[2 * k + 1] = s_hi;
    pending[k] = npred;
  }
}

__device__ __forceinline__ void sync_all() {
  __threadfence();
  __syncthreads();
}

__global__ __launch_bounds__(1024) void k_resolve_dag(ResolveArgs a) {
  __shared__ int s_len;
  __shared__ unsigned long long s_res, s_frag;
  const int tid = threadIdx.x, nt = blockDim.x;
  const uint32_t found = a.ctr->pair_count;
  const int cnt = (int)min((int64_t)found, a.pair_cap);
  if (tid == 0) s_len = 0, s_res = 0, s_frag = 0;
  __syncthreads();
  int32_t *cur = a.list_a, *nxt = a.list_b;
  for (int k = tid; k < cnt; k += nt)
    if (a.pending[k] == 0) cur[atomicAdd(&s_len, 1)] = k;
  sync_all();
  int m = s_len;
  ResolveTally t;
  while (m > 0) {
    __syncthreads();  // every thread has read s_len
    if (tid == 0) s_len = 0;
    __syncthreads();
    for (int q = tid; q < m; q += nt) {
      const int k = cur[q];
      const uint64_t key = a.pairs[k];
      resolve_pair(a, (int)(key >> 32), (int)(key & 0xffffffffu), t);
      for (int e = 0; e < 2; e++) {
        const int sk = a.succ[2 * k + e];
        if (sk >= 0 && atomicSub(&a.pending[sk], 1) == 1) nxt[atomicAdd(&s_len, 1)] = sk;
      }
    }
    sync_all();
    m = s_len;
    int32_t *sw = cur;
    cur = nxt;
    nxt = sw;
  }
  atomicAdd(&s_res, (unsigned long long)t.resolved);
  atomicAdd(&s_frag, (unsigned long long)t.fragreq);
  __syncthreads();
  if (tid == 0) {
    ResolveTally all;
    all.resolved = s_res;
    all.fragreq = s_frag;
    add_tally(a.ctr, found, all);
  }
}

hipError_t launch_resolve(const ResolveArgs &a, hipStream_t st) {
  if (a.events || a.serial || !a.succ) {
    hipLaunchKernelGGL(k_resolve, dim3(1), dim3(64), 0, st, a);
    return hipGetLastError();
  }
  // the chains: swapped keys, sorted; successors and predecessor counts; then the rounds
  const unsigned g = (unsigned)std::min<int64_t>(4096, (a.pair_cap + 255) / 256);
  hipLaunchKernelGGL(k_pairs_swap, dim3(g), dim3(256), 0, st, a.pairs, a.pairs_sw, a.pair_cap,
                     a.ctr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_sort_pairs(a.pairs_sw, a.sort_tmp, a.pair_cap, a.ctr, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_pair_dag, dim3(g), dim3(256), 0, st, a.pairs, a.pairs_sw, a.pair_cap,
                     a.ctr, a.succ, a.pending);
  hipLaunchKernelGGL(k_resolve_dag, dim3(1), dim3(1024), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Update: Body.Update (body.go:114-139) + rewrite of the FP32 source stream.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float4 stream_entry(bool act, double x, double y, double z, double m) {
  return act ? make_float4((float)x, (float)y, (float)z, (float)m)
             : make_float4(kSentinel, kSentinel, kSentinel, 0.f);
}

__global__ __launch_bounds__(256) void k_update(UpdateArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.npad) return;
  if (i >= a.n) {
    a.js[i] = stream_entry(false, 0, 0, 0, 0);
    return;
  }
  uint8_t f = a.flags[i];
  if (!(f & NB_F_EXISTS)) {
    a.js[i] = stream_entry(false, 0, 0, 0, 0);
    return;
  }
  const double dt = a.dt;
  double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
  if (!a.collided[i]) {
    const double mi = a.m[i];
    double fx = 0, fy = 0, fz = 0;
    if (a.fst && (f & NB_F_FRAGMENTING)) {  // the force Body.fx kept from its last Compute
      fx = a.fst[i].fx;
      fy = a.fst[i].fy;
      fz = a.fst[i].fz;
    } else {
      for (int s = 0; s < a.split; s++) {
        const double *o = a.acc + (size_t)s * 3 * a.npad;
        fx += o[i];
        fy += o[a.npad + i];
        fz += o[2 * a.npad + i];
      }
    }
    if (a.force_mode == 1) {
      // FP64 path: acc holds the force; Vx += timeScaling * fx / Mass (body.go:120-122).
      vx += dt * fx / mi;
      vy += dt * fy / mi;
      vz += dt * fz / mi;
    } else {
      // FP32 path: acc holds sum mj dx / r^3; the kick is dt * G * a.  A massless body gets the
      // reference's 0/0 (its force G*0*mj/d^2 is zero), which the NaN guard below removes.
      if (mi == 0.0) {
        vx = vy = vz = __builtin_nan("");
      } else {
        vx += dt * (a.G * fx);
        vy += dt * (a.G * fy);
        vz += dt * (a.G * fz);
      }
    }
    a.vx[i] = vx;
    a.vy[i] = vy;
    a.vz[i] = vz;
  }
  const double x = a.x[i] + dt * vx;
  const double y = a.y[i] + dt * vy;
  const double z = a.z[i] + dt * vz;
  a.x[i] = x;
  a.y[i] = y;
  a.z[i] = z;
  a.collided[i] = 0;
  a.rest[i] = a.R;
  if (isnan(x) || isnan(y) || isnan(z)) {
    f &= (uint8_t)~NB_F_EXISTS;
    a.flags[i] = f;
    atomicAdd((unsigned long long *)&a.ctr->removed_total, 1ull);
  }
  a.js[i] = stream_entry(active_f(f), x, y, z, a.m[i]);
}

hipError_t launch_update(const UpdateArgs &a, hipStream_t st) {
  hipLaunchKernelGGL(k_update, dim3((unsigned)(a.npad / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

__global__ void k_write_stream(const double *x, const double *y, const double *z,
                               const double *m, const uint8_t *flags, float4 *js, int n,
                               int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  js[i] = i < n ? stream_entry(active_f(flags[i]), x[i], y[i], z[i], m[i])
                : stream_entry(false, 0, 0, 0, 0);
}

hipError_t launch_write_stream(const double *x, const double *y, const double *z,
                               const double *m, const uint8_t *flags, float4 *js, int n,
                               int npad, hipStream_t st) {
  hipLaunchKernelGGL(k_write_stream, dim3((unsigned)(npad / 256)), dim3(256), 0, st, x, y, z, m,
                     flags, js, n, npad);
  return hipGetLastError();
}

// Per-tile max radius over active bodies, rounded up to the next float.
__global__ __launch_bounds__(256) void k_rmax_tile(const double *r, const uint8_t *flags, int n,
                                                   int tile_j, float *out) {
  __shared__ double sh[256];
  const int t = blockIdx.x;
  double mx = 0;
  for (int q = threadIdx.x; q < tile_j; q += blockDim.x) {
    const int j = t * tile_j + q;
    if (j < n && active_f(flags[j])) mx = fmax(mx, r[j]);
  }
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float f = (float)sh[0];
    if ((double)f < sh[0]) f = __uint_as_float(__float_as_uint(f) + 1u);
    out[t] = f;
  }
}

hipError_t launch_rmax_tile(const double *r, const uint8_t *flags, int n, int tile_j, int tiles,
                            float *out, hipStream_t st) {
  hipLaunchKernelGGL(k_rmax_tile, dim3((unsigned)tiles), dim3(256), 0, st, r, flags, n, tile_j,
                     out);
  return hipGetLastError();
}

__global__ void k_zero_counters(DevCounters *c, int full) {
  if (threadIdx.x != 0) return;
  if (full) {
    DevCounters z{};
    *c = z;
  } else {
    c->pair_count = 0;
    c->rec_count = 0;
  }
}

hipError_t launch_zero_counters(DevCounters *ctr, int full, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, ctr, full);
  return hipGetLastError();
}

// Force per body (Body.fx/fy/fz) from the per-slice accumulators.
__global__ void k_combine_force(const double *acc, int split, int npad, int n, const double *m,
                                double G, int force_mode, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double f[3] = {0, 0, 0};
  for (int s = 0; s < split; s++)
    for (int c = 0; c < 3; c++) f[c] += acc[((size_t)s * 3 + c) * npad + i];
  for (int c = 0; c < 3; c++)
    out[(size_t)c * n + i] = force_mode == 1 ? f[c] : m[i] * (G * f[c]);
}

hipError_t launch_combine_force(const double *acc, int split, int npad, int n, const double *m,
                                double G, int force_mode, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_combine_force, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, acc,
                     split, npad, n, m, G, force_mode, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Diagnostics: per-body FP64 terms then a fixed-order tree reduction (deterministic).
// scratch layout: [0, 8*n) per-body terms (ke, pe, px, py, pz, m, active, 0), then the result.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag_terms(const double *x, const double *y,
                                                    const double *z, const double *vx,
                                                    const double *vy, const double *vz,
                                                    const double *m, const uint8_t *flags, int n,
                                                    double eps, double *t) {
  // sources staged 256 at a time in LDS; each body sums its pair terms in ascending j
  __shared__ double sx[256], sy[256], sz[256], sm[256];
  __shared__ uint8_t se[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const bool ex = valid && (flags[i] & NB_F_EXISTS);
  double ke = 0, pe = 0, px = 0, py = 0, pz = 0, mm = 0, act = 0;
  double xi = 0, yi = 0, zi = 0, mi = 0;
  if (ex) {
    mi = m[i];
    xi = x[i];
    yi = y[i];
    zi = z[i];
    ke = 0.5 * mi * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    px = mi * vx[i];
    py = mi * vy[i];
    pz = mi * vz[i];
    mm = mi;
    act = 1;
  }
  const double eps2 = eps * eps;
  for (int j0 = 0; j0 < n; j0 += 256) {
    __syncthreads();
    const int jl = j0 + (int)threadIdx.x;
    if (jl < n) {
      sx[threadIdx.x] = x[jl];
      sy[threadIdx.x] = y[jl];
      sz[threadIdx.x] = z[jl];
      sm[threadIdx.x] = m[jl];
      se[threadIdx.x] = (flags[jl] & NB_F_EXISTS) ? 1 : 0;
    }
    __syncthreads();
    if (!ex) continue;
    const int cnt = min(256, n - j0);
    for (int jj = 0; jj < cnt; jj++) {
      if (j0 + jj == i || !se[jj]) continue;
      const double dx = sx[jj] - xi, dy = sy[jj] - yi, dz = sz[jj] - zi;
      const double d2 = dx * dx + dy * dy + dz * dz + eps2;
      if (d2 > 0) pe -= mi * sm[jj] / __builtin_sqrt(d2);
    }
  }
  if (!valid) return;
  pe *= 0.5;
  double *o = t + (size_t)i * 8;
  o[0] = ke;
  o[1] = pe;
  o[2] = px;
  o[3] = py;
  o[4] = pz;
  o[5] = mm;
  o[6] = act;
  o[7] = 0;
}

__global__ __launch_bounds__(256) void k_diag_reduce(const double *t, int n, double G,
                                                     double *out) {
  __shared__ double sh[256][8];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // Fixed assignment of bodies to threads and a fixed tree: the result does not depend on
  // scheduling.
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    for (int c = 0; c < 8; c++) acc[c] += t[(size_t)i * 8 + c];
  for (int c = 0; c < 8; c++) sh[threadIdx.x][c] = acc[c];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 8; c++) sh[threadIdx.x][c] += sh[threadIdx.x + w][c];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int c = 0; c < 8; c++) out[c] = sh[0][c];
    out[1] = sh[0][1] * G;
  }
}

hipError_t launch_diag(const double *x, const double *y, const double *z, const double *vx,
                       const double *vy, const double *vz, const double *m, const uint8_t *flags,
                       int n, double G, double eps, double *scratch, int64_t scratch_len,
                       hipStream_t st) {
  if (scratch_len < (int64_t)n * 8 + 8) return hipErrorInvalidValue;
  if (n > 0)
    hipLaunchKernelGGL(k_diag_terms, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y,
                       z, vx, vy, vz, m, flags, n, eps, scratch);
  hipLaunchKernelGGL(k_diag_reduce, dim3(1), dim3(256), 0, st, scratch, n, G,
                     scratch + (size_t)n * 8);
  return hipGetLastError();
}

__global__ void k_render(const double *x, const double *y, const double *z, const uint8_t *flags,
                         int n, float *xyz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool e = flags[i] & NB_F_EXISTS;
  // NewRenderable (renderable.go:22-40): float32 positions; zero for a non-existent body.
  xyz[3 * (size_t)i] = e ? (float)x[i] : 0.f;
  xyz[3 * (size_t)i + 1] = e ? (float)y[i] : 0.f;
  xyz[3 * (size_t)i + 2] = e ? (float)z[i] : 0.f;
}

hipError_t launch_render(const double *x, const double *y, const double *z, const uint8_t *flags,
                         int n, float *xyz, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_render, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y, z,
                     flags, n, xyz);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Collection lifecycle on the device: compaction (BodyCollection.Cycle) and removals
// (ComputationRunner.processDeletes).
// ---------------------------------------------------------------------------------------------
__global__ void k_keep_flags(const uint8_t *flags, int n, int32_t *keep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keep[i] = (flags[i] & NB_F_EXISTS) ? 1 : 0;
}

__global__ void k_compact(CompactArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n || !(a.flags_src[i] & NB_F_EXISTS)) return;
  const int p = a.pos[i];
#pragma unroll
  for (int k = 0; k < 10; k++) a.dst[k][p] = a.src[k][i];
  a.beh_dst[p] = a.beh_src[i];
  a.flags_dst[p] = a.flags_src[i];
  if (a.fstep_dst) a.fstep_dst[p] = a.fstep_src[i];
  if (a.fst_dst) a.fst_dst[p] = a.fst_src[i];
}

hipError_t launch_compact(const CompactArgs &a, hipStream_t st) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_compact, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_keep_flags(const uint8_t *flags, int n, int32_t *keep, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_keep_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, flags, n,
                     keep);
  return hipGetLastError();
}

// processDeletes (computation-runner.go:176-216) without its sequential walk: the walk sets
// shouldRemove at every index that is a multiple of step and clears it at the first existing,
// unpinned body at or after that point.  So the removed bodies are, for each stride point p, the
// first such body in [p, p + step) -- or, if there is none, the one the next stride point finds
// (the flag is still set when the walk gets there).  One thread per stride point scans its own
// interval; the last one scans to the end.  all: the -1 request, every existing body.
__global__ void k_remove_walk(uint8_t *flags, double *m, int n, int step, int npoints, int all,
                              unsigned long long *removed) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npoints) return;
  int lo, hi;
  if (all) {
    lo = k;
    hi = k + 1;
  } else {
    lo = k * step;
    hi = (k == npoints - 1) ? n : min(n, lo + step);
  }
  for (int i = lo; i < hi; i++) {
    const uint8_t f = flags[i];
    if ((f & NB_F_EXISTS) && (all || !(f & NB_F_PINNED))) {
      flags[i] = (uint8_t)(f & ~NB_F_EXISTS);  // SetNotExists (body.go:93-96)
      m[i] = 0.0;
      atomicAdd(&removed[0], 1ull);
      if (!(f & NB_F_FRAGMENTING)) atomicAdd(&removed[1], 1ull);
      break;
    }
  }
}

hipError_t launch_remove_walk(uint8_t *flags, double *m, int n, int step, int npoints, int all,
                              unsigned long long *removed, hipStream_t st) {
  if (npoints <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_remove_walk, dim3((unsigned)((npoints + 255) / 256)), dim3(256), 0, st,
                     flags, m, n, step, npoints, all, removed);
  return hipGetLastError();
}

// Largest radius over active bodies, as the bits of a non-negative double (their unsigned order
// is the numeric order), for the exact value the host computes when it loads the bodies.
__global__ __launch_bounds__(256) void k_rmax_active(const double *r, const uint8_t *flags, int n,
                                                     unsigned long long *out) {
  __shared__ unsigned long long sh[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (i < n && active_f(flags[i]) && r[i] > 0.0) v = (unsigned long long)__double_as_longlong(r[i]);
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; k++) v = sh[k] > v ? sh[k] : v;
    if (v) atomicMax(out, v);
  }
}

hipError_t launch_rmax_active(const double *r, const uint8_t *flags, int n,
                              unsigned long long *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rmax_active, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, r, flags,
                     n, out);
  return hipGetLastError();
}

}  // namespace nb
