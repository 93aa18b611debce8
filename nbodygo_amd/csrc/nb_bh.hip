// nb_bh.hip -- the reference's Barnes-Hut octree (cmd/body/octree.go) on the device, with Build
// wired into BuildTree (the reference only allocates an empty tree, body_collection.go:298-300,
// so its default mode computes no gravity).  FP64 throughout, with the reference's arithmetic
// for every interaction.
//
// Build, every force phase:
//   box     min/max over existing, non-fragmenting bodies; the root's centre and half exactly as
//           Build computes them (octree.go:136-194)
//   keys    each body descends from the root with the reference's centre arithmetic (octantOf,
//           childCenter: octree.go:55-88), one octal digit per level, max_depth levels
//   sort    stable radix sort of the keys: every node's bodies are one contiguous run, and a
//           bucket's bodies (the depth limit) stay in array order
//   levels  one launch per level splits each node's run into its eight octant runs by binary
//           search on that level's digit.  Runs of two or more bodies become nodes, single
//           bodies leaves -- the tree that leaf-capacity-1 insertion builds (insert, subdivide:
//           octree.go:198-231): same topology, node centres, halves and leaf depths.
//   mass    bottom-up per level.  A leaf's mass and centre of mass are updateMass's from zero,
//           (0 + x*m) / (0 + m) (octree.go:91-97); a node sums its children in octant order.
//           (The reference folds bodies into each node one at a time in array order, so a
//           node's centre of mass can differ from it in the last bits; leaves are identical.)
//           A bucket folds its bodies in array order as updateMass does.
// Links, after the masses: each live child slot gets the slot to visit when the walk opens it
// (its first live child) and the one to visit when it is accepted or done (its next live
// sibling, else its parent's), so a depth-first walk needs no stack.
// Force (accumForce, octree.go:276-309): one lane per body, bodies in key order.  Each lane walks
// the tree on its own by those links: it accepts or opens a node by its own test
// (2*half/dist) < theta, and children come in octant order, so every body's sum has the
// reference's terms in the reference's order.  Suns sum directly over all bodies in array order
// (CalcForce, octree.go:252-267).
// Collisions: a second walk opens every node whose box lies within R_i + R_max of the body and
// applies the reference's test at the leaves (Collided, body.go:192-208); pairs go to the same
// canonical list as the direct path.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "nb_internal.hpp"

namespace nb {

constexpr int kBhLevels = 21;        // octal digits in a 64-bit key
constexpr int kBhStack = 160;        // collision walk: per-wave stack entries (bound 7 * 20 + 8)
constexpr int32_t kBucket = INT32_MIN;
constexpr int kSunChunk = 64;

struct BhCounters {
  int32_t M;                         // bodies in the tree
  int32_t depth_max;                 // deepest leaf
  int32_t level[kBhLevels + 2];      // first node of each level
  int32_t lcount[kBhLevels + 2];     // nodes at each level
  uint32_t nsun;
  uint32_t buckets;
  uint32_t stack_overflow;
  uint32_t node_overflow;
  int32_t root_first;                // the root's first live child slot (-1: none)
  unsigned long long interactions;   // accepted nodes + leaf and direct terms, all targets
};

// A child as the force walk needs it, eight per node: centre of mass and mass (a leaf's from
// updateMass from zero), the child's code (-1 empty, <= -2 the leaf of sorted body -2 - code,
// else a node), the walk's links: `first` (slot of its first live child, for a node that is
// opened) and `skip` (slot visited next once this subtree is done; -1 ends the walk), and `dep`:
// a node's depth (its half is the root's halved dep times, exactly as the build halves it), -1
// for a bucket, 0 for a leaf or an empty slot.  A slot is live when its code is not -1 and its
// mass is > 0.  48 bytes: three 16-byte loads per visit instead of four, the half coming from
// a per-depth table.
struct alignas(16) BhSlot {
  double x, y, z, m;
  int32_t code, first, skip, dep;
};
static_assert(sizeof(BhSlot) == 48, "three 16-byte words");

// 2 half of each depth (the root's half, halved once per level: exact, as in the build), for
// d = 0..kBhLevels, into a workgroup's LDS table.
__device__ __forceinline__ void bh_half2_table(const double4 *nbox, double *h2s) {
  if (threadIdx.x <= kBhLevels) {
    double h = nbox[0].w;
    for (int d = 0; d < (int)threadIdx.x; d++) h = h / 2;
    h2s[threadIdx.x] = 2 * h;
  }
  __syncthreads();
}

struct Bh {
  nb_bh_params p{};
  int64_t cap = 0, node_cap = 0;
  uint64_t *keys_in = nullptr, *keys = nullptr;
  int32_t *perm_in = nullptr, *perm = nullptr;
  double *sx = nullptr, *sy = nullptr, *sz = nullptr, *sm = nullptr;
  double *lcx = nullptr, *lcy = nullptr, *lcz = nullptr;
  int32_t *depth = nullptr, *suns = nullptr;
  double4 *nbox = nullptr, *nmc = nullptr;
  int32_t *nchild = nullptr;
  BhSlot *slots = nullptr;
  int32_t *pslot = nullptr;  // each node's slot in its parent
  int2 *nrange = nullptr;
  double *bpart = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_bytes = 0;
  BhCounters *cnt = nullptr;
  int64_t builds = 0;
  int cus = 256;  // compute units of the device
};

template <class T>
static hipError_t balloc(T **p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  return hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
}

__device__ __forceinline__ bool bh_active(uint8_t f) {
  return (f & NB_F_EXISTS) && !(f & NB_F_FRAGMENTING);
}

__device__ __forceinline__ void bh_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---- box ---------------------------------------------------------------------------------------
// Partial min/max (x, y, z) and count over existing, non-fragmenting bodies: 7 values per block.
__global__ __launch_bounds__(256) void k_bh_box_part(const double *x, const double *y,
                                                     const double *z, const uint8_t *flags, int n,
                                                     double *part) {
  __shared__ double sh[7][256];
  double v[7] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (!bh_active(flags[i])) continue;
    const double px = x[i], py = y[i], pz = z[i];
    v[0] = px < v[0] ? px : v[0];
    v[1] = py < v[1] ? py : v[1];
    v[2] = pz < v[2] ? pz : v[2];
    v[3] = px > v[3] ? px : v[3];
    v[4] = py > v[4] ? py : v[4];
    v[5] = pz > v[5] ? pz : v[5];
    v[6] += 1.0;
  }
  for (int q = 0; q < 7; q++) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const int o = threadIdx.x + s;
      for (int q = 0; q < 3; q++) {
        const double a = sh[q][threadIdx.x], b = sh[q][o];
        sh[q][threadIdx.x] = b < a ? b : a;
      }
      for (int q = 3; q < 6; q++) {
        const double a = sh[q][threadIdx.x], b = sh[q][o];
        sh[q][threadIdx.x] = b > a ? b : a;
      }
      sh[6][threadIdx.x] += sh[6][o];
    }
    __syncthreads();
  }
  if (threadIdx.x < 7) part[blockIdx.x * 8 + threadIdx.x] = sh[threadIdx.x][0];
}

// The root cell (Build: cx = (minX + maxX) / 2, half = max extent / 2 * 1.05) and the level
// bookkeeping.  One block of 256: the partials are folded per thread, then in LDS (min, max and
// sums of whole numbers are exact in any order).
__global__ __launch_bounds__(256) void k_bh_root(const double *part, int nparts,
                                                 BhCounters *cnt, double4 *nbox, int2 *nrange,
                                                 int32_t *nchild) {
  __shared__ double sh[7][256];
  double v[7] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX, 0.0};
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
    const double *p = part + b * 8;
    for (int q = 0; q < 3; q++) v[q] = p[q] < v[q] ? p[q] : v[q];
    for (int q = 3; q < 6; q++) v[q] = p[q] > v[q] ? p[q] : v[q];
    v[6] += p[6];
  }
  for (int q = 0; q < 7; q++) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      const int o = threadIdx.x + st;
      for (int q = 0; q < 3; q++) {
        const double a = sh[q][threadIdx.x], b = sh[q][o];
        sh[q][threadIdx.x] = b < a ? b : a;
      }
      for (int q = 3; q < 6; q++) {
        const double a = sh[q][threadIdx.x], b = sh[q][o];
        sh[q][threadIdx.x] = b > a ? b : a;
      }
      sh[6][threadIdx.x] += sh[6][o];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  for (int q = 0; q < 7; q++) v[q] = sh[q][0];
  const int M = (int)v[6];
  cnt->M = M;
  cnt->depth_max = 0;
  for (int d = 0; d < kBhLevels + 2; d++) {
    cnt->level[d] = 0;
    cnt->lcount[d] = 0;
  }
  cnt->level[1] = 1;
  cnt->lcount[0] = M >= 2 ? 1 : 0;
  cnt->nsun = 0;
  cnt->buckets = 0;
  cnt->stack_overflow = 0;
  cnt->node_overflow = 0;
  cnt->interactions = 0;
  if (M == 0) return;
  const double ex = v[3] - v[0], ey = v[4] - v[1], ez = v[5] - v[2];
  const double mx = ey > ez ? ey : ez;  // math.Max on finite values
  const double half = (ex > mx ? ex : mx) / 2 * 1.05;
  nbox[0] = make_double4((v[0] + v[3]) / 2, (v[1] + v[4]) / 2, (v[2] + v[5]) / 2, half);
  nrange[0] = make_int2(0, M);
  for (int c = 0; c < 8; c++) nchild[c] = -1;
}

// Octant path of each body from the root, octantOf / childCenter op by op; D digits,
// left-aligned in 63 bits.  Bodies outside the tree get the largest key.
__global__ __launch_bounds__(256) void k_bh_keys(const double *x, const double *y,
                                                 const double *z, const uint8_t *flags, int n,
                                                 const double4 *nbox, int D, uint64_t *keys,
                                                 int32_t *perm, int32_t *depth) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  perm[i] = i;
  depth[i] = -1;
  if (!bh_active(flags[i])) {
    keys[i] = ~0ull;
    return;
  }
  const double4 r = nbox[0];
  double cx = r.x, cy = r.y, cz = r.z, half = r.w;
  const double px = x[i], py = y[i], pz = z[i];
  uint64_t key = 0;
  for (int l = 0; l < D; l++) {
    const int o = (px >= cx ? 1 : 0) | (py >= cy ? 2 : 0) | (pz >= cz ? 4 : 0);
    key = (key << 3) | (uint64_t)o;
    const double q = half / 2;
    cx = (o & 1) ? cx + q : cx - q;
    cy = (o & 2) ? cy + q : cy - q;
    cz = (o & 4) ? cz + q : cz - q;
    half = half / 2;
  }
  keys[i] = key << (3 * (kBhLevels - D));
}

// Sorted copies of position and mass, and each body's leaf centre of mass (updateMass from
// zero: (0*0 + x*m) / (0 + m)).
__global__ __launch_bounds__(256) void k_bh_gather(const int32_t *perm, const double *x,
                                                   const double *y, const double *z,
                                                   const double *m, const BhCounters *cnt,
                                                   double *sx, double *sy, double *sz,
                                                   double *sm, double *lcx, double *lcy,
                                                   double *lcz, int32_t *depth) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= cnt->M) return;
  const int i = perm[s];
  if (cnt->M == 1) depth[i] = 0;  // the root is the lone body's leaf
  const double px = x[i], py = y[i], pz = z[i], mm = m[i];
  const double zero = 0.0;
  const double tot = zero + mm;
  sx[s] = px;
  sy[s] = py;
  sz[s] = pz;
  sm[s] = mm;
  lcx[s] = (zero * zero + px * mm) / tot;
  lcy[s] = (zero * zero + py * mm) / tot;
  lcz[s] = (zero * zero + pz * mm) / tot;
}

// One level of the tree: each node of level d splits its run into octant runs.
struct BhLevelArgs {
  const uint64_t *keys;
  const int32_t *perm;
  BhCounters *cnt;
  double4 *nbox;
  int2 *nrange;
  int32_t *nchild;
  int32_t *depth;
  int64_t node_cap;
  int d, D;
};

// Eight lanes per node: lane c (1..7) finds by binary search the first body of the run whose
// digit d is >= c, the group's boundaries go round by shuffles, and lane c makes child c.
__global__ __launch_bounds__(256) void k_bh_level(BhLevelArgs a) {
  BhCounters *cnt = a.cnt;
  const int d = a.d;
  const int begin = cnt->level[d], end = begin + cnt->lcount[d];
  if (blockIdx.x == 0 && threadIdx.x == 0) cnt->level[d + 1] = end;
  const int shift = 60 - 3 * d;
  const int last = (int)min((int64_t)end, a.node_cap);  // nodes past the capacity were not made
  const int c = threadIdx.x & 7;
  const int groups = gridDim.x * (blockDim.x >> 3);
  // every lane runs the same number of iterations, so the shuffles see full groups
  const int first = begin + (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 3);
  for (int k0 = begin; k0 < last; k0 += groups) {
    const int k = first + (k0 - begin);
    const bool live = k < last;
    int2 rg = make_int2(0, 0);
    int lb = 0;
    if (live) {
      rg = a.nrange[k];
      if (c == 0) {
        lb = rg.x;
      } else {  // first body whose digit is >= c
        int lo = rg.x, hi = rg.y;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if ((int)((a.keys[mid] >> shift) & 7) < c)
            lo = mid + 1;
          else
            hi = mid;
        }
        lb = lo;
      }
    }
    int ub = __shfl_down(lb, 1, 8);
    if (c == 7) ub = rg.y;
    if (!live) continue;
    const double4 bx = a.nbox[k];
    const double q = bx.w / 2;
    const int cn = ub - lb;
    int32_t code = -1;
    if (cn == 1) {
      code = -2 - lb;
      a.depth[a.perm[lb]] = d + 1;
      atomicMax(&cnt->depth_max, d + 1);
    } else if (cn >= 2) {
      const int id = end + atomicAdd(&cnt->lcount[d + 1], 1);
      if (id >= a.node_cap) {
        atomicOr(&cnt->node_overflow, 1u);
      } else {
        a.nrange[id] = make_int2(lb, ub);
        a.nbox[id] = make_double4((c & 1) ? bx.x + q : bx.x - q, (c & 2) ? bx.y + q : bx.y - q,
                                  (c & 4) ? bx.z + q : bx.z - q, bx.w / 2);
        int32_t *ch = a.nchild + (size_t)id * 8;
        if (d + 1 == a.D) {  // depth limit: a bucket of bodies in array order
          ch[0] = kBucket;
          ch[1] = lb;
          ch[2] = cn;
          for (int q2 = 3; q2 < 8; q2++) ch[q2] = -1;
          for (int s = lb; s < ub; s++) a.depth[a.perm[s]] = d + 1;
          atomicAdd(&cnt->buckets, 1u);
          atomicMax(&cnt->depth_max, d + 1);
        } else {
          for (int q2 = 0; q2 < 8; q2++) ch[q2] = -1;
        }
        code = id;
      }
    }
    a.nchild[(size_t)k * 8 + c] = code;
  }
}

// Mass and centre of mass of the nodes of level d, from their children (or, for a bucket, its
// bodies folded in array order by updateMass).
struct BhMassArgs {
  const BhCounters *cnt;
  const int32_t *nchild;
  const int2 *nrange;
  const double4 *nbox;
  double4 *nmc;
  BhSlot *slots;
  const double *sx, *sy, *sz, *sm, *lcx, *lcy, *lcz;
  int64_t node_cap;
  int d;
};

__global__ __launch_bounds__(256) void k_bh_mass(BhMassArgs a) {
  const int begin = a.cnt->level[a.d];
  const int end = (int)min((int64_t)begin + a.cnt->lcount[a.d], a.node_cap);
  for (int k = begin + blockIdx.x * blockDim.x + threadIdx.x; k < end;
       k += gridDim.x * blockDim.x) {
    const int32_t *ch = a.nchild + (size_t)k * 8;
    if (ch[0] == kBucket) {
      double mt = 0, cx = 0, cy = 0, cz = 0;
      for (int s = ch[1]; s < ch[1] + ch[2]; s++) {
        const double mm = a.sm[s];
        const double nt = mt + mm;
        cx = (cx * mt + a.sx[s] * mm) / nt;
        cy = (cy * mt + a.sy[s] * mm) / nt;
        cz = (cz * mt + a.sz[s] * mm) / nt;
        mt = nt;
      }
      a.nmc[k] = make_double4(cx, cy, cz, mt);
      continue;
    }
    double mt = 0, wx = 0, wy = 0, wz = 0;
    BhSlot *sl = a.slots + (size_t)k * 8;
    for (int c = 0; c < 8; c++) {
      const int32_t code = ch[c];
      BhSlot o;
      o.code = code;
      o.first = o.skip = -1;
      o.dep = 0;
      if (code == -1) {
        o.x = o.y = o.z = o.m = 0.0;
        sl[c] = o;
        continue;
      }
      if (code <= -2) {
        const int s = -2 - code;
        o.m = 0.0 + a.sm[s];
        o.x = a.lcx[s];
        o.y = a.lcy[s];
        o.z = a.lcz[s];
      } else {
        const double4 v = a.nmc[code];
        o.m = v.w;
        o.x = v.x;
        o.y = v.y;
        o.z = v.z;
        o.dep = a.nchild[(size_t)code * 8] == kBucket ? -1 : a.d + 1;
      }
      sl[c] = o;
      mt += o.m;
      wx += o.x * o.m;
      wy += o.y * o.m;
      wz += o.z * o.m;
    }
    a.nmc[k] = make_double4(wx / mt, wy / mt, wz / mt, mt);
  }
}

// ---- links -------------------------------------------------------------------------------------
__device__ __forceinline__ int bh_nodes(const BhCounters *cnt, int D, int64_t node_cap) {
  return (int)min((int64_t)cnt->level[D] + cnt->lcount[D], node_cap);
}

__device__ __forceinline__ bool bh_live(const BhSlot &sl) { return sl.code != -1 && sl.m > 0; }

// Each node's slot in its parent (the root has none).
__global__ __launch_bounds__(256) void k_bh_parent(const BhCounters *cnt, const int32_t *nchild,
                                                   int D, int64_t node_cap, int32_t *pslot) {
  const int nodes = bh_nodes(cnt, D, node_cap);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nodes; k += gridDim.x * blockDim.x) {
    const int32_t *ch = nchild + (size_t)k * 8;
    if (ch[0] == kBucket) continue;
    for (int c = 0; c < 8; c++)
      if (ch[c] >= 0) pslot[ch[c]] = k * 8 + c;
  }
}

// first and skip of every live slot; the root's first live child goes to the counters.
__global__ __launch_bounds__(256) void k_bh_links(BhCounters *cnt, const int32_t *nchild,
                                                  const int32_t *pslot, int D, int64_t node_cap,
                                                  BhSlot *slots) {
  const int nodes = bh_nodes(cnt, D, node_cap);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nodes; k += gridDim.x * blockDim.x) {
    if (nchild[(size_t)k * 8] == kBucket) continue;
    BhSlot *sl = slots + (size_t)k * 8;
    if (k == 0) {
      int f = -1;
      for (int c = 0; c < 8 && f < 0; c++)
        if (bh_live(sl[c])) f = c;
      cnt->root_first = f;
    }
    for (int c = 0; c < 8; c++) {
      if (!bh_live(sl[c])) continue;
      int first = -1;
      const int32_t code = sl[c].code;
      if (code >= 0 && sl[c].dep >= 0) {
        const BhSlot *cs = slots + (size_t)code * 8;
        for (int c2 = 0; c2 < 8 && first < 0; c2++)
          if (bh_live(cs[c2])) first = code * 8 + c2;
      }
      // next live sibling of this slot, else of the nearest ancestor that has one
      int skip = -1, node = k, cc = c;
      for (;;) {
        const BhSlot *ps = slots + (size_t)node * 8;
        for (int c2 = cc + 1; c2 < 8 && skip < 0; c2++)
          if (bh_live(ps[c2])) skip = node * 8 + c2;
        if (skip >= 0 || node == 0) break;
        const int up = pslot[node];
        node = up >> 3;
        cc = up & 7;
      }
      sl[c].first = first;
      sl[c].skip = skip;
    }
  }
}

// ---- force -------------------------------------------------------------------------------------
struct BhForceArgs {
  const BhCounters *cnt;
  BhCounters *cnt_w;
  const int32_t *perm;
  const uint64_t *keys;  // sorted keys (fast mode's group cut)
  const uint8_t *flags;
  const double *sx, *sy, *sz, *sm;
  const double4 *nbox, *nmc;
  const int32_t *nchild;
  const BhSlot *slots;
  double *acc;
  int npad;
  int gsz;  // bodies per wave (lanes gsz..63 idle)
  double G, theta, soft2;
  double theta2;  // RN(theta^2) when it is a normal number, else NaN (exact test always)
  int64_t node_cap;
};

// The reference's opening test RN(2 half / RN(sqrt(distSq))) < theta (octree.go:299-306),
// decided from squares where it is clear-cut.  Let a = RN((2h)^2) and b = RN(RN(theta^2)
// distSq), with RN(theta^2) normal (the host passes NaN otherwise) and b in [2^-1000, 2^1000].
// Then each rounding (a when normal, theta^2, b, the margin product) and the square root's is
// within 2^-53 relative, so a < b (1 - 2^-40) gives 2h / dist < theta (1 - 2^-43), below
// theta's predecessor, and a > b (1 + 2^-40) gives 2h / dist > theta (1 + 2^-44), above its
// successor: the rounded quotient lands on the same side of theta as the exact one.  a needs
// no range check: below the normal range its absolute error (2^-1075) is 2^-74 of b, and an
// infinite a exceeds any such b.  Anything else (within 2^-40 of the boundary, b out of range,
// NaN) is left undecided for the reference's own expression.
__device__ __forceinline__ void bh_quick_test(double h2, double distSq, double theta2,
                                              bool &yes, bool &no) {
  const double a = h2 * h2, b = theta2 * distSq;
  const bool normal = b >= 0x1p-1000 && b <= 0x1p1000;
  yes = normal && a < b * (1.0 - 0x1p-40);
  no = normal && a > b * (1.0 + 0x1p-40);
}

// One interaction, op for op as accumForce / directForce (octree.go:292-298, 314-325).
__device__ __forceinline__ void bh_term(double dx, double dy, double dz, double distSq,
                                        double dist, double Gm, double pm, double &fx, double &fy,
                                        double &fz) {
  const double force = Gm * pm / distSq;
  fx += force * dx / dist;
  fy += force * dy / dist;
  fz += force * dz / dist;
}

// The same interaction with the three divisions by dist sharing one reciprocal.  A correctly
// rounded FP64 division n / d compiles to: d' = v_div_scale(d, d, n); y = v_rcp(d'); two Newton
// steps y = fma(y, fma(-d', y, 1), y); n' = v_div_scale(n, d, n); q = n' y; r = fma(-d', q, n');
// q = v_div_fmas(r, y, q); v_div_fixup(q, d, n).  With d in [2^-400, 2^400] and |n| in
// [2^-400, 2^300] both scalings return their operand unchanged and clear VCC (no zero operand, no
// denormal d, 1/d or n/d, exponent gap below 768, n not tiny), v_div_fmas is then a plain fma
// and v_div_fixup returns q with the sign of n / d, which q already has.  So within that range
// the steps that depend on d alone are computed once and each quotient takes three
// instructions, bit-identical to the division; any operand outside it takes the division.
// (The conditions are combined in int, as one lane mask rather than a chain of branches.)
__device__ __forceinline__ int bh_div_range(double n) {
  return (int)(fabs(n) >= 0x1p-400) & (int)(fabs(n) <= 0x1p300);
}

__device__ __forceinline__ void bh_term_shared(double dx, double dy, double dz, double distSq,
                                               double dist, double Gm, double pm, double &fx,
                                               double &fy, double &fz) {
  // the reciprocal does not wait for the numerators: it runs beside the first division
  double y = __builtin_amdgcn_rcp(dist);
  y = __builtin_fma(y, __builtin_fma(-dist, y, 1.0), y);
  y = __builtin_fma(y, __builtin_fma(-dist, y, 1.0), y);
  const double force = Gm * pm / distSq;
  const double nx = force * dx, ny = force * dy, nz = force * dz;
  const int shared = (int)(dist >= 0x1p-400) & (int)(dist <= 0x1p400) & bh_div_range(nx) &
                     bh_div_range(ny) & bh_div_range(nz);
  if (shared) {
    auto quot = [&](double n) {
      const double q = n * y;
      return __builtin_fma(__builtin_fma(-dist, q, n), y, q);
    };
    fx += quot(nx);
    fy += quot(ny);
    fz += quot(nz);
  } else {
    fx += nx / dist;
    fy += ny / dist;
    fz += nz / dist;
  }
}

// The exact walk: one lane per body, bodies in key order.  Each step decides its slot (a leaf
// is taken unless it is the body itself; a node by the opening test from squares, with the
// reference's expression behind a wave-uniform branch only near the boundary), requests the
// next slot and only then computes the taken slot's term, so the term's square root and
// divisions run while the load is in flight.  Terms and their order are the reference's.
// SHARED selects the term with one reciprocal for its three divisions: fewer instructions, a
// longer step (measured on one box: 10.63 against 10.94 ms at N = 262,144, four waves per SIMD;
// 3.31 against 3.14 ms at N = 65,536, one wave per SIMD).
template <bool SHARED>
__global__ __launch_bounds__(256) void k_bh_force(BhForceArgs a) {
  const int M = a.cnt->M;
  if (M < 2) return;  // a lone body feels nothing; forces were zeroed
  __shared__ double h2s[kBhLevels + 1];
  bh_half2_table(a.nbox, h2s);
  const int lane = threadIdx.x & 63;
  const int s = (blockIdx.x * 4 + (int)(threadIdx.x >> 6)) * a.gsz + lane;
  bool act = lane < a.gsz && s < M;
  int i = -1;
  if (act) {
    i = a.perm[s];
    act = !(a.flags[i] & NB_F_SUN);  // suns sum directly
  }
  const unsigned long long nact = __builtin_amdgcn_ballot_w64(act);
  double fx = 0, fy = 0, fz = 0;
  unsigned long long nint = 0;
  if (act) {
    const double xi = a.sx[s], yi = a.sy[s], zi = a.sz[s], Gm = a.G * a.sm[s];
    const double soft2 = a.soft2, theta = a.theta, theta2 = a.theta2;
    const double4 root = a.nmc[0];
    int q = -1;
    if (root.w != 0.0) {  // accumForce: n.totalMass == 0 returns
      const double dx = root.x - xi, dy = root.y - yi, dz = root.z - zi;
      const double distSq = dx * dx + dy * dy + dz * dz + soft2;
      const double dist = __builtin_sqrt(distSq);
      if ((2 * a.nbox[0].w / dist) < theta) {
        bh_term(dx, dy, dz, distSq, dist, Gm, root.w, fx, fy, fz);
        nint++;
      } else {
        const int f = a.cnt->root_first;
        q = f < 0 ? -1 : f;
      }
    }
    int4 w0, w1, w2;
    auto fetch = [&](int slot) {
      const int4 *w = reinterpret_cast<const int4 *>(a.slots + slot);
      w0 = w[0];
      w1 = w[1];
      w2 = w[2];
    };
    // Every step requests a slot, slot 0 (node 0's first slot, always allocated) once the walk is
    // done: an unconditional load lands straight in the registers the next step reads, where a
    // conditional one is copied there and waited for before the term.
    fetch(max(q, 0));
    uint32_t nterm = 0;  // fewer than 2^32: one per visited slot
    while (q >= 0) {
      BhSlot sl;
      __builtin_memcpy(&sl, &w0, 16);
      __builtin_memcpy(reinterpret_cast<char *>(&sl) + 16, &w1, 16);
      __builtin_memcpy(reinterpret_cast<char *>(&sl) + 32, &w2, 16);
      const bool leaf = sl.code <= -2;
      if (!leaf && sl.dep < 0) {  // depth limit: the bucket's bodies one by one (directForce)
        const int32_t *ch = a.nchild + (size_t)sl.code * 8;
        const int b0 = ch[1], bn = ch[2];
        for (int s2 = b0; s2 < b0 + bn; s2++) {
          if (s2 == s) continue;
          const double ex = a.sx[s2] - xi, ey = a.sy[s2] - yi, ez = a.sz[s2] - zi;
          const double d2 = ex * ex + ey * ey + ez * ez + soft2;
          bh_term(ex, ey, ez, d2, __builtin_sqrt(d2), Gm, a.sm[s2], fx, fy, fz);
          nint++;
        }
        q = sl.skip;
        fetch(max(q, 0));
        continue;
      }
      const double dx = sl.x - xi, dy = sl.y - yi, dz = sl.z - zi;
      const double distSq = dx * dx + dy * dy + dz * dz + soft2;
      const double h2 = h2s[leaf ? 0 : sl.dep];
      bool yes, no;
      bh_quick_test(h2, distSq, theta2, yes, no);
      const bool other = -2 - sl.code != s;
      bool take = (leaf & other) | (!leaf & yes);
      const bool und = !leaf & !yes & !no;
      if (__builtin_amdgcn_ballot_w64(und)) {
        __asm__ volatile("");  // a real branch: this division is rare
        if (und) take = (h2 / __builtin_sqrt(distSq)) < theta;
      }
      q = (leaf | take) ? sl.skip : sl.first;
      fetch(max(q, 0));
      if (take) {
        if constexpr (SHARED)
          bh_term_shared(dx, dy, dz, distSq, __builtin_sqrt(distSq), Gm, sl.m, fx, fy, fz);
        else
          bh_term(dx, dy, dz, distSq, __builtin_sqrt(distSq), Gm, sl.m, fx, fy, fz);
        nterm++;
      }
    }
    nint += nterm;
    a.acc[i] = fx;
    a.acc[a.npad + i] = fy;
    a.acc[2 * a.npad + i] = fz;
  }
  if (nact == 0) return;
  for (int o = 32; o > 0; o >>= 1) nint += __shfl_xor(nint, o);
  if (lane == 0) atomicAdd(&a.cnt_w->interactions, nint);
}

// ---- fast mode: one walk per group of bodies, FP32 terms ------------------------------------------
// The 64 bodies of a wave (consecutive in key order) are cut into groups at tree-cell boundaries:
// a run whose smallest enclosing cell holds kBhGroupSpan or more bodies on either side of the run
// is split between that cell's children, and again, until every run lies in a cell of fewer than
// 64 + 2 kBhGroupSpan bodies (or is one body, or one bucket).  Each group shares one traversal.
// Without the cut, a run that straddles two large cells takes a box spanning both and opens
// everything inside it to the leaves (on a Plummer sphere one such wave met 261,464 of the
// 262,144 bodies and set the kernel's duration).  A node is accepted for the whole group when 2 half < theta * sqrt(d_box^2 + soft^2), d_box being
// its centre of mass's distance to the group's bounding box: every body of the group is then at
// least that far away, so each would accept it by the reference's own test (octree.go:299-309);
// otherwise the node is opened.  The walk is breadth-parallel: each step pops up to 8 nodes from
// a per-wave LDS stack and the 64 lanes examine their 64 child slots at once; accepted nodes and
// leaves go to an LDS list of terms, opened nodes back to the stack.  Every 64 listed terms are
// evaluated by all lanes against their own bodies like the direct FP32 kernel (sources broadcast
// from LDS, coordinates relative to the group's centre, r^-3 from v_rsq_f32), the list's partial
// sums are added into FP64 per lane, and the self term is skipped by body index.
constexpr int kBhGroupSpan = 512;   // see the group cut above
// Per-wave stack entries.  The stack stays sorted by depth (each step pops the deepest entries
// and pushes their children on top), so entries of depth d can only be pushed by a step that
// popped every deeper-or-equal entry: at most 64 entries per depth, 21 depths below the root.
constexpr int kBhFastStack = 64 * kBhLevels;
constexpr int kBhFastList = 128;    // per-wave term list (< 64 kept + up to 64 added per step)

struct BhFastLds {
  int32_t stack[kBhFastStack];
  float4 term[kBhFastList];
  int32_t tidx[kBhFastList];  // body (sorted index) of a leaf term, -1 for a node
  int32_t top, len;
};

__device__ __forceinline__ void fast_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The first `cnt` listed terms against the lane's body; the rest of the list moves to the front.
__device__ __forceinline__ void fast_flush(BhFastLds &L, int cnt, float tx, float ty, float tz,
                                           float soft2, int s, bool act, double &ax, double &ay,
                                           double &az, unsigned long long &nint) {
  const int lane = threadIdx.x & 63;
  float px = 0.f, py = 0.f, pz = 0.f;
  const float sf = to_vgpr(soft2);
  for (int j = 0; j < cnt; j++) {
    const float4 e = L.term[j];
    const int id = L.tidx[j];
    const float dx = e.x - tx, dy = e.y - ty, dz = e.z - tz;
    float r2 = __builtin_fmaf(dx, dx, sf);
    r2 = __builtin_fmaf(dy, dy, r2);
    r2 = __builtin_fmaf(dz, dz, r2);
    const float inv = __builtin_amdgcn_rsqf(r2);
    float w = e.w * inv;
    w = w * (inv * inv);
    w = id == s ? 0.f : w;
    px = __builtin_fmaf(dx, w, px);
    py = __builtin_fmaf(dy, w, py);
    pz = __builtin_fmaf(dz, w, pz);
  }
  if (act) {
    ax += px;
    ay += py;
    az += pz;
    nint += (unsigned long long)cnt;
  }
  fast_sync();
  const int rest = L.len - cnt;
  float4 mv = make_float4(0.f, 0.f, 0.f, 0.f);
  int mi = -1;
  if (lane < rest) {
    mv = L.term[cnt + lane];
    mi = L.tidx[cnt + lane];
  }
  fast_sync();
  if (lane < rest) {
    L.term[lane] = mv;
    L.tidx[lane] = mi;
  }
  if (lane == 0) L.len = rest;
  fast_sync();
}

// Appends one term per lane that has one (wave-wide, in lane order).
__device__ __forceinline__ void fast_append(BhFastLds &L, bool has, float4 t, int id) {
  const unsigned long long bal = __builtin_amdgcn_ballot_w64(has);
  const int base = L.len;
  if (has) {
    const int pos = base + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0));
    L.term[pos] = t;
    L.tidx[pos] = id;
  }
  fast_sync();
  if ((threadIdx.x & 63) == 0) L.len = base + __builtin_popcountll(bal);
  fast_sync();
}

__global__ __launch_bounds__(256) void k_bh_force_fast(BhForceArgs a) {
  __shared__ BhFastLds lds[4];
  __shared__ double h2s[kBhLevels + 1];
  const int M = a.cnt->M;
  if (M < 2) return;
  bh_half2_table(a.nbox, h2s);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  BhFastLds &L = lds[wv];
  const int s = (blockIdx.x * 4 + wv) * 64 + lane;
  bool act = s < M;
  int i = -1;
  if (act) {
    i = a.perm[s];
    act = !(a.flags[i] & NB_F_SUN);  // suns sum directly
  }
  if (__builtin_amdgcn_ballot_w64(act) == 0) return;
  const double px = act ? a.sx[s] : 0, py = act ? a.sy[s] : 0, pz = act ? a.sz[s] : 0;
  // the group cut (wave-uniform): bit j of `starts` marks a group starting at lane j
  const int wbase = s - lane, nl = min(64, M - wbase);
  const uint64_t kl = lane < nl ? a.keys[s] : ~0ull;
  const uint64_t kprev = __shfl_up(kl, 1);
  unsigned long long starts = 1ull, fin = 0ull;
  for (;;) {
    const unsigned long long todo = starts & ~fin;
    if (todo == 0) break;
    const int ga = __builtin_ctzll(todo);
    const unsigned long long later = starts & ~((2ull << ga) - 1ull);
    const int gb = later ? __builtin_ctzll(later) : nl;
    const uint64_t ka = __shfl(kl, ga), kb = __shfl(kl, gb - 1);
    const uint64_t x = ka ^ kb;  // live keys use bits 0..62, so x < 2^63
    bool cut = false;
    if (gb - ga > 1 && x != 0) {
      // bits below the run's common prefix: whole octal digits down from the highest difference
      const int lb = 3 * ((63 - __builtin_clzll(x)) / 3 + 1);
      const uint64_t pre = ka >> lb;
      const int pa = wbase + ga - kBhGroupSpan, pb = wbase + gb - 1 + kBhGroupSpan;
      const bool big = (pa >= 0 && (a.keys[pa] >> lb) == pre) ||
                       (pb < M && (a.keys[pb] >> lb) == pre);
      if (big) {
        // cut where the next digit changes: the run's parts in the cell's children
        starts |= __builtin_amdgcn_ballot_w64(lane > ga && lane < gb &&
                                              ((kl ^ kprev) >> (lb - 3)) != 0);
        cut = true;
      }
    }
    if (!cut) fin |= 1ull << ga;
  }
  const double th2 = a.theta * a.theta, soft2 = a.soft2;
  double ax = 0, ay = 0, az = 0;
  unsigned long long nint = 0;
  bool overflow = false;
  const int node_lim = (int)min(a.node_cap, (int64_t)INT32_MAX / 8);
  for (unsigned long long segs = starts; segs && !overflow;) {
    const int ga = __builtin_ctzll(segs);
    segs &= segs - 1;
    const int gb = segs ? __builtin_ctzll(segs) : nl;
    const bool mine = act && lane >= ga && lane < gb;
    if (__builtin_amdgcn_ballot_w64(mine) == 0) continue;
    // the group's box (FP64)
    double glo[3] = {mine ? px : DBL_MAX, mine ? py : DBL_MAX, mine ? pz : DBL_MAX};
    double ghi[3] = {mine ? px : -DBL_MAX, mine ? py : -DBL_MAX, mine ? pz : -DBL_MAX};
    for (int o = 32; o > 0; o >>= 1)
      for (int d = 0; d < 3; d++) {
        glo[d] = fmin(glo[d], __shfl_xor(glo[d], o));
        ghi[d] = fmax(ghi[d], __shfl_xor(ghi[d], o));
      }
    const double cx = 0.5 * (glo[0] + ghi[0]), cy = 0.5 * (glo[1] + ghi[1]),
                 cz = 0.5 * (glo[2] + ghi[2]);
    const double hx = 0.5 * (ghi[0] - glo[0]), hy = 0.5 * (ghi[1] - glo[1]),
                 hz = 0.5 * (ghi[2] - glo[2]);
    const float tx = (float)(px - cx), ty = (float)(py - cy), tz = (float)(pz - cz);
    // a node accepted for the group: (2 half)^2 < theta^2 (d_box^2 + soft^2)
    auto accept = [&](double x, double y, double z, double half) {
      const double ex = fmax(fabs(x - cx) - hx, 0.0), ey = fmax(fabs(y - cy) - hy, 0.0),
                   ez = fmax(fabs(z - cz) - hz, 0.0);
      return 4.0 * half * half < th2 * (ex * ex + ey * ey + ez * ez + soft2);
    };
    if (lane == 0) {
      L.top = 0;
      L.len = 0;
    }
    fast_sync();
    const double4 root = a.nmc[0];
    if (root.w != 0.0) {
      if (accept(root.x, root.y, root.z, a.nbox[0].w)) {
        fast_append(L, lane == 0, make_float4((float)(root.x - cx), (float)(root.y - cy),
                                              (float)(root.z - cz), (float)root.w), -1);
      } else {
        if (lane == 0) {
          L.stack[0] = 0;
          L.top = 1;
        }
        fast_sync();
      }
    }
    for (;;) {
      const int top = L.top;
      if (top == 0) break;
      const int nn = min(8, top);
      const int k = lane >> 3, c = lane & 7;
      const int node = k < nn ? L.stack[top - nn + k] : -1;
      fast_sync();
      if (lane == 0) L.top = top - nn;
      fast_sync();
      bool term = false, open = false, bucket = false;
      float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
      int id = -1, code = -1;
      if (node >= 0 && node < node_lim) {
        const BhSlot sl = a.slots[(size_t)node * 8 + c];
        code = sl.code;
        if (code != -1 && sl.m > 0) {
          if (code <= -2) {  // a leaf: the body's own centre of mass (x*m)/m and mass 0 + m
            term = true;
            id = -2 - code;
          } else if (sl.dep < 0) {
            bucket = code < node_lim;
          } else if (accept(sl.x, sl.y, sl.z, 0.5 * h2s[sl.dep])) {
            term = true;
          } else {
            open = code < node_lim;
          }
          if (term)
            t = make_float4((float)(sl.x - cx), (float)(sl.y - cy), (float)(sl.z - cz),
                            (float)sl.m);
        }
      }
      fast_append(L, term, t, id);
      if (L.len >= 64) fast_flush(L, 64, tx, ty, tz, (float)soft2, s, mine, ax, ay, az, nint);
      // opened nodes onto the stack
      const unsigned long long bo = __builtin_amdgcn_ballot_w64(open);
      const int base = L.top, nopen = __builtin_popcountll(bo);
      if (base + nopen > kBhFastStack) {
        overflow = true;
        break;
      }
      if (open)
        L.stack[base + (int)__builtin_amdgcn_mbcnt_hi(
                           (uint32_t)(bo >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bo, 0))] =
            code;
      fast_sync();
      if (lane == 0) L.top = base + nopen;
      fast_sync();
      // buckets (the depth limit): their bodies one by one, 64 at a time
      unsigned long long bb = __builtin_amdgcn_ballot_w64(bucket);
      while (bb) {
        const int src = __builtin_ctzll(bb);
        bb &= bb - 1;
        const int bnode = __shfl(code, src);
        const int32_t *ch = a.nchild + (size_t)bnode * 8;
        const int b0 = ch[1], bn = ch[2];
        for (int q0 = 0; q0 < bn; q0 += 64) {
          const int q = b0 + q0 + lane;
          const bool h = q0 + lane < bn;
          float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
          if (h)
            e = make_float4((float)(a.sx[q] - cx), (float)(a.sy[q] - cy), (float)(a.sz[q] - cz),
                            (float)a.sm[q]);
          fast_append(L, h, e, q);
          if (L.len >= 64)
            fast_flush(L, 64, tx, ty, tz, (float)soft2, s, mine, ax, ay, az, nint);
        }
      }
    }
    if (L.len > 0) fast_flush(L, L.len, tx, ty, tz, (float)soft2, s, mine, ax, ay, az, nint);
  }
  if (overflow && lane == 0) atomicOr(&a.cnt_w->stack_overflow, 1u);
  if (act) {
    const double Gm = a.G * a.sm[s];
    a.acc[i] = Gm * ax;
    a.acc[a.npad + i] = Gm * ay;
    a.acc[2 * a.npad + i] = Gm * az;
  }
  for (int o = 32; o > 0; o >>= 1) nint += __shfl_xor(nint, o);
  if (lane == 0) atomicAdd(&a.cnt_w->interactions, nint);
}

// Suns: directForce from every other existing, non-fragmenting body in array order.
__global__ __launch_bounds__(256) void k_bh_sun_list(const uint8_t *flags, int n, BhCounters *cnt,
                                                     int32_t *suns) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && bh_active(flags[i]) && (flags[i] & NB_F_SUN)) suns[atomicAdd(&cnt->nsun, 1u)] = i;
}

struct BhSunArgs {
  BhCounters *cnt;
  const int32_t *suns;
  const double *x, *y, *z, *m;
  const uint8_t *flags;
  double *acc;
  int n, npad;
  double G, soft2;
};

__global__ __launch_bounds__(64) void k_bh_sun(BhSunArgs a) {
  __shared__ double t[3][kSunChunk];
  __shared__ int valid[kSunChunk];
  if (a.cnt->M < 2) return;
  const int lane = threadIdx.x;
  const int nsun = (int)a.cnt->nsun;
  for (int k = blockIdx.x; k < nsun; k += gridDim.x) {
    const int b = a.suns[k];
    const double xb = a.x[b], yb = a.y[b], zb = a.z[b], Gm = a.G * a.m[b];
    double sum = 0;  // lane c < 3 accumulates component c
    unsigned long long nint = 0;
    for (int j0 = 0; j0 < a.n; j0 += kSunChunk) {
      const int j = j0 + lane;
      const bool v = j < a.n && j != b && bh_active(a.flags[j]);
      double tx = 0, ty = 0, tz = 0;
      if (v) {
        const double dx = a.x[j] - xb;
        const double dy = a.y[j] - yb;
        const double dz = a.z[j] - zb;
        const double distSq = dx * dx + dy * dy + dz * dz + a.soft2;
        const double dist = __builtin_sqrt(distSq);
        const double force = Gm * a.m[j] / distSq;
        tx = force * dx / dist;
        ty = force * dy / dist;
        tz = force * dz / dist;
        nint++;
      }
      t[0][lane] = tx;
      t[1][lane] = ty;
      t[2][lane] = tz;
      valid[lane] = v ? 1 : 0;
      bh_wave_sync();
      if (lane < 3)
        for (int q = 0; q < kSunChunk; q++)
          if (valid[q]) sum += t[lane][q];
      bh_wave_sync();
    }
    if (lane < 3) a.acc[(size_t)lane * a.npad + b] = sum;
    for (int o = 32; o > 0; o >>= 1) nint += __shfl_xor(nint, o);
    if (lane == 0) atomicAdd(&a.cnt->interactions, nint);
  }
}

// ---- collision pairs ---------------------------------------------------------------------------
struct BhDetectArgs {
  const BhCounters *cnt;
  const int32_t *perm;
  const double *x, *y, *z, *r;
  const int32_t *beh;
  const double4 *nbox;
  const int32_t *nchild;
  uint64_t *pairs;
  DevCounters *ctr;
  int64_t pair_cap;
  double rmax;
  int events;
};

__device__ __forceinline__ void bh_emit(const BhDetectArgs &a, int i, int j) {
  const uint32_t q = atomicAdd(&a.ctr->pair_count, 1u);
  if ((int64_t)q < a.pair_cap)
    a.pairs[q] = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
  else
    atomicOr(&a.ctr->pair_overflow, 1u);
}

// Collided (body.go:192-208) from body i's side, pairs i < j, dispatched as Compute does.
__device__ __forceinline__ void bh_test_pair(const BhDetectArgs &a, int i, int32_t bi, double xi,
                                             double yi, double zi, double ri, int s2) {
  const int j = a.perm[s2];
  if (j <= i) return;
  const double dx = a.x[j] - xi;
  const double dy = a.y[j] - yi;
  const double dz = a.z[j] - zi;
  const double dist = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
  const double rj = a.r[j];
  if (dist <= ri + rj && nb_pair_event(bi, a.beh[j], ri, rj, dist, a.events)) bh_emit(a, i, j);
}

__global__ __launch_bounds__(256) void k_bh_detect(BhDetectArgs a) {
  __shared__ int32_t stk_code[4][kBhStack];
  __shared__ unsigned long long stk_mask[4][kBhStack];
  const int M = a.cnt->M;
  if (M < 2) return;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  int i = -1;
  bool act = false;
  int32_t bi = NB_CB_NONE;
  double xi = 0, yi = 0, zi = 0, ri = 0, lim = 0;
  if (s < M) {
    i = a.perm[s];
    bi = a.beh[i];
    act = nb_may_pair(bi, a.events);
    xi = a.x[i];
    yi = a.y[i];
    zi = a.z[i];
    ri = a.r[i];
    // a box farther than R_i + R_max holds no partner; the slack covers the box test's rounding
    const double reach = ri + a.rmax;
    lim = reach + 1e-12 * (fabs(xi) + fabs(yi) + fabs(zi) + reach) + 1e-300;
  }
  const unsigned long long start = __builtin_amdgcn_ballot_w64(act);
  if (start == 0) return;
  if (lane == 0) {
    stk_code[w][0] = 0;
    stk_mask[w][0] = start;
  }
  int sp = 1;
  bh_wave_sync();
  while (sp > 0) {
    sp--;
    const int code = __builtin_amdgcn_readfirstlane(stk_code[w][sp]);
    const unsigned long long msk = stk_mask[w][sp];
    const bool on = (msk >> lane) & 1ull;
    if (code <= -2) {
      if (on) bh_test_pair(a, i, bi, xi, yi, zi, ri, -2 - code);
      continue;
    }
    const int32_t *ch = a.nchild + (size_t)code * 8;
    if (ch[0] == kBucket) {
      if (on)
        for (int s2 = ch[1]; s2 < ch[1] + ch[2]; s2++)
          if (s2 != s) bh_test_pair(a, i, bi, xi, yi, zi, ri, s2);
      continue;
    }
    const double4 bx = a.nbox[code];
    bool pass = false;
    if (on) {
      const double ex = fmax(fabs(xi - bx.x) - bx.w, 0.0);
      const double ey = fmax(fabs(yi - bx.y) - bx.w, 0.0);
      const double ez = fmax(fabs(zi - bx.z) - bx.w, 0.0);
      const double slack = 1e-12 * (fabs(bx.x) + fabs(bx.y) + fabs(bx.z) + bx.w);
      const double l = lim + slack;
      pass = ex * ex + ey * ey + ez * ez <= l * l;
    }
    const unsigned long long pm = __builtin_amdgcn_ballot_w64(pass);
    if (pm && sp + 8 <= kBhStack) {
      for (int c = 7; c >= 0; c--) {
        const int32_t cc = ch[c];
        if (cc == -1) continue;
        if (lane == 0) {
          stk_code[w][sp] = cc;
          stk_mask[w][sp] = pm;
        }
        sp++;
      }
    }
    bh_wave_sync();
  }
}

__global__ void k_bh_zero(double *acc, size_t len) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < len;
       k += (size_t)gridDim.x * blockDim.x)
    acc[k] = 0.0;
}

// ---- host --------------------------------------------------------------------------------------
static int nb256(int64_t n) { return (int)std::max<int64_t>(1, (n + 255) / 256); }

void bh_free(Bh *b) {
  if (!b) return;
  void *ps[] = {b->keys_in, b->keys, b->perm_in, b->perm, b->sx,     b->sy,     b->sz,
                b->sm,      b->lcx,  b->lcy,     b->lcz,  b->depth,  b->suns,   b->nbox,
                b->nmc,     b->nchild, b->nrange, b->bpart, b->sort_tmp, b->cnt, b->slots,
                b->pslot};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  delete b;
}

static int bh_ensure(Ctx *c) {
  Bh *b = c->bh;
  if (!b->cnt) {
    NB_HIP(c, hipMalloc((void **)&b->cnt, sizeof(BhCounters)));
    NB_HIP(c, hipMemset(b->cnt, 0, sizeof(BhCounters)));
    NB_HIP(c, balloc(&b->bpart, 1024 * 8));
  }
  if (b->cap >= c->cap) return NB_OK;
  const size_t n = (size_t)c->cap;
  NB_HIP(c, balloc(&b->keys_in, n));
  NB_HIP(c, balloc(&b->keys, n));
  NB_HIP(c, balloc(&b->perm_in, n));
  NB_HIP(c, balloc(&b->perm, n));
  NB_HIP(c, balloc(&b->sx, n));
  NB_HIP(c, balloc(&b->sy, n));
  NB_HIP(c, balloc(&b->sz, n));
  NB_HIP(c, balloc(&b->sm, n));
  NB_HIP(c, balloc(&b->lcx, n));
  NB_HIP(c, balloc(&b->lcy, n));
  NB_HIP(c, balloc(&b->lcz, n));
  NB_HIP(c, balloc(&b->depth, n));
  NB_HIP(c, balloc(&b->suns, n));
  // every node holds 2+ bodies and the nodes of one level are disjoint: at most n/2 per level
  b->node_cap = (int64_t)(n / 2) * kBhLevels + 64;
  NB_HIP(c, balloc(&b->nbox, (size_t)b->node_cap));
  NB_HIP(c, balloc(&b->nmc, (size_t)b->node_cap));
  NB_HIP(c, balloc(&b->nchild, (size_t)b->node_cap * 8));
  NB_HIP(c, balloc(&b->slots, (size_t)b->node_cap * 8));
  NB_HIP(c, balloc(&b->pslot, (size_t)b->node_cap));
  NB_HIP(c, balloc(&b->nrange, (size_t)b->node_cap));
  size_t bytes = 0;
  NB_HIP(c, rocprim::radix_sort_pairs(nullptr, bytes, b->keys_in, b->keys, b->perm_in, b->perm,
                                      n, 0, 64, c->stream));
  void *tmp = nullptr;
  NB_HIP(c, hipMalloc(&tmp, bytes ? bytes : 1));
  if (b->sort_tmp) (void)hipFree(b->sort_tmp);
  b->sort_tmp = tmp;
  b->sort_bytes = bytes;
  b->cap = c->cap;
  return NB_OK;
}

int bh_force_phase(Ctx *c, bool timed, int ev_slot) {
  Bh *b = c->bh;
  hipStream_t st = c->stream;
  NB_HIP(c, launch_zero_counters(c->ctr, 0, st));
  int rc = bh_ensure(c);
  if (rc) return rc;
  const int n = (int)c->n;
  const int D = b->p.max_depth;
  const size_t acc_len = (size_t)3 * c->npad;
  hipLaunchKernelGGL(k_bh_zero, dim3(nb256((int64_t)std::min<size_t>(acc_len, 1 << 20))),
                     dim3(256), 0, st, c->acc, acc_len);
  if (n == 0) return NB_OK;
  const int nparts = std::min(1024, nb256(n));
  hipLaunchKernelGGL(k_bh_box_part, dim3(nparts), dim3(256), 0, st, c->x, c->y, c->z, c->flags,
                     n, b->bpart);
  hipLaunchKernelGGL(k_bh_root, dim3(1), dim3(256), 0, st, b->bpart, nparts, b->cnt, b->nbox,
                     b->nrange, b->nchild);
  hipLaunchKernelGGL(k_bh_keys, dim3(nb256(n)), dim3(256), 0, st, c->x, c->y, c->z, c->flags, n,
                     b->nbox, D, b->keys_in, b->perm_in, b->depth);
  NB_HIP(c, hipGetLastError());
  size_t bytes = b->sort_bytes;
  NB_HIP(c, rocprim::radix_sort_pairs(b->sort_tmp, bytes, b->keys_in, b->keys, b->perm_in,
                                      b->perm, (size_t)n, 0, 64, st));
  hipLaunchKernelGGL(k_bh_gather, dim3(nb256(n)), dim3(256), 0, st, b->perm, c->x, c->y, c->z,
                     c->m, b->cnt, b->sx, b->sy, b->sz, b->sm, b->lcx, b->lcy, b->lcz, b->depth);
  const int gl = std::min(2048, nb256(std::max(1, n / 2)));          // one thread per node
  const int gl8 = std::min(4096, nb256((int64_t)std::max(1, n / 2) * 8));  // eight per node
  BhLevelArgs la;
  la.keys = b->keys;
  la.perm = b->perm;
  la.cnt = b->cnt;
  la.nbox = b->nbox;
  la.nrange = b->nrange;
  la.nchild = b->nchild;
  la.depth = b->depth;
  la.node_cap = b->node_cap;
  la.D = D;
  for (int d = 0; d < D; d++) {
    la.d = d;
    hipLaunchKernelGGL(k_bh_level, dim3(gl8), dim3(256), 0, st, la);
  }
  BhMassArgs ma;
  ma.cnt = b->cnt;
  ma.nchild = b->nchild;
  ma.nrange = b->nrange;
  ma.nmc = b->nmc;
  ma.nbox = b->nbox;
  ma.slots = b->slots;
  ma.sx = b->sx; ma.sy = b->sy; ma.sz = b->sz; ma.sm = b->sm;
  ma.lcx = b->lcx; ma.lcy = b->lcy; ma.lcz = b->lcz;
  ma.node_cap = b->node_cap;
  for (int d = D; d >= 0; d--) {
    ma.d = d;
    hipLaunchKernelGGL(k_bh_mass, dim3(gl), dim3(256), 0, st, ma);
  }
  const int gn = std::min(4096, nb256(std::max(1, n)));  // nodes < n
  hipLaunchKernelGGL(k_bh_parent, dim3(gn), dim3(256), 0, st, b->cnt, b->nchild, D, b->node_cap,
                     b->pslot);
  hipLaunchKernelGGL(k_bh_links, dim3(gn), dim3(256), 0, st, b->cnt, b->nchild, b->pslot, D,
                     b->node_cap, b->slots);
  NB_HIP(c, hipGetLastError());
  BhForceArgs fa;
  fa.cnt = b->cnt;
  fa.cnt_w = b->cnt;
  fa.perm = b->perm;
  fa.keys = b->keys;
  fa.flags = c->flags;
  fa.sx = b->sx; fa.sy = b->sy; fa.sz = b->sz; fa.sm = b->sm;
  fa.nbox = b->nbox;
  fa.nmc = b->nmc;
  fa.nchild = b->nchild;
  fa.slots = b->slots;
  // 64 bodies per wave unless asked: smaller groups only repeat the shared part of the walk in
  // more waves (measured 6.9 / 8.4 / 10.7 ms at N = 65,536 for 64 / 32 / 16)
  fa.gsz = b->p.group ? b->p.group : 64;
  fa.node_cap = b->node_cap;
  fa.acc = c->acc;
  fa.npad = (int)c->npad;
  fa.G = c->p.G;
  fa.theta = b->p.theta;
  {
    const double t2 = b->p.theta * b->p.theta;
    fa.theta2 = (t2 >= 0x1p-1000 && t2 <= 0x1p1000) ? t2 : std::nan("");
  }
  fa.soft2 = b->p.softening * b->p.softening;
  if (timed) NB_HIP(c, hipEventRecord(c->ev[2 * ev_slot], st));
  if (b->p.mode == 1)
    hipLaunchKernelGGL(k_bh_force_fast, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, fa);
  else {
    // by default the shared-reciprocal term once the walk has four waves per SIMD (issue-bound)
    const int64_t waves = (n + fa.gsz - 1) / fa.gsz;
    const bool shared = b->p.term == 2 || (b->p.term == 0 && waves >= (int64_t)4 * 4 * b->cus);
    hipLaunchKernelGGL(shared ? k_bh_force<true> : k_bh_force<false>,
                       dim3((unsigned)((n + 4 * fa.gsz - 1) / (4 * fa.gsz))), dim3(256), 0, st,
                       fa);
  }
  hipLaunchKernelGGL(k_bh_sun_list, dim3(nb256(n)), dim3(256), 0, st, c->flags, n, b->cnt,
                     b->suns);
  BhSunArgs sa;
  sa.cnt = b->cnt;
  sa.suns = b->suns;
  sa.x = c->x; sa.y = c->y; sa.z = c->z; sa.m = c->m;
  sa.flags = c->flags;
  sa.acc = c->acc;
  sa.n = n;
  sa.npad = (int)c->npad;
  sa.G = c->p.G;
  sa.soft2 = fa.soft2;
  hipLaunchKernelGGL(k_bh_sun, dim3(64), dim3(64), 0, st, sa);
  if (timed) NB_HIP(c, hipEventRecord(c->ev[2 * ev_slot + 1], st));
  NB_HIP(c, hipGetLastError());
  if (c->p.collisions) {
    BhDetectArgs da;
    da.cnt = b->cnt;
    da.perm = b->perm;
    da.x = c->x; da.y = c->y; da.z = c->z; da.r = c->r;
    da.beh = c->beh;
    da.nbox = b->nbox;
    da.nchild = b->nchild;
    da.pairs = c->pairs;
    da.ctr = c->ctr;
    da.pair_cap = c->pair_cap;
    da.rmax = c->rmax;
    da.events = c->p.events;
    hipLaunchKernelGGL(k_bh_detect, dim3(nb256(n)), dim3(256), 0, st, da);
    NB_HIP(c, hipGetLastError());
    {
      const int rcs = sort_pairs(c);
      if (rcs) return rcs;
    }
  }
  b->builds++;
  return NB_OK;
}

}  // namespace nb

using namespace nb;

extern "C" {

int nb_bh_params_init(nb_bh_params *p) {
  if (!p) return NB_E_INVALID;
  memset(p, 0, sizeof *p);
  p->theta = 0.3;
  p->softening = 1e-4;
  p->max_depth = kBhLevels;
  return NB_OK;
}

int nb_bh_enable(nb_ctx *c, const nb_bh_params *p) {
  if (!c) return NB_E_INVALID;
  nb_bh_params q;
  if (p)
    q = *p;
  else
    nb_bh_params_init(&q);
  if (!(q.theta >= 0) || !std::isfinite(q.theta) || !(q.softening >= 0) ||
      !std::isfinite(q.softening) || q.max_depth < 1 || q.max_depth > kBhLevels ||
      !(q.group == 0 || q.group == 16 || q.group == 32 || q.group == 64) ||
      (q.mode != 0 && q.mode != 1) || q.term < 0 || q.term > 2)
    return fail_msg(c, NB_E_INVALID,
                    "nb_bh_enable: theta, softening >= 0, max_depth in [1, 21], group 0/16/32/64, "
                    "mode 0/1, term 0/1/2");
  if (c->p.precision != NB_PREC_FP64)
    return fail_msg(c, NB_E_UNSUPPORTED, "nb_bh_enable: the octree path is FP64 (precision fp64)");
  if (c->sur_on)
    return fail_msg(c, NB_E_UNSUPPORTED, "nb_bh_enable: the surrogate is on; disable it first");
  NB_HIP(c, hipSetDevice(c->device));
  if (!c->bh) {
    c->bh = new Bh();
    NB_HIP(c, hipDeviceGetAttribute(&c->bh->cus, hipDeviceAttributeMultiprocessorCount,
                                    c->device));
  }
  c->bh->p = q;
  c->bh_on = true;
  return NB_OK;
}

int nb_bh_disable(nb_ctx *c) {
  if (!c) return NB_E_INVALID;
  c->bh_on = false;
  return NB_OK;
}

int nb_bh_get_stats(nb_ctx *c, nb_bh_stats *out) {
  if (!c || !out) return NB_E_INVALID;
  memset(out, 0, sizeof *out);
  if (!c->bh || !c->bh->cnt) return NB_OK;
  NB_HIP(c, hipSetDevice(c->device));
  BhCounters h;
  NB_HIP(c, hipStreamSynchronize(c->stream));
  NB_HIP(c, hipMemcpy(&h, c->bh->cnt, sizeof h, hipMemcpyDeviceToHost));
  int64_t nodes = 0;
  for (int d = 0; d <= kBhLevels; d++) nodes += h.lcount[d];
  out->builds = c->bh->builds;
  out->bodies = h.M;
  out->nodes = nodes;
  out->buckets = h.buckets;
  out->interactions = (int64_t)h.interactions;
  out->depth_max = h.depth_max;
  out->overflow = (int32_t)(h.stack_overflow | h.node_overflow);
  return NB_OK;
}

int nb_bh_get_depths(nb_ctx *c, int32_t *depth) {
  if (!c || !depth) return NB_E_INVALID;
  if (!c->bh || !c->bh->depth || c->bh->builds == 0)
    return fail_msg(c, NB_E_STATE, "nb_bh_get_depths: no tree built yet");
  if (c->n == 0) return NB_OK;
  NB_HIP(c, hipSetDevice(c->device));
  NB_HIP(c, hipMemcpyAsync(depth, c->bh->depth, (size_t)c->n * sizeof(int32_t),
                           hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  return NB_OK;
}

}  // extern "C"
