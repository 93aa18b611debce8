// nb_batch.hip -- many independent simulations of uneven size in one persistent launch.
//
// No counterpart in the reference (it runs one BodyCollection per process); each simulation
// follows the same per-cycle semantics as the single-context path: FP32 pairwise force with Kahan
// tile sums (Body.Compute's force loop, body.go:156-170, :214-225, softened per octree.go:314-325)
// and the FP64 Body.Update (body.go:114-139).  Collisions are not part of this path.
//
// Layout: the simulations' bodies are concatenated; each simulation is padded to a multiple of 64
// with far-away massless sentinels (flags = 0), so no 64-source tile straddles two simulations.
// Work items are (simulation, block of 64*IPL targets), sorted longest first (cost = targets x
// sources) and pulled from a device queue by a persistent grid: every workgroup takes items until
// the queue is exhausted, so the grid drains on its own.  The W waves of a workgroup (W = 1, 2 or
// 4) share the item's targets and split its sources into W contiguous ranges; their
// Kahan-compensated partial sums are combined through LDS in wave order, so results do not
// depend on scheduling (with W = 1 each wave takes items of its own).  The target-block width
// (IPL) and the Kahan tile (64, 128 or 256 sources) are chosen per size bucket, W per launch, by
// nb_batch_autotune.  When every item has the same (IPL, tile) the launch uses the kernel
// specialised to that shape, whose registers and LDS are sized for it alone.
//
// One nb_batch_step call runs nsteps cycles as one hipGraph launch: per cycle a fill kernel resets
// the queue (no memset nodes), then the force kernel, then the update kernel.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "nb_internal.hpp"

namespace nb {

constexpr int kSimPad = 64;
constexpr int kBatchMaxTargets = 256; // 64 lanes * IPL (IPL <= 4)
constexpr int kBuckets = 3;           // size buckets for shape selection
constexpr int kBatchMaxTile = 256;    // sources per Kahan tile (64, 128 or 256)

struct BatchItem {
  int32_t t0;     // first target (global index)
  int32_t tn;     // targets in this block (<= 64 * ipl)
  int32_t s0;     // first source of the simulation (global index)
  int32_t sn;     // sources (padded simulation size, multiple of 64)
  int32_t ipl;    // 1, 2 or 4
  int32_t tile;   // sources per Kahan tile: 64, 128 or 256
  int32_t pad[2];
};

struct BatchQueue {
  uint32_t next;
  uint32_t pad[63];  // keep the counter on its own cache line
};


// Orders this wave's LDS writes before its later reads (and reads before later overwrites) of its
// private tile; the waitcnt of a workgroup-scope fence covers the LDS queue.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// One work item on the W waves of a workgroup: its 64 * IPL targets against the simulation's
// sources, which the waves split into W contiguous ranges of TILE-source tiles (the last tile of
// a range may be shorter: the range ends on a 64-source boundary).  Each wave stages its tile in
// its own LDS buffer (one 16-byte load per lane per 64 sources), reads it back as wave-uniform
// broadcasts, keeps plain FP32 partial sums within the tile and Kahan-adds them at the tile's
// end.  With W > 1 the waves' FP64 results are added in wave order through LDS.
template <int W, int IPL, int TILE, int NT, int NS>
__device__ void batch_item(const BatchItem &it, const float4 *__restrict__ js, double *acc,
                           int npad, float eps2_s, double (*lds)[3][NT],
                           float4 (*tiles_lds)[NS]) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;          // the workgroup's wave: its LDS tile
  const int w = W == 1 ? 0 : wave;            // its share of the item's sources
  const float eps2 = to_vgpr(eps2_s);
  float4 *tl = tiles_lds[wave];
  float xi[IPL], yi[IPL], zi[IPL], sx[IPL], sy[IPL], sz[IPL], cx[IPL], cy[IPL], cz[IPL];
#pragma unroll
  for (int k = 0; k < IPL; k++) {
    const int local = lane + 64 * k;
    const float4 b = local < it.tn ? js[it.t0 + local]
                                   : make_float4(kSentinel, kSentinel, kSentinel, 0.f);
    xi[k] = b.x;
    yi[k] = b.y;
    zi[k] = b.z;
    sx[k] = sy[k] = sz[k] = cx[k] = cy[k] = cz[k] = 0.f;
  }
  // wave w takes the 64-source blocks [w*blocks/W, (w+1)*blocks/W) of the simulation
  const int blocks = it.sn / 64;
  const int ja = (w * blocks / W) * 64, jb = ((w + 1) * blocks / W) * 64;
  for (int t0 = ja; t0 < jb; t0 += TILE) {
    const int cnt = min(TILE, jb - t0);  // a multiple of 64
    float px[IPL], py[IPL], pz[IPL];
#pragma unroll
    for (int k = 0; k < IPL; k++) px[k] = py[k] = pz[k] = 0.f;
    for (int q = lane; q < cnt; q += 64) tl[q] = js[it.s0 + t0 + q];
    wave_lds_sync();
    auto interact = [&](const float4 b) {
      float dx[IPL], dy[IPL], dz[IPL], inv[IPL];
#pragma unroll
      for (int k = 0; k < IPL; k++) {
        dx[k] = b.x - xi[k];
        dy[k] = b.y - yi[k];
        dz[k] = b.z - zi[k];
        float r2 = __builtin_fmaf(dx[k], dx[k], eps2);
        r2 = __builtin_fmaf(dy[k], dy[k], r2);
        r2 = __builtin_fmaf(dz[k], dz[k], r2);
        inv[k] = __builtin_amdgcn_rsqf(r2);
      }
#pragma unroll
      for (int k = 0; k < IPL; k++) {
        float s = b.w * inv[k];
        const float i2 = inv[k] * inv[k];
        s = s * i2;
        px[k] = __builtin_fmaf(dx[k], s, px[k]);
        py[k] = __builtin_fmaf(dy[k], s, py[k]);
        pz[k] = __builtin_fmaf(dz[k], s, pz[k]);
      }
    };
    // cnt is a multiple of 64.  One target per lane: eight sources per step are read first
    // (eight ds_read_b128 in flight, one base address, 63 VGPRs).  Two or four: sixteen sources
    // per unrolled iteration share one base address (the compiler keeps fewer reads ahead, which
    // holds the registers of those shapes lowest).
    if constexpr (IPL == 1) {
      for (int j8 = 0; j8 < cnt; j8 += 8) {
        float4 b8[8];
#pragma unroll
        for (int u = 0; u < 8; u++) b8[u] = tl[j8 + u];
#pragma unroll
        for (int u = 0; u < 8; u++) interact(b8[u]);
      }
    } else {
#pragma unroll 16
      for (int jj = 0; jj < cnt; jj++) interact(tl[jj]);
    }
#pragma unroll
    for (int k = 0; k < IPL; k++) {
      float y, s;
      y = px[k] - cx[k]; s = sx[k] + y; cx[k] = (s - sx[k]) - y; sx[k] = s;
      y = py[k] - cy[k]; s = sy[k] + y; cy[k] = (s - sy[k]) - y; sy[k] = s;
      y = pz[k] - cz[k]; s = sz[k] + y; cz[k] = (s - sz[k]) - y; sz[k] = s;
    }
    wave_lds_sync();  // all reads of this tile precede the next tile's writes
  }
  if (W == 1) {
#pragma unroll
    for (int k = 0; k < IPL; k++) {
      const int local = lane + 64 * k;
      if (local < it.tn) {
        const int i = it.t0 + local;
        acc[i] = (double)sx[k] - (double)cx[k];
        acc[npad + i] = (double)sy[k] - (double)cy[k];
        acc[2 * (size_t)npad + i] = (double)sz[k] - (double)cz[k];
      }
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < IPL; k++) {
    const int local = lane + 64 * k;
    lds[w][0][local] = (double)sx[k] - (double)cx[k];
    lds[w][1][local] = (double)sy[k] - (double)cy[k];
    lds[w][2][local] = (double)sz[k] - (double)cz[k];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < it.tn; q += blockDim.x) {
    double f[3];
    for (int c = 0; c < 3; c++) {
      double v = lds[0][c][q];
      for (int ww = 1; ww < W; ww++) v += lds[ww][c][q];
      f[c] = v;
    }
    const int i = it.t0 + q;
    acc[i] = f[0];
    acc[npad + i] = f[1];
    acc[2 * (size_t)npad + i] = f[2];
  }
}

template <int W, int IPL>
__device__ __forceinline__ void batch_item_tile(const BatchItem &it, const float4 *js,
                                                double *acc, int npad, float eps2,
                                                double (*lds)[3][kBatchMaxTargets],
                                                float4 (*tiles_lds)[kBatchMaxTile]) {
  constexpr int NT = kBatchMaxTargets, NS = kBatchMaxTile;
  if (it.tile == 256) batch_item<W, IPL, 256, NT, NS>(it, js, acc, npad, eps2, lds, tiles_lds);
  else if (it.tile == 128) batch_item<W, IPL, 128, NT, NS>(it, js, acc, npad, eps2, lds, tiles_lds);
  else batch_item<W, IPL, 64, NT, NS>(it, js, acc, npad, eps2, lds, tiles_lds);
}

// Persistent grid: each workgroup (W waves) takes items from the queue until it is exhausted, so
// the grid drains on its own.  W = 1: the four waves of a 256-thread workgroup are independent
// and each takes its own items.  IPL = 0 reads each item's shape; a launch whose items all share
// one shape uses the kernel specialised to it (IPL, TILE > 0), whose registers are sized for
// that shape alone.
template <int W, int IPL, int TILE>
__global__ __launch_bounds__(256) void k_batch_force(const BatchItem *__restrict__ items,
                                                     int nitems, BatchQueue *queue,
                                                     const float4 *__restrict__ js, double *acc,
                                                     int npad, float eps2) {
  // LDS sized for the shape: the specialised kernel's item targets and tile only
  constexpr int NT = IPL > 0 ? 64 * IPL : kBatchMaxTargets;
  constexpr int NS = TILE > 0 ? TILE : kBatchMaxTile;
  __shared__ double lds[W > 1 ? W : 1][3][NT];
  __shared__ float4 tiles_lds[W > 1 ? W : 4][NS];
  __shared__ int cur[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (;;) {
    int q;
    if (W == 1) {  // per wave: lane 0 takes the next item, the wave reads it back
      if (lane == 0) cur[w] = (int)atomicAdd(&queue->next, 1u);
      wave_lds_sync();
      q = cur[w];
      wave_lds_sync();
    } else {
      __syncthreads();
      if (threadIdx.x == 0) cur[0] = (int)atomicAdd(&queue->next, 1u);
      __syncthreads();
      q = cur[0];
    }
    if (q >= nitems) break;
    const BatchItem it = items[q];
    if constexpr (IPL > 0) {
      batch_item<W, (IPL > 0 ? IPL : 1), (TILE > 0 ? TILE : 64), NT, NS>(it, js, acc, npad, eps2,
                                                                         lds, tiles_lds);
    } else {
      if (it.ipl == 4)
        batch_item_tile<W, 4>(it, js, acc, npad, eps2, lds, tiles_lds);
      else if (it.ipl == 2)
        batch_item_tile<W, 2>(it, js, acc, npad, eps2, lds, tiles_lds);
      else
        batch_item_tile<W, 1>(it, js, acc, npad, eps2, lds, tiles_lds);
    }
  }
}

__global__ void k_batch_reset(BatchQueue *q) {
  if (threadIdx.x == 0) q->next = 0;
}

}  // namespace nb

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
using namespace nb;

struct nb_batch {
  nb_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int nsims = 0;
  int64_t n = 0;           // padded total (multiple of 64, then of 1024)
  std::vector<int64_t> off;     // caller offsets (unpadded), nsims + 1
  std::vector<int64_t> poff;    // padded offsets, nsims + 1
  double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *m = nullptr, *rest = nullptr, *acc = nullptr;
  uint8_t *flags = nullptr, *collided = nullptr;
  float4 *js = nullptr;
  DevCounters *ctr = nullptr;
  BatchItem *items = nullptr;
  BatchQueue *queue = nullptr;
  int nitems = 0;
  int bucket_ipl[kBuckets] = {2, 4, 4};
  int bucket_tile[kBuckets] = {64, 64, 64};
  int waves = 4;  // waves per work item (1, 2 or 4)
  int items_ipl = 0, items_tile = 0;  // the shape every current item shares (0: mixed)
  int cus = 0;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = -1;
  double graph_dt = 0, graph_R = 0;
  double interactions_per_step = 0;
  int64_t n_active = 0;
};

namespace {

int bfail(nb_batch *b, int code, const std::string &m) {
  b->err = m;
  return code;
}

bool bhip(nb_batch *b, hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  b->err = buf;
  return false;
}
#define NB_BHIP(b, call)                               \
  do {                                                 \
    if (!bhip((b), (call), #call)) return NB_E_HIP;    \
  } while (0)

int bucket_of(int64_t nk) { return nk < 1024 ? 0 : nk < 4096 ? 1 : 2; }

void drop_graph(nb_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_steps = -1;
}

void free_batch(nb_batch *b) {
  drop_graph(b);
  void *ptrs[] = {b->x, b->y, b->z, b->vx, b->vy, b->vz, b->m, b->rest, b->acc,
                  b->flags, b->collided, b->js, b->ctr, b->items, b->queue};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  b->x = b->y = b->z = b->vx = b->vy = b->vz = b->m = b->rest = b->acc = nullptr;
  b->flags = b->collided = nullptr;
  b->js = nullptr;
  b->ctr = nullptr;
  b->items = nullptr;
  b->queue = nullptr;
}

// Build the item list (longest first) with the per-bucket IPL and tile.
int build_items(nb_batch *b, int only_bucket, int force_ipl, int force_tile = 0) {
  std::vector<BatchItem> v;
  for (int k = 0; k < b->nsims; k++) {
    const int64_t nk = b->off[k + 1] - b->off[k];
    if (nk == 0) continue;
    const int bk = bucket_of(nk);
    if (only_bucket >= 0 && bk != only_bucket) continue;
    const int ipl = force_ipl ? force_ipl : b->bucket_ipl[bk];
    const int tile = force_tile ? force_tile : b->bucket_tile[bk];
    const int64_t sn = b->poff[k + 1] - b->poff[k];
    const int64_t blk = 64 * ipl;
    for (int64_t t = 0; t < nk; t += blk) {
      BatchItem it{};
      it.t0 = (int32_t)(b->poff[k] + t);
      it.tn = (int32_t)std::min<int64_t>(blk, nk - t);
      it.s0 = (int32_t)b->poff[k];
      it.sn = (int32_t)sn;
      it.ipl = ipl;
      it.tile = tile;
      v.push_back(it);
    }
  }
  std::stable_sort(v.begin(), v.end(), [](const BatchItem &a, const BatchItem &c) {
    return (int64_t)a.tn * a.sn > (int64_t)c.tn * c.sn;
  });
  b->nitems = (int)v.size();
  b->items_ipl = v.empty() ? 0 : v[0].ipl;
  b->items_tile = v.empty() ? 0 : v[0].tile;
  for (const BatchItem &it : v)
    if (it.ipl != b->items_ipl || it.tile != b->items_tile) {
      b->items_ipl = b->items_tile = 0;
      break;
    }
  if (!v.empty())
    NB_BHIP(b, hipMemcpyAsync(b->items, v.data(), v.size() * sizeof(BatchItem),
                              hipMemcpyHostToDevice, b->stream));
  NB_BHIP(b, hipStreamSynchronize(b->stream));
  drop_graph(b);
  return NB_OK;
}

template <int W>
static const void *batch_kernel(int ipl, int tile) {
  switch (ipl * 1000 + tile) {
    case 1064: return (const void *)k_batch_force<W, 1, 64>;
    case 1128: return (const void *)k_batch_force<W, 1, 128>;
    case 1256: return (const void *)k_batch_force<W, 1, 256>;
    case 2064: return (const void *)k_batch_force<W, 2, 64>;
    case 2128: return (const void *)k_batch_force<W, 2, 128>;
    case 2256: return (const void *)k_batch_force<W, 2, 256>;
    case 4064: return (const void *)k_batch_force<W, 4, 64>;
    case 4128: return (const void *)k_batch_force<W, 4, 128>;
    case 4256: return (const void *)k_batch_force<W, 4, 256>;
    default: return (const void *)k_batch_force<W, 0, 0>;
  }
}

hipError_t launch_batch_force(nb_batch *b) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, b->stream, b->queue);
  float eps2 = (float)(b->p.eps * b->p.eps);
  // items of one shape: the kernel specialised to it; otherwise the per-item dispatch
  const int ipl = b->items_ipl, tile = b->items_tile;
  // 32 waves per CU: 8 workgroups of 4 waves (W = 4, and W = 1 with independent waves), 16 of 2
  const void *fn = b->waves == 2   ? batch_kernel<2>(ipl, tile)
                   : b->waves == 1 ? batch_kernel<1>(ipl, tile)
                                   : batch_kernel<4>(ipl, tile);
  const dim3 grid((unsigned)(b->cus * (b->waves == 2 ? 16 : 8)));
  const dim3 block(b->waves == 2 ? 128 : 256);
  const BatchItem *items = b->items;
  int nitems = b->nitems, npad = (int)b->n;
  BatchQueue *queue = b->queue;
  const float4 *js = b->js;
  double *acc = b->acc;
  void *args[] = {&items, &nitems, &queue, &js, &acc, &npad, &eps2};
  return hipLaunchKernel(fn, grid, block, args, 0, b->stream);
}

hipError_t enqueue_cycle(nb_batch *b, double dt, double R) {
  hipError_t e = launch_batch_force(b);
  if (e != hipSuccess) return e;
  UpdateArgs u;
  u.x = b->x; u.y = b->y; u.z = b->z; u.vx = b->vx; u.vy = b->vy; u.vz = b->vz;
  u.m = b->m;
  u.rest = b->rest;
  u.flags = b->flags;
  u.collided = b->collided;
  u.acc = b->acc;
  u.split = 1;
  u.npad = (int)b->n;
  u.n = (int)b->n;
  u.G = b->p.G;
  u.dt = dt;
  u.R = R;
  u.force_mode = 0;
  u.js = b->js;
  u.ctr = b->ctr;
  return launch_update(u, b->stream);
}

}  // namespace

extern "C" {

int nb_batch_create(const nb_params *p, nb_batch **out) {
  if (!out || !p) return NB_E_INVALID;
  *out = nullptr;
  if (!(p->eps > 0.0) || !std::isfinite(p->eps) || (float)(p->eps * p->eps) < 1e-30f)
    return NB_E_INVALID;  // the batched path is the softened FP32 force
  if (p->collisions || p->overlap_cutoff || p->events || p->precision != NB_PREC_FP32)
    return NB_E_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || p->device < 0 || p->device >= ndev)
    return NB_E_HIP;
  nb_batch *b = new (std::nothrow) nb_batch();
  if (!b) return NB_E_NOMEM;
  b->p = *p;
  b->device = p->device;
  if (hipSetDevice(b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return NB_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) {
    (void)hipStreamDestroy(b->stream);
    delete b;
    return NB_E_HIP;
  }
  b->cus = prop.multiProcessorCount;
  *out = b;
  return NB_OK;
}

void nb_batch_destroy(nb_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_batch(b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char *nb_batch_last_error(const nb_batch *b) { return b ? b->err.c_str() : "NULL batch"; }

int nb_batch_load(nb_batch *b, int32_t nsims, const int64_t *offsets, const double *x,
                  const double *y, const double *z, const double *vx, const double *vy,
                  const double *vz, const double *m, const uint8_t *flags) {
  if (!b) return NB_E_INVALID;
  if (nsims < 1 || !offsets) return bfail(b, NB_E_INVALID, "nb_batch_load: need nsims >= 1");
  if (offsets[0] != 0) return bfail(b, NB_E_INVALID, "nb_batch_load: offsets[0] must be 0");
  for (int k = 0; k < nsims; k++)
    if (offsets[k + 1] < offsets[k])
      return bfail(b, NB_E_INVALID, "nb_batch_load: offsets must be non-decreasing");
  const int64_t total = offsets[nsims];
  if (total > ((int64_t)1 << 30)) return bfail(b, NB_E_INVALID, "nb_batch_load: too many bodies");
  if (total > 0 && (!x || !y || !z || !vx || !vy || !vz || !m))
    return bfail(b, NB_E_INVALID, "nb_batch_load: NULL array");
  NB_BHIP(b, hipSetDevice(b->device));
  free_batch(b);
  b->nsims = nsims;
  b->off.assign(offsets, offsets + nsims + 1);
  b->poff.assign(nsims + 1, 0);
  for (int k = 0; k < nsims; k++)
    b->poff[k + 1] = b->poff[k] + round_up(offsets[k + 1] - offsets[k], kSimPad);
  b->n = std::max<int64_t>(round_up(b->poff[nsims], kPad), kPad);
  const int64_t n = b->n;
  // host staging in the padded layout
  std::vector<double> h[7];
  for (auto &v : h) v.assign((size_t)n, 0.0);
  std::vector<uint8_t> hf((size_t)n, 0);
  const double *src[7] = {x, y, z, vx, vy, vz, m};
  b->n_active = 0;
  b->interactions_per_step = 0;
  for (int k = 0; k < nsims; k++) {
    const int64_t nk = offsets[k + 1] - offsets[k];
    for (int64_t q = 0; q < nk; q++) {
      const size_t d = (size_t)(b->poff[k] + q), s = (size_t)(offsets[k] + q);
      for (int c = 0; c < 7; c++) h[c][d] = src[c][s];
      hf[d] = flags ? flags[s] : (uint8_t)NB_F_EXISTS;
      if ((hf[d] & NB_F_EXISTS) && !(hf[d] & NB_F_FRAGMENTING)) b->n_active++;
    }
    b->interactions_per_step += (double)nk * (double)(nk - 1);
  }
  double **dst[7] = {&b->x, &b->y, &b->z, &b->vx, &b->vy, &b->vz, &b->m};
  for (int c = 0; c < 7; c++) {
    NB_BHIP(b, hipMalloc((void **)dst[c], (size_t)n * sizeof(double)));
    NB_BHIP(b, hipMemcpyAsync(*dst[c], h[c].data(), (size_t)n * sizeof(double),
                              hipMemcpyHostToDevice, b->stream));
  }
  NB_BHIP(b, hipMalloc((void **)&b->rest, (size_t)n * sizeof(double)));
  NB_BHIP(b, hipMalloc((void **)&b->acc, (size_t)3 * n * sizeof(double)));
  NB_BHIP(b, hipMalloc((void **)&b->flags, (size_t)n));
  NB_BHIP(b, hipMalloc((void **)&b->collided, (size_t)n));
  NB_BHIP(b, hipMalloc((void **)&b->js, (size_t)n * sizeof(float4)));
  NB_BHIP(b, hipMalloc((void **)&b->ctr, sizeof(DevCounters)));
  NB_BHIP(b, hipMalloc((void **)&b->queue, sizeof(BatchQueue)));
  NB_BHIP(b, hipMemcpyAsync(b->flags, hf.data(), (size_t)n, hipMemcpyHostToDevice, b->stream));
  NB_BHIP(b, hipMemsetAsync(b->collided, 0, (size_t)n, b->stream));
  NB_BHIP(b, hipMemsetAsync(b->acc, 0, (size_t)3 * n * sizeof(double), b->stream));
  // worst case item count: one item per 64 targets
  NB_BHIP(b, hipMalloc((void **)&b->items, (size_t)(n / 64 + nsims + 1) * sizeof(BatchItem)));
  NB_BHIP(b, launch_zero_counters(b->ctr, 1, b->stream));
  NB_BHIP(b, launch_write_stream(b->x, b->y, b->z, b->m, b->flags, b->js, (int)n, (int)n,
                                 b->stream));
  return build_items(b, -1, 0);
}

int nb_batch_set_graph(nb_batch *b, int32_t enable) {
  if (!b) return NB_E_INVALID;
  b->use_graph = enable != 0;
  drop_graph(b);
  return NB_OK;
}

int nb_batch_step(nb_batch *b, double dt, double R, int32_t nsteps, nb_stats *out) {
  if (!b) return NB_E_INVALID;
  if (!b->items) return bfail(b, NB_E_STATE, "nb_batch_step: nothing loaded");
  if (nsteps < 0 || !std::isfinite(dt) || !std::isfinite(R))
    return bfail(b, NB_E_INVALID, "nb_batch_step: bad arguments");
  NB_BHIP(b, hipSetDevice(b->device));
  auto t0 = std::chrono::steady_clock::now();
  DevCounters before, after;
  NB_BHIP(b, hipMemcpyAsync(&before, b->ctr, sizeof before, hipMemcpyDeviceToHost, b->stream));
  if (nsteps > 0) {
    if (b->use_graph) {
      if (!b->exec || b->graph_steps != nsteps || b->graph_dt != dt || b->graph_R != R) {
        drop_graph(b);
        NB_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int s = 0; s < nsteps && e == hipSuccess; s++) e = enqueue_cycle(b, dt, R);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(b->stream, &g);
        NB_BHIP(b, e);
        NB_BHIP(b, e2);
        b->graph = g;
        NB_BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_steps = nsteps;
        b->graph_dt = dt;
        b->graph_R = R;
      }
      NB_BHIP(b, hipGraphLaunch(b->exec, b->stream));
    } else {
      for (int s = 0; s < nsteps; s++) NB_BHIP(b, enqueue_cycle(b, dt, R));
    }
  }
  NB_BHIP(b, hipMemcpyAsync(&after, b->ctr, sizeof after, hipMemcpyDeviceToHost, b->stream));
  NB_BHIP(b, hipStreamSynchronize(b->stream));
  auto t1 = std::chrono::steady_clock::now();
  const int64_t removed = (int64_t)(after.removed_total - before.removed_total);
  if (out) {
    memset(out, 0, sizeof *out);
    out->steps = nsteps;
    out->n = b->off[b->nsims];
    out->interactions = b->interactions_per_step * nsteps;
    out->removed_total = removed;
    b->n_active -= removed;
    out->n_active = b->n_active;
    out->wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  } else {
    b->n_active -= removed;
  }
  return NB_OK;
}

int nb_batch_forces(nb_batch *b, double *fx, double *fy, double *fz) {
  if (!b) return NB_E_INVALID;
  if (!b->items) return bfail(b, NB_E_STATE, "nb_batch_forces: nothing loaded");
  NB_BHIP(b, hipSetDevice(b->device));
  NB_BHIP(b, launch_batch_force(b));
  const int64_t n = b->n;
  // the per-body accelerations and masses come back to the host, where the force is formed
  // exactly as the combine kernel forms it (m * (G * a))
  std::vector<double> a((size_t)3 * n), mh((size_t)n);
  NB_BHIP(b, hipMemcpyAsync(a.data(), b->acc, a.size() * sizeof(double), hipMemcpyDeviceToHost,
                            b->stream));
  NB_BHIP(b, hipMemcpyAsync(mh.data(), b->m, mh.size() * sizeof(double), hipMemcpyDeviceToHost,
                            b->stream));
  NB_BHIP(b, hipStreamSynchronize(b->stream));
  std::vector<double> h((size_t)3 * n);
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++)
      h[(size_t)c * n + i] = mh[(size_t)i] * (b->p.G * a[(size_t)c * n + i]);
  double *o[3] = {fx, fy, fz};
  for (int k = 0; k < b->nsims; k++)
    for (int64_t q = 0; q < b->off[k + 1] - b->off[k]; q++)
      for (int c = 0; c < 3; c++)
        o[c][b->off[k] + q] = h[(size_t)c * n + (size_t)(b->poff[k] + q)];
  return NB_OK;
}

int nb_batch_store(nb_batch *b, double *x, double *y, double *z, double *vx, double *vy,
                   double *vz, uint8_t *flags) {
  if (!b) return NB_E_INVALID;
  if (!b->items) return bfail(b, NB_E_STATE, "nb_batch_store: nothing loaded");
  NB_BHIP(b, hipSetDevice(b->device));
  const int64_t n = b->n;
  double *dst[6] = {x, y, z, vx, vy, vz};
  const double *src[6] = {b->x, b->y, b->z, b->vx, b->vy, b->vz};
  std::vector<double> h((size_t)n);
  for (int c = 0; c < 6; c++) {
    if (!dst[c]) continue;
    NB_BHIP(b, hipMemcpy(h.data(), src[c], (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < b->nsims; k++)
      for (int64_t q = 0; q < b->off[k + 1] - b->off[k]; q++)
        dst[c][b->off[k] + q] = h[(size_t)(b->poff[k] + q)];
  }
  if (flags) {
    std::vector<uint8_t> hf((size_t)n);
    NB_BHIP(b, hipMemcpy(hf.data(), b->flags, (size_t)n, hipMemcpyDeviceToHost));
    for (int k = 0; k < b->nsims; k++)
      for (int64_t q = 0; q < b->off[k + 1] - b->off[k]; q++)
        flags[b->off[k] + q] = hf[(size_t)(b->poff[k] + q)];
  }
  return NB_OK;
}

int nb_batch_get_shapes(nb_batch *b, int32_t *ipl3) {
  if (!b || !ipl3) return NB_E_INVALID;
  for (int k = 0; k < kBuckets; k++) ipl3[k] = b->bucket_ipl[k];
  return NB_OK;
}

int nb_batch_set_shapes(nb_batch *b, const int32_t *ipl3) {
  if (!b || !ipl3) return NB_E_INVALID;
  for (int k = 0; k < kBuckets; k++)
    if (ipl3[k] != 1 && ipl3[k] != 2 && ipl3[k] != 4)
      return bfail(b, NB_E_INVALID, "nb_batch_set_shapes: IPL must be 1, 2 or 4");
  for (int k = 0; k < kBuckets; k++) b->bucket_ipl[k] = ipl3[k];
  return b->items ? build_items(b, -1, 0) : NB_OK;
}

int nb_batch_get_tuning(nb_batch *b, int32_t *cfg7) {
  if (!b || !cfg7) return NB_E_INVALID;
  cfg7[0] = b->waves;
  for (int k = 0; k < kBuckets; k++) {
    cfg7[1 + 2 * k] = b->bucket_ipl[k];
    cfg7[2 + 2 * k] = b->bucket_tile[k];
  }
  return NB_OK;
}

int nb_batch_set_tuning(nb_batch *b, const int32_t *cfg7) {
  if (!b || !cfg7) return NB_E_INVALID;
  if (cfg7[0] != 1 && cfg7[0] != 2 && cfg7[0] != 4)
    return bfail(b, NB_E_INVALID, "nb_batch_set_tuning: waves must be 1, 2 or 4");
  for (int k = 0; k < kBuckets; k++) {
    const int ipl = cfg7[1 + 2 * k], tile = cfg7[2 + 2 * k];
    if ((ipl != 1 && ipl != 2 && ipl != 4) || (tile != 64 && tile != 128 && tile != 256))
      return bfail(b, NB_E_INVALID, "nb_batch_set_tuning: IPL in {1, 2, 4}, tile in {64, 128, 256}");
  }
  b->waves = cfg7[0];
  for (int k = 0; k < kBuckets; k++) {
    b->bucket_ipl[k] = cfg7[1 + 2 * k];
    b->bucket_tile[k] = cfg7[2 + 2 * k];
  }
  drop_graph(b);
  return b->items ? build_items(b, -1, 0) : NB_OK;
}

// For each waves-per-item choice: every bucket's items alone under each (IPL, tile), then two
// kinds of whole-batch candidate, each timed as a full launch: the per-bucket winners together
// (the per-item dispatch kernel, or the specialised one when they coincide) and every single
// (IPL, tile) for all items (the specialised kernel).  The fastest full launch is kept.
int nb_batch_autotune(nb_batch *b, int32_t reps, int32_t *ipl3) {
  if (!b) return NB_E_INVALID;
  if (!b->items) return bfail(b, NB_E_STATE, "nb_batch_autotune: nothing loaded");
  if (reps < 1) reps = 1;
  NB_BHIP(b, hipSetDevice(b->device));
  hipEvent_t e0, e1;
  NB_BHIP(b, hipEventCreate(&e0));
  NB_BHIP(b, hipEventCreate(&e1));
  // mean ms of `reps` launches of the current items, after one warm-up launch
  auto time_launch = [&](float &ms) -> int {
    NB_BHIP(b, launch_batch_force(b));
    NB_BHIP(b, hipEventRecord(e0, b->stream));
    for (int r = 0; r < reps; r++) NB_BHIP(b, launch_batch_force(b));
    NB_BHIP(b, hipEventRecord(e1, b->stream));
    NB_BHIP(b, hipEventSynchronize(e1));
    NB_BHIP(b, hipEventElapsedTime(&ms, e0, e1));
    ms /= (float)reps;
    return NB_OK;
  };
  float best_total = 1e30f;
  int best_w = b->waves, best_ipl[kBuckets], best_tile[kBuckets];
  for (int k = 0; k < kBuckets; k++) best_ipl[k] = b->bucket_ipl[k], best_tile[k] = b->bucket_tile[k];
  for (int w : {1, 2, 4}) {
    b->waves = w;
    int ipl_w[kBuckets], tile_w[kBuckets];
    for (int bk = 0; bk < kBuckets; bk++) {
      float best = 1e30f;
      ipl_w[bk] = b->bucket_ipl[bk];
      tile_w[bk] = b->bucket_tile[bk];
      for (int ipl : {1, 2, 4})
        for (int tile : {64, 128, 256}) {
          int rc = build_items(b, bk, ipl, tile);
          if (rc) return rc;
          if (b->nitems == 0) break;
          float ms = 0;
          if ((rc = time_launch(ms))) return rc;
          if (ms < best) {
            best = ms;
            ipl_w[bk] = ipl;
            tile_w[bk] = tile;
          }
        }
    }
    // the per-bucket winners together
    for (int k = 0; k < kBuckets; k++) b->bucket_ipl[k] = ipl_w[k], b->bucket_tile[k] = tile_w[k];
    int rc = build_items(b, -1, 0);
    if (rc) return rc;
    float ms = 0;
    if ((rc = time_launch(ms))) return rc;
    if (ms < best_total) {
      best_total = ms;
      best_w = w;
      for (int k = 0; k < kBuckets; k++) best_ipl[k] = ipl_w[k], best_tile[k] = tile_w[k];
    }
    // one shape for all items
    for (int ipl : {1, 2, 4})
      for (int tile : {64, 128, 256}) {
        if ((rc = build_items(b, -1, ipl, tile))) return rc;
        if ((rc = time_launch(ms))) return rc;
        if (ms < best_total) {
          best_total = ms;
          best_w = w;
          for (int k = 0; k < kBuckets; k++) best_ipl[k] = ipl, best_tile[k] = tile;
        }
      }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->waves = best_w;
  for (int k = 0; k < kBuckets; k++) b->bucket_ipl[k] = best_ipl[k], b->bucket_tile[k] = best_tile[k];
  int rc = build_items(b, -1, 0);
  if (rc) return rc;
  if (ipl3)
    for (int k = 0; k < kBuckets; k++) ipl3[k] = b->bucket_ipl[k];
  return NB_OK;
}

}  // extern "C"
