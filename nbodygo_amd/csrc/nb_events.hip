// nb_events.hip -- the reference's subsume and fragmentation events on the device
// (nb_params.events = 1).  Detection (the nb_pair_event dispatch in every pair pass) and the
// resolution of subsume pairs and the start of fragmentation (k_resolve) live with the direct
// path; this file holds fragment(): each cycle every fragmenting body creates up to 101 fragment
// bodies (fragcalc.go:97-117) into a device queue that nb_cycle appends, newest first.
//
// Fragment positions: GetVectorEven (cmd/util/vectorutil.go:32-42) draws x, y, z uniform in
// [-1, 1) until x^2 + y^2 + z^2 <= 1 and returns centre + radius * (x, y, z).  The reference seeds
// math/rand from the clock; here every draw comes from a counter-based generator keyed by the
// fragmentation's id, the fragment's ordinal and the draw index (SplitMix64 finaliser), so a run
// is reproducible and the oracle draws the same numbers.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "nb_internal.hpp"

namespace nb {

__host__ __device__ __forceinline__ uint64_t ev_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Uniform double in [0, 1) with 53 random bits, as rand.Float64 returns.
__device__ __forceinline__ double ev_uniform(uint64_t id, uint64_t ordinal, uint64_t draw) {
  const uint64_t h = ev_mix(ev_mix(id) ^ (ordinal * 0xD1B54A32D192ED03ull) ^
                            (draw * 0xA0761D6478BD642Full));
  return (double)(h >> 11) * 0x1.0p-53;
}

constexpr int kEvMaxDraws = 64;  // rejection rounds (each accepts with probability pi/6)

// Fragments each fragmenting body creates this cycle: min(left, 101), or 0.
__global__ __launch_bounds__(256) void k_frag_count(const uint8_t *flags, const FragState *fst,
                                                    int n, int32_t *e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t f = flags[i];
  int32_t k = 0;
  if ((f & NB_F_EXISTS) && (f & NB_F_FRAGMENTING)) {
    const int64_t left = fst[i].left;
    k = left > 0 ? (int32_t)min(left, (int64_t)kFragPerCycle) : 0;
  }
  e[i] = k;
}

struct EmitArgs {
  const double *vx, *vy, *vz;
  uint8_t *flags;
  FragState *fst;
  const int32_t *e, *off;
  double *fq;
  int64_t fq_cap;
  DevCounters *ctr;
  int n;
};

// fragment (fragcalc.go:97-117) for the fragmenting bodies, slots in body order.  If the queue
// cannot take this cycle's fragments, none are made (the bodies keep their counts) and the
// overflow is flagged.
__global__ __launch_bounds__(256) void k_frag_emit(EmitArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint8_t f = a.flags[i];
  if (!((f & NB_F_EXISTS) && (f & NB_F_FRAGMENTING))) return;
  const int64_t total = (int64_t)a.off[a.n - 1] + a.e[a.n - 1];
  const int64_t pend = a.ctr->frag_pending;
  const bool fits = pend + total <= a.fq_cap;
  FragState st = a.fst[i];
  const int k = a.e[i];
  if (k > 0) {
    if (!fits) {
      a.ctr->frag_overflow = 1u;
      return;
    }
    const double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
    const double rad = st.radius * .9;
    const size_t cap = (size_t)a.fq_cap;
    for (int q = 0; q < k; q++) {
      const uint64_t ordinal = (uint64_t)(st.made + q);
      double x = 0, y = 0, z = 0;
      for (int d = 0; d < kEvMaxDraws; d++) {  // for d := 2; d > 1 (vectorutil.go:35-40)
        x = ev_uniform((uint64_t)st.id, ordinal, 3 * d) * 2 - 1;
        y = ev_uniform((uint64_t)st.id, ordinal, 3 * d + 1) * 2 - 1;
        z = ev_uniform((uint64_t)st.id, ordinal, 3 * d + 2) * 2 - 1;
        if (!(x * x + y * y + z * z > 1)) break;
      }
      const size_t slot = (size_t)(pend + a.off[i] + q);
      a.fq[slot] = (x * rad) + st.px;
      a.fq[cap + slot] = (y * rad) + st.py;
      a.fq[2 * cap + slot] = (z * rad) + st.pz;
      a.fq[3 * cap + slot] = vx;
      a.fq[4 * cap + slot] = vy;
      a.fq[5 * cap + slot] = vz;
      a.fq[6 * cap + slot] = st.mass;
      a.fq[7 * cap + slot] = kFragRadius;
    }
    st.left -= k;
    st.made += k;
    a.fst[i] = st;
  }
  if (st.left <= 0) {  // b.Exists = false, the mass kept (not SetNotExists)
    a.flags[i] = (uint8_t)(f & ~NB_F_EXISTS);
    atomicAdd((unsigned long long *)&a.ctr->frag_finished, 1ull);
  }
}

__global__ void k_frag_commit(const int32_t *e, const int32_t *off, int n, int64_t cap,
                              DevCounters *ctr) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t total = (int64_t)off[n - 1] + e[n - 1];
  if ((int64_t)ctr->frag_pending + total <= cap) {
    ctr->frag_pending += (uint32_t)total;
    ctr->frag_emitted += (uint64_t)total;
  }
}

// The queue, newest first, into the new arrays at [at, at + nq): each fragment's ordinal in the
// queue is reversed, as Cycle walks the event list from its front.
struct AppendArgs {
  const double *fq;
  int64_t fq_cap, nq, at;
  double *dst[10];
  int32_t *beh;
  uint8_t *flags;
  double *fstep;
  FragState *fst;
  double R;
};

__global__ __launch_bounds__(256) void k_append_frags(AppendArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nq) return;
  const size_t src = (size_t)(a.nq - 1 - t), cap = (size_t)a.fq_cap;
  const size_t p = (size_t)(a.at + t);
  for (int k = 0; k < 8; k++) a.dst[k][p] = a.fq[k * cap + src];  // x y z vx vy vz m r
  a.dst[8][p] = a.R;   // Cycle: arr[j].r = R (body_collection.go:276)
  a.dst[9][p] = 0.0;   // FragFactor 0
  a.beh[p] = NB_CB_ELASTIC;
  a.flags[p] = NB_F_EXISTS;
  if (a.fstep) a.fstep[p] = 0.0;  // FragStep 0
  if (a.fst) {
    FragState z{};
    a.fst[p] = z;
  }
}

__global__ void k_frag_drain(DevCounters *ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctr->frag_pending = 0;
    ctr->frag_overflow = 0;
  }
}

// ---- host ----------------------------------------------------------------------------------------
static int nb256e(int64_t n) { return (int)std::max<int64_t>(1, (n + 255) / 256); }

int events_alloc(Ctx *c, int64_t cap) {
  if (c->fstep) (void)hipFree(c->fstep);
  if (c->fst) (void)hipFree(c->fst);
  c->fstep = nullptr;
  c->fst = nullptr;
  NB_HIP(c, hipMalloc((void **)&c->fstep, (size_t)cap * sizeof(double)));
  NB_HIP(c, hipMalloc((void **)&c->fst, (size_t)cap * sizeof(FragState)));
  NB_HIP(c, hipMemsetAsync(c->fstep, 0, (size_t)cap * sizeof(double), c->stream));
  NB_HIP(c, hipMemsetAsync(c->fst, 0, (size_t)cap * sizeof(FragState), c->stream));
  return NB_OK;
}

// Room for the fragments of the next cycles: what is pending plus 101 per fragmenting body (and
// some more for bodies that start fragmenting), at least 4096.  Grows, keeping the contents.
int events_ensure_queue(Ctx *c) {
  const int64_t want = std::max<int64_t>(
      4096, c->evs.fragments_pending + (int64_t)kFragPerCycle * (c->evs.fragmenting + 64) * 4);
  if (c->fq && c->fq_cap >= want) return NB_OK;
  const int64_t cap = std::max<int64_t>(want, 2 * c->fq_cap);
  double *q = nullptr;
  NB_HIP(c, hipMalloc((void **)&q, (size_t)cap * 8 * sizeof(double)));
  if (c->fq && c->evs.fragments_pending > 0) {
    for (int k = 0; k < 8; k++)
      NB_HIP(c, hipMemcpyAsync(q + (size_t)k * cap, c->fq + (size_t)k * c->fq_cap,
                               (size_t)c->evs.fragments_pending * sizeof(double),
                               hipMemcpyDeviceToDevice, c->stream));
  }
  NB_HIP(c, hipStreamSynchronize(c->stream));
  if (c->fq) (void)hipFree(c->fq);
  c->fq = q;
  c->fq_cap = cap;
  return NB_OK;
}

void events_free(Ctx *c) {
  if (c->fstep) (void)hipFree(c->fstep);
  if (c->fst) (void)hipFree(c->fst);
  if (c->fq) (void)hipFree(c->fq);
  c->fstep = nullptr;
  c->fst = nullptr;
  c->fq = nullptr;
  c->fq_cap = 0;
}

void events_refresh_host(Ctx *c, const DevCounters &d) {
  c->evs.subsumed_total = (int64_t)d.subsumed_total;
  c->evs.fragmentations = (int64_t)d.frag_started;
  c->evs.fragments_emitted = (int64_t)d.frag_emitted;
  c->evs.fragments_pending = (int64_t)d.frag_pending;
  c->evs.fragmenting = (int64_t)d.frag_started - (int64_t)d.frag_finished;
  c->evs.queue_overflow = d.frag_overflow ? 1 : 0;
}

// One cycle of fragment(): counts, their prefix sum, the fragments, the queue total.
int events_emit(Ctx *c) {
  const int n = (int)c->n;
  if (n == 0 || !c->fst) return NB_OK;
  hipStream_t st = c->stream;
  int32_t *e = reinterpret_cast<int32_t *>(c->diag);  // scratch: 2 n ints fit in diag
  int32_t *off = e + n;
  hipLaunchKernelGGL(k_frag_count, dim3(nb256e(n)), dim3(256), 0, st, c->flags, c->fst, n, e);
  size_t bytes = 0;
  NB_HIP(c, rocprim::exclusive_scan(nullptr, bytes, e, off, 0, (size_t)n,
                                    rocprim::plus<int32_t>(), st));
  if (bytes > c->ev_scan_bytes) {
    if (c->ev_scan_tmp) (void)hipFree(c->ev_scan_tmp);
    c->ev_scan_tmp = nullptr;
    NB_HIP(c, hipMalloc(&c->ev_scan_tmp, bytes));
    c->ev_scan_bytes = bytes;
  }
  NB_HIP(c, rocprim::exclusive_scan(c->ev_scan_tmp, bytes, e, off, 0, (size_t)n,
                                    rocprim::plus<int32_t>(), st));
  EmitArgs a;
  a.vx = c->vx;
  a.vy = c->vy;
  a.vz = c->vz;
  a.flags = c->flags;
  a.fst = c->fst;
  a.e = e;
  a.off = off;
  a.fq = c->fq;
  a.fq_cap = c->fq_cap;
  a.ctr = c->ctr;
  a.n = n;
  hipLaunchKernelGGL(k_frag_emit, dim3(nb256e(n)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_frag_commit, dim3(1), dim3(64), 0, st, e, off, n, c->fq_cap, c->ctr);
  NB_HIP(c, hipGetLastError());
  return NB_OK;
}

hipError_t launch_append_frags(Ctx *c, int64_t nq, double *const dst[10], int32_t *beh,
                               uint8_t *flags, double *fstep, FragState *fst, int64_t at,
                               double R, hipStream_t st) {
  if (nq <= 0) return hipSuccess;
  AppendArgs a;
  a.fq = c->fq;
  a.fq_cap = c->fq_cap;
  a.nq = nq;
  a.at = at;
  for (int k = 0; k < 10; k++) a.dst[k] = dst[k];
  a.beh = beh;
  a.flags = flags;
  a.fstep = fstep;
  a.fst = fst;
  a.R = R;
  hipLaunchKernelGGL(k_append_frags, dim3(nb256e(nq)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_frag_drain, dim3(1), dim3(64), 0, st, c->ctr);
  return hipGetLastError();
}

}  // namespace nb

using namespace nb;

extern "C" {

int nb_set_frag_steps(nb_ctx *c, const double *frag_step) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded) return fail_msg(c, NB_E_STATE, "nb_set_frag_steps: no bodies loaded");
  if (!c->p.events)
    return fail_msg(c, NB_E_UNSUPPORTED, "nb_set_frag_steps: the context has events off");
  if (c->n == 0) return NB_OK;
  NB_HIP(c, hipSetDevice(c->device));
  if (frag_step)
    NB_HIP(c, hipMemcpyAsync(c->fstep, frag_step, (size_t)c->n * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  else
    NB_HIP(c, hipMemsetAsync(c->fstep, 0, (size_t)c->n * sizeof(double), c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  return NB_OK;
}

int nb_get_event_stats(nb_ctx *c, nb_event_stats *out) {
  if (!c || !out) return NB_E_INVALID;
  *out = c->evs;
  return NB_OK;
}

}  // extern "C"
