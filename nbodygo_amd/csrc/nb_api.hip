// nb_api.hip -- C-ABI of libnbgpu.so: context, body transfer, the cycle driver and the shape
// autotuner.  See include/nbody_gpu.h for the contract of each entry point.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include <rocprim/rocprim.hpp>

#include "nb_internal.hpp"

using nb::Ctx;

namespace nb {

static thread_local std::string g_create_err;

bool hip_ok(Ctx *c, hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  if (c) c->err = buf;
  else g_create_err = buf;
  return false;
}

static int fail(Ctx *c, int code, const std::string &msg) {
  c->err = msg;
  return code;
}

int fail_msg(Ctx *c, int code, const char *msg) { return fail(c, code, msg); }

void set_global_error(const char *msg) { g_create_err = msg; }

template <class T>
static hipError_t dalloc(T **p, size_t count) {
  if (*p) {
    hipError_t e = hipFree(*p);
    *p = nullptr;
    if (e != hipSuccess) return e;
  }
  if (count == 0) count = 1;
  return hipMalloc((void **)p, count * sizeof(T));
}

static void free_spare(Ctx *c) {
  auto &S = c->spare;
  for (auto *&q : S.a)
    if (q) (void)hipFree(q), q = nullptr;
  void *ps[] = {S.beh, S.flags, S.collided, S.fstep, S.fst, c->cyc_keep, c->cyc_pos, c->cyc_tmp};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  S.beh = nullptr;
  S.flags = S.collided = nullptr;
  S.fstep = nullptr;
  S.fst = nullptr;
  S.cap = 0;
  c->cyc_keep = c->cyc_pos = nullptr;
  c->cyc_tmp = nullptr;
  c->cyc_tmp_bytes = 0;
  c->cyc_cap = 0;
}

static void free_all(Ctx *c) {
  free_spare(c);
  void *ptrs[] = {c->x,    c->y,      c->z,     c->vx,   c->vy,        c->vz,   c->m,
                  c->r,    c->rest,   c->frag,  c->beh,  c->flags,     c->collided,
                  c->js,   c->rmax_tile, c->acc, c->fcnt, c->flist, c->recs, c->pairs, c->ctr,
                  c->diag, c->pairs_tmp, c->pairs_sw, c->res_words};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  c->x = c->y = c->z = c->vx = c->vy = c->vz = c->m = c->r = c->rest = c->frag = nullptr;
  c->beh = nullptr;
  c->flags = c->collided = nullptr;
  c->js = nullptr;
  c->rmax_tile = nullptr;
  c->acc = nullptr;
  c->fcnt = c->flist = nullptr;
  c->recs = nullptr;
  c->pairs = c->pairs_tmp = c->pairs_sw = nullptr;
  c->res_words = nullptr;
  c->ctr = nullptr;
  c->diag = nullptr;
}

static bool detect_needed(const Ctx *c) {
  return c->p.collisions || c->p.overlap_cutoff || c->p.eps == 0.0;
}
static bool drop_needed(const Ctx *c) { return c->p.overlap_cutoff || c->p.eps == 0.0; }

static int tiles_of(const Ctx *c) { return (int)(c->npad / c->shape.tile_j); }

// Default launch shape: the one nb_autotune picks at 65,536 and 262,144 bodies (LDS-staged
// sources, one target per lane, 512-source tiles), with enough source slices for about 16
// workgroups per CU on 256 CUs.  npad is a multiple of 1024, so every field divides it.
static ForceShape default_shape(int64_t npad) {
  ForceShape s;
  s.ipl = 1;
  s.block = 256;
  s.tile_j = 512;
  s.variant = 2;
  int64_t bx = npad / (s.block * s.ipl);
  int64_t tiles = npad / s.tile_j;
  int64_t want = (4096 + bx - 1) / bx;
  s.jsplit = (int)std::max<int64_t>(1, std::min<int64_t>({want, (int64_t)kMaxSplit, tiles}));
  return s;
}

static bool shape_valid(const ForceShape &s, int64_t npad) {
  if (s.ipl != 1 && s.ipl != 2 && s.ipl != 4) return false;
  if (s.block < 64 || s.block > 256 || (s.block & 63)) return false;
  if (s.tile_j < 64 || s.tile_j > 1024 || (s.tile_j & 63) || npad % s.tile_j) return false;
  if (npad % (s.block * s.ipl)) return false;
  if (s.jsplit < 1 || s.jsplit > kMaxSplit || s.jsplit > npad / s.tile_j) return false;
  if (s.variant != 1 && s.variant != 2) return false;
  return true;
}

// (Re)size the per-slice and per-tile buffers for the current shape.
static int prepare_shape(Ctx *c) {
  if (!shape_valid(c->shape, c->npad))
    return fail(c, NB_E_INVALID, "invalid FP32 force kernel shape for this body count");
  const int split = c->p.precision == NB_PREC_FP64 ? 1 : c->shape.jsplit;
  if (split > c->acc_split || !c->acc) {
    const int alloc_split = std::max(split, 1);
    NB_HIP(c, dalloc(&c->acc, (size_t)alloc_split * 3 * c->cap));
    NB_HIP(c, dalloc(&c->fcnt, (size_t)alloc_split * c->cap));
    NB_HIP(c, dalloc(&c->flist, (size_t)alloc_split * c->cap * kFlagSlots));
    NB_HIP(c, dalloc(&c->recs, (size_t)alloc_split * c->cap));
    c->acc_split = alloc_split;
  }
  const int tiles = tiles_of(c);
  if (c->rmax_tile_j != c->shape.tile_j) {
    NB_HIP(c, dalloc(&c->rmax_tile, (size_t)(c->cap / 64)));
    c->rmax_tile_j = c->shape.tile_j;
  }
  NB_HIP(c, launch_rmax_tile(c->r, c->flags, (int)c->n, c->shape.tile_j, tiles, c->rmax_tile,
                             c->stream));
  return NB_OK;
}

// Keeps a tuned shape that still fits the padded count; else the default with the params'
// pinned fields.
static void choose_shape(Ctx *c) {
  if (c->shape_tuned && shape_valid(c->shape, c->npad)) return;
  c->shape = default_shape(c->npad);
  c->shape_tuned = false;
  const nb_params &p = c->p;
  if (p.ipl) c->shape.ipl = p.ipl;
  if (p.tile_j) c->shape.tile_j = p.tile_j;
  if (p.jsplit) c->shape.jsplit = p.jsplit;
  if (p.block) c->shape.block = p.block;
  if (p.variant) c->shape.variant = p.variant;
}

// Pair buffer for n bodies: a power of two so the global bitonic network can pad in place.
static int size_pairs(Ctx *c, int64_t n) {
  const int64_t pcap =
      c->p.pair_capacity > 0 ? c->p.pair_capacity : std::max<int64_t>(4 * n, 65536);
  int64_t pow2 = 1;
  while (pow2 < pcap) pow2 <<= 1;
  if (pow2 > c->pair_cap_pow2) {
    NB_HIP(c, dalloc(&c->pairs, (size_t)pow2));
    NB_HIP(c, dalloc(&c->pairs_tmp, (size_t)pow2));
    NB_HIP(c, dalloc(&c->pairs_sw, (size_t)pow2));
    NB_HIP(c, dalloc(&c->res_words, (size_t)pow2 * 5));
    c->pair_cap_pow2 = pow2;
  }
  c->pair_cap = pcap;
  return NB_OK;
}

ForceArgs make_force_args(Ctx *c) {
  ForceArgs a;
  a.js = c->js;
  a.rmax_tile = c->rmax_tile;
  a.r = c->r;
  a.acc = c->acc;
  a.fcnt = c->fcnt;
  a.flist = c->flist;
  a.npad = (int)c->npad;
  a.n = (int)c->n;
  a.tile_j = c->shape.tile_j;
  a.tiles = tiles_of(c);
  a.tiles_per_slice = (a.tiles + c->shape.jsplit - 1) / c->shape.jsplit;
  const double e2 = c->p.eps * c->p.eps;
  a.eps2 = (float)e2;
  a.detect = detect_needed(c) ? 1 : 0;
  a.drop_flagged = drop_needed(c) ? 1 : 0;
  a.recs = c->recs;
  a.rec_cap = (int64_t)c->acc_split * c->cap;
  a.ctr = c->ctr;
  a.run_flag = nullptr;
  return a;
}

// Force phase of one cycle (force + detection + sorted pair list).
static int force_phase(Ctx *c, bool timed, int ev_slot) {
  if (c->sur_on) return sur_force_phase(c, timed, ev_slot);
  if (c->bh_on) return bh_force_phase(c, timed, ev_slot);
  NB_HIP(c, launch_zero_counters(c->ctr, 0, c->stream));
  if (c->p.precision == NB_PREC_FP64) {
    Force64Args a;
    a.x = c->x; a.y = c->y; a.z = c->z; a.m = c->m; a.r = c->r;
    a.beh = c->beh;
    a.flags = c->flags;
    a.acc = c->acc;
    a.pairs = c->pairs;
    a.ctr = c->ctr;
    a.pair_cap = c->pair_cap;
    a.n = (int)c->n;
    a.npad = (int)c->npad;
    a.G = c->p.G;
    a.eps = c->p.eps;
    a.cutoff = c->p.overlap_cutoff;
    a.collisions = c->p.collisions;
    a.events = c->p.events;
    if (timed) NB_HIP(c, hipEventRecord(c->ev[2 * ev_slot], c->stream));
    NB_HIP(c, launch_force_f64(a, c->stream));
    if (timed) NB_HIP(c, hipEventRecord(c->ev[2 * ev_slot + 1], c->stream));
  } else {
    ForceArgs a = make_force_args(c);
    if (timed) NB_HIP(c, hipEventRecord(c->ev[2 * ev_slot], c->stream));
    NB_HIP(c, launch_force_f32(c->shape, a, c->stream));
    if (timed) NB_HIP(c, hipEventRecord(c->ev[2 * ev_slot + 1], c->stream));
    if (a.detect) {
      FixupArgs f;
      f.x = c->x; f.y = c->y; f.z = c->z; f.m = c->m; f.r = c->r;
      f.beh = c->beh;
      f.flags = c->flags;
      f.acc = c->acc;
      f.fcnt = c->fcnt;
      f.flist = c->flist;
      f.recs = c->recs;
      f.rec_cap = a.rec_cap;
      f.pairs = c->pairs;
      f.ctr = c->ctr;
      f.pair_cap = c->pair_cap;
      f.npad = a.npad;
      f.n = a.n;
      f.tiles = a.tiles;
      f.tiles_per_slice = a.tiles_per_slice;
      f.tile_j = a.tile_j;
      f.split = c->shape.jsplit;
      f.eps2 = c->p.eps * c->p.eps;
      f.recompute = a.drop_flagged;
      f.cutoff = c->p.overlap_cutoff;
      f.collisions = c->p.collisions;
      f.events = c->p.events;
      NB_HIP(c, launch_fixup(f, c->stream));
    }
  }
  if (c->p.collisions) return sort_pairs(c);
  return NB_OK;
}

int sort_pairs(Ctx *c) {
  NB_HIP(c, launch_sort_pairs(c->pairs, c->pairs_tmp, c->pair_cap, c->ctr, c->stream));
  return NB_OK;
}

static int split_now(const Ctx *c) {
  return (c->p.precision == NB_PREC_FP64 || c->sur_on) ? 1 : c->shape.jsplit;
}

static int64_t count_active(int64_t n, const uint8_t *flags) {
  int64_t a = 0;
  for (int64_t i = 0; i < n; i++)
    if ((flags[i] & NB_F_EXISTS) && !(flags[i] & NB_F_FRAGMENTING)) a++;
  return a;
}

}  // namespace nb

using namespace nb;


extern "C" {

int nb_params_init(nb_params *p) {
  if (!p) return NB_E_INVALID;
  memset(p, 0, sizeof *p);
  p->G = 6.673e-11;
  p->eps = 0.0;
  p->overlap_cutoff = 1;
  p->collisions = 1;
  p->precision = NB_PREC_FP32;
  p->device = 0;
  return NB_OK;
}

const char *nb_last_error(const nb_ctx *ctx) {
  if (!ctx) return g_create_err.c_str();
  return ctx->err.c_str();
}

int nb_create(const nb_params *p, nb_ctx **out) {
  if (!out) return NB_E_INVALID;
  *out = nullptr;
  if (!p) {
    g_create_err = "nb_create: params is NULL";
    return NB_E_INVALID;
  }
  if (!(p->eps >= 0.0) || !std::isfinite(p->eps) || !std::isfinite(p->G)) {
    g_create_err = "nb_create: eps must be finite and >= 0, G finite";
    return NB_E_INVALID;
  }
  if (p->events != 0 && p->events != 1) {
    g_create_err = "nb_create: events must be 0 or 1";
    return NB_E_INVALID;
  }
  if (p->precision != NB_PREC_FP32 && p->precision != NB_PREC_FP64) {
    g_create_err = "nb_create: unknown precision mode";
    return NB_E_INVALID;
  }
  if (p->resolve != 0 && p->resolve != 1) {
    g_create_err = "nb_create: resolve must be 0 or 1";
    return NB_E_INVALID;
  }
  if (p->precision == NB_PREC_FP32 && p->eps > 0.0 && (float)(p->eps * p->eps) < 1e-30f) {
    g_create_err = "nb_create: eps^2 underflows FP32; use eps = 0 or NB_PREC_FP64";
    return NB_E_INVALID;
  }
  int ndev = 0;
  if (!hip_ok(nullptr, hipGetDeviceCount(&ndev), "hipGetDeviceCount")) return NB_E_HIP;
  if (p->device < 0 || p->device >= ndev) {
    g_create_err = "nb_create: no such HIP device";
    return NB_E_INVALID;
  }
  nb_ctx *c = new (std::nothrow) nb_ctx();
  if (!c) {
    g_create_err = "nb_create: out of host memory";
    return NB_E_NOMEM;
  }
  c->p = *p;
  c->device = p->device;
  if (!hip_ok(nullptr, hipSetDevice(c->device), "hipSetDevice") ||
      !hip_ok(nullptr, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking),
              "hipStreamCreate") ||
      !hip_ok(nullptr, hipMalloc((void **)&c->ctr, sizeof(DevCounters)), "hipMalloc") ||
      !hip_ok(nullptr, hipMalloc((void **)&c->small, 8 * sizeof(unsigned long long)),
              "hipMalloc")) {
    if (c->ctr) (void)hipFree(c->ctr);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return NB_E_HIP;
  }
  *out = c;
  return NB_OK;
}

void nb_destroy(nb_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  sur_free(c->sur);
  c->sur = nullptr;
  bh_free(c->bh);
  c->bh = nullptr;
  events_free(c);
  if (c->ev_scan_tmp) (void)hipFree(c->ev_scan_tmp);
  c->ev_scan_tmp = nullptr;
  free_all(c);
  if (c->small) (void)hipFree(c->small);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->ev_res) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int nb_load(nb_ctx *c, int64_t n, const double *x, const double *y, const double *z,
            const double *vx, const double *vy, const double *vz, const double *m,
            const double *r, const int32_t *beh, const uint8_t *flags, const double *rest,
            const double *frag) {
  if (!c) return NB_E_INVALID;
  if (n < 0 || n > (int64_t)1 << 30)
    return fail(c, NB_E_INVALID, "nb_load: n out of range [0, 2^30]");
  if (n > 0 && (!x || !y || !z || !vx || !vy || !vz || !m || !r || !beh || !flags))
    return fail(c, NB_E_INVALID, "nb_load: NULL array");
  NB_HIP(c, hipSetDevice(c->device));
  const int64_t npad = std::max<int64_t>(round_up(n, kPad), kPad);
  if (npad > c->cap) {
    free_all(c);
    c->cap = npad;
    c->acc_split = 0;
    c->rmax_tile_j = 0;
    c->pair_cap_pow2 = 0;
    NB_HIP(c, dalloc(&c->x, c->cap));
    NB_HIP(c, dalloc(&c->y, c->cap));
    NB_HIP(c, dalloc(&c->z, c->cap));
    NB_HIP(c, dalloc(&c->vx, c->cap));
    NB_HIP(c, dalloc(&c->vy, c->cap));
    NB_HIP(c, dalloc(&c->vz, c->cap));
    NB_HIP(c, dalloc(&c->m, c->cap));
    NB_HIP(c, dalloc(&c->r, c->cap));
    NB_HIP(c, dalloc(&c->rest, c->cap));
    NB_HIP(c, dalloc(&c->frag, c->cap));
    NB_HIP(c, dalloc(&c->beh, c->cap));
    NB_HIP(c, dalloc(&c->flags, c->cap));
    NB_HIP(c, dalloc(&c->collided, c->cap));
    NB_HIP(c, dalloc(&c->js, c->cap));
    NB_HIP(c, hipMalloc((void **)&c->ctr, sizeof(DevCounters)));
    c->diag_len = c->cap * 8 + 64;
    NB_HIP(c, dalloc(&c->diag, c->diag_len));
  }
  {
    const int rc = size_pairs(c, n);
    if (rc) return rc;
  }
  c->n = n;
  c->npad = npad;
  hipStream_t s = c->stream;
  const size_t db = (size_t)n * sizeof(double);
  double *dsts[8] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m, c->r};
  const double *srcs[8] = {x, y, z, vx, vy, vz, m, r};
  for (int k = 0; k < 8; k++) {
    NB_HIP(c, hipMemsetAsync(dsts[k], 0, (size_t)c->cap * sizeof(double), s));
    if (n) NB_HIP(c, hipMemcpyAsync(dsts[k], srcs[k], db, hipMemcpyHostToDevice, s));
  }
  NB_HIP(c, hipMemsetAsync(c->beh, 0, (size_t)c->cap * sizeof(int32_t), s));
  NB_HIP(c, hipMemsetAsync(c->flags, 0, (size_t)c->cap, s));
  NB_HIP(c, hipMemsetAsync(c->collided, 0, (size_t)c->cap, s));
  NB_HIP(c, hipMemsetAsync(c->frag, 0, (size_t)c->cap * sizeof(double), s));
  if (n) {
    NB_HIP(c, hipMemcpyAsync(c->beh, beh, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    NB_HIP(c, hipMemcpyAsync(c->flags, flags, (size_t)n, hipMemcpyHostToDevice, s));
    if (frag) NB_HIP(c, hipMemcpyAsync(c->frag, frag, db, hipMemcpyHostToDevice, s));
  }
  if (rest) {
    if (n) NB_HIP(c, hipMemcpyAsync(c->rest, rest, db, hipMemcpyHostToDevice, s));
  } else {
    std::vector<double> ones((size_t)std::max<int64_t>(n, 1), 1.0);
    if (n) NB_HIP(c, hipMemcpyAsync(c->rest, ones.data(), db, hipMemcpyHostToDevice, s));
    NB_HIP(c, hipStreamSynchronize(s));
  }
  NB_HIP(c, launch_write_stream(c->x, c->y, c->z, c->m, c->flags, c->js, (int)n, (int)npad, s));
  choose_shape(c);
  int rc = prepare_shape(c);
  if (rc) return rc;
  NB_HIP(c, launch_zero_counters(c->ctr, 1, s));
  if (c->p.events) {
    int rc2 = events_alloc(c, c->cap);
    if (rc2) return rc2;
    c->evs = nb_event_stats{};
    rc2 = events_ensure_queue(c);
    if (rc2) return rc2;
  }
  NB_HIP(c, hipStreamSynchronize(s));
  c->n_active = n ? count_active(n, flags) : 0;
  c->rmax = 0;
  for (int64_t i = 0; i < n; i++)
    if ((flags[i] & NB_F_EXISTS) && !(flags[i] & NB_F_FRAGMENTING) && r[i] > c->rmax)
      c->rmax = r[i];
  c->loaded = true;
  return NB_OK;
}

int nb_step(nb_ctx *c, double dt, double R, int32_t nsteps, nb_stats *out) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_step: no bodies loaded");
  if (nsteps < 0) return fail(c, NB_E_INVALID, "nb_step: nsteps < 0");
  if (!std::isfinite(dt) || !std::isfinite(R))
    return fail(c, NB_E_INVALID, "nb_step: dt and R must be finite");
  NB_HIP(c, hipSetDevice(c->device));
  auto t0 = std::chrono::steady_clock::now();
  DevCounters before;
  NB_HIP(c, hipMemcpyAsync(&before, c->ctr, sizeof before, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  const bool timed = c->timing && nsteps > 0 && c->n > 0;
  if (timed) {
    while ((int)c->ev.size() < 2 * nsteps) {
      hipEvent_t e;
      NB_HIP(c, hipEventCreate(&e));
      c->ev.push_back(e);
    }
    while (c->p.collisions && (int)c->ev_res.size() < 2 * nsteps) {
      hipEvent_t e;
      NB_HIP(c, hipEventCreate(&e));
      c->ev_res.push_back(e);
    }
  }
  const int64_t na0 = c->n_active;
  if (c->p.events) {
    int rc = events_ensure_queue(c);
    if (rc) return rc;
  }
  for (int s = 0; s < nsteps && c->n > 0; s++) {
    if (c->p.events) {  // fragment() of the fragmenting bodies, in Compute (body.go:152-154)
      int rc = events_emit(c);
      if (rc) return rc;
    }
    int rc = force_phase(c, timed, s);
    if (rc) return rc;
    if (c->p.collisions) {
      ResolveArgs ra;
      ra.x = c->x; ra.y = c->y; ra.z = c->z; ra.vx = c->vx; ra.vy = c->vy; ra.vz = c->vz;
      ra.m = c->m; ra.r = c->r; ra.rest = c->rest; ra.frag = c->frag;
      ra.beh = c->beh;
      ra.flags = c->flags;
      ra.collided = c->collided;
      ra.pairs = c->pairs;
      ra.ctr = c->ctr;
      ra.pair_cap = c->pair_cap;
      ra.events = c->p.events;
      ra.fstep = c->fstep;
      ra.fst = c->fst;
      ra.acc = c->acc;
      ra.split = split_now(c);
      ra.npad = (int)c->npad;
      ra.pairs_sw = c->pairs_sw;
      ra.sort_tmp = c->pairs_tmp;
      ra.succ = c->res_words;
      ra.pending = c->res_words + 2 * c->pair_cap_pow2;
      ra.list_a = c->res_words + 3 * c->pair_cap_pow2;
      ra.list_b = c->res_words + 4 * c->pair_cap_pow2;
      ra.serial = c->p.resolve == 1 ? 1 : 0;
      if (timed) NB_HIP(c, hipEventRecord(c->ev_res[2 * s], c->stream));
      NB_HIP(c, launch_resolve(ra, c->stream));
      if (timed) NB_HIP(c, hipEventRecord(c->ev_res[2 * s + 1], c->stream));
    }
    UpdateArgs u;
    u.x = c->x; u.y = c->y; u.z = c->z; u.vx = c->vx; u.vy = c->vy; u.vz = c->vz;
    u.m = c->m;
    u.rest = c->rest;
    u.flags = c->flags;
    u.collided = c->collided;
    u.acc = c->acc;
    u.split = split_now(c);
    u.npad = (int)c->npad;
    u.n = (int)c->n;
    u.G = c->p.G;
    u.dt = dt;
    u.R = R;
    u.force_mode = c->p.precision == NB_PREC_FP64 ? 1 : 0;
    u.js = c->js;
    u.ctr = c->ctr;
    u.fst = c->p.events ? c->fst : nullptr;
    NB_HIP(c, launch_update(u, c->stream));
  }
  DevCounters after;
  NB_HIP(c, hipMemcpyAsync(&after, c->ctr, sizeof after, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  auto t1 = std::chrono::steady_clock::now();
  double force_ms = 0, resolve_ms = 0;
  if (timed) {
    for (int s = 0; s < nsteps; s++) {
      float ms = 0;
      NB_HIP(c, hipEventElapsedTime(&ms, c->ev[2 * s], c->ev[2 * s + 1]));
      force_ms += ms;
      if (c->p.collisions) {
        NB_HIP(c, hipEventElapsedTime(&ms, c->ev_res[2 * s], c->ev_res[2 * s + 1]));
        resolve_ms += ms;
      }
    }
  }
  const int64_t removed = (int64_t)(after.removed_total - before.removed_total) +
                          (int64_t)(after.inactive_events - before.inactive_events);
  c->n_active = std::max<int64_t>(0, c->n_active - removed);
  if (c->p.events) events_refresh_host(c, after);
  if (out) {
    memset(out, 0, sizeof *out);
    out->steps = nsteps;
    out->n = c->n;
    out->n_active = c->n_active;
    out->interactions = (double)nsteps * (double)na0 * (double)std::max<int64_t>(na0 - 1, 0);
    out->pairs_total = (int64_t)(after.pairs_total - before.pairs_total);
    out->pairs_last = c->p.collisions ? (int64_t)after.pair_count : 0;
    out->resolved_total = (int64_t)(after.resolved_total - before.resolved_total);
    out->fragment_requests = (int64_t)(after.fragment_requests - before.fragment_requests);
    out->removed_total = removed;
    out->fixup_tiles = (int64_t)(after.fixup_tiles - before.fixup_tiles);
    out->pair_overflow = after.pair_overflow && !before.pair_overflow ? 1 : (int)after.pair_overflow;
    out->wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    out->force_ms = force_ms;
    out->force_launches = timed ? nsteps : 0;
    out->resolve_ms = resolve_ms;
  }
  if (after.pair_overflow)
    return fail(c, NB_E_PAIR_OVERFLOW,
                "collision pairs exceeded pair_capacity; pairs beyond capacity were dropped");
  return NB_OK;
}

int nb_compute_forces(nb_ctx *c, double *fx, double *fy, double *fz) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_compute_forces: no bodies loaded");
  if (c->n && (!fx || !fy || !fz)) return fail(c, NB_E_INVALID, "nb_compute_forces: NULL output");
  if (c->n == 0) return NB_OK;
  NB_HIP(c, hipSetDevice(c->device));
  int rc = force_phase(c, false, 0);
  if (rc) return rc;
  double *tmp = c->diag;  // 3 n doubles of the context's scratch
  NB_HIP(c, launch_combine_force(c->acc, split_now(c), (int)c->npad, (int)c->n, c->m, c->p.G,
                                 c->p.precision == NB_PREC_FP64 ? 1 : 0, tmp, c->stream));
  const size_t db = (size_t)c->n * sizeof(double);
  NB_HIP(c, hipMemcpyAsync(fx, tmp, db, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipMemcpyAsync(fy, tmp + c->n, db, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipMemcpyAsync(fz, tmp + 2 * c->n, db, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  DevCounters ctr;
  NB_HIP(c, hipMemcpy(&ctr, c->ctr, sizeof ctr, hipMemcpyDeviceToHost));
  if (ctr.pair_overflow)
    return fail(c, NB_E_PAIR_OVERFLOW, "collision pairs exceeded pair_capacity");
  return NB_OK;
}

int nb_accel_fp64(nb_ctx *c, const int64_t *idx, int64_t nidx, double *ax, double *ay,
                  double *az) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_accel_fp64: no bodies loaded");
  if (nidx < 0 || (nidx && (!idx || !ax || !ay || !az)))
    return fail(c, NB_E_INVALID, "nb_accel_fp64: bad arguments");
  if (nidx == 0) return NB_OK;
  for (int64_t k = 0; k < nidx; k++)
    if (idx[k] < 0 || idx[k] >= c->n) return fail(c, NB_E_INVALID, "nb_accel_fp64: index out of range");
  NB_HIP(c, hipSetDevice(c->device));
  // staged through the context's scratch, in chunks of diag_len / 4 targets (an index and three
  // outputs each)
  const int64_t chunk = c->diag_len / 4;
  int64_t *didx = reinterpret_cast<int64_t *>(c->diag);
  double *dout = c->diag + chunk;
  for (int64_t k0 = 0; k0 < nidx; k0 += chunk) {
    const int64_t k = std::min(chunk, nidx - k0);
    NB_HIP(c, hipMemcpyAsync(didx, idx + k0, (size_t)k * sizeof(int64_t), hipMemcpyHostToDevice,
                             c->stream));
    Accel64Args a;
    a.x = c->x; a.y = c->y; a.z = c->z; a.m = c->m; a.r = c->r;
    a.flags = c->flags;
    a.idx = didx;
    a.nidx = k;
    a.out = dout;
    a.n = (int)c->n;
    a.G = c->p.G;
    a.eps = c->p.eps;
    a.cutoff = c->p.overlap_cutoff;
    NB_HIP(c, launch_accel_f64(a, c->stream));
    const size_t db = (size_t)k * sizeof(double);
    NB_HIP(c, hipMemcpyAsync(ax + k0, dout, db, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipMemcpyAsync(ay + k0, dout + k, db, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipMemcpyAsync(az + k0, dout + 2 * k, db, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
  }
  return NB_OK;
}

int nb_get_pairs(nb_ctx *c, int32_t *ij, int64_t cap, int64_t *n) {
  if (!c || !n || cap < 0 || (cap > 0 && !ij)) return c ? fail(c, NB_E_INVALID, "nb_get_pairs: bad arguments") : NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_get_pairs: no bodies loaded");
  NB_HIP(c, hipSetDevice(c->device));
  DevCounters ctr;
  NB_HIP(c, hipMemcpyAsync(&ctr, c->ctr, sizeof ctr, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  const int64_t stored = std::min<int64_t>((int64_t)ctr.pair_count, c->pair_cap);
  *n = c->p.collisions ? (int64_t)ctr.pair_count : 0;
  const int64_t k = c->p.collisions ? std::min(stored, cap) : 0;
  if (k > 0) {
    std::vector<uint64_t> keys((size_t)k);
    NB_HIP(c, hipMemcpy(keys.data(), c->pairs, (size_t)k * sizeof(uint64_t), hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < k; q++) {
      ij[2 * q] = (int32_t)(keys[q] >> 32);
      ij[2 * q + 1] = (int32_t)(keys[q] & 0xffffffffu);
    }
  }
  return NB_OK;
}

int nb_store(nb_ctx *c, double *x, double *y, double *z, double *vx, double *vy, double *vz,
             uint8_t *flags, double *rest) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_store: no bodies loaded");
  if (c->n == 0) return NB_OK;
  NB_HIP(c, hipSetDevice(c->device));
  const size_t db = (size_t)c->n * sizeof(double);
  double *dst[7] = {x, y, z, vx, vy, vz, rest};
  const double *src[7] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->rest};
  for (int k = 0; k < 7; k++)
    if (dst[k]) NB_HIP(c, hipMemcpyAsync(dst[k], src[k], db, hipMemcpyDeviceToHost, c->stream));
  if (flags) NB_HIP(c, hipMemcpyAsync(flags, c->flags, (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  return NB_OK;
}

int nb_store_bodies(nb_ctx *c, double *x, double *y, double *z, double *vx, double *vy,
                    double *vz, double *m, double *r, int32_t *beh, uint8_t *flags, double *rest,
                    double *frag, double *frag_step) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_store_bodies: no bodies loaded");
  if (c->n == 0) return NB_OK;
  NB_HIP(c, hipSetDevice(c->device));
  const size_t db = (size_t)c->n * sizeof(double);
  double *dst[10] = {x, y, z, vx, vy, vz, m, r, rest, frag};
  const double *src[10] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m, c->r, c->rest, c->frag};
  for (int k = 0; k < 10; k++)
    if (dst[k]) NB_HIP(c, hipMemcpyAsync(dst[k], src[k], db, hipMemcpyDeviceToHost, c->stream));
  if (beh)
    NB_HIP(c, hipMemcpyAsync(beh, c->beh, (size_t)c->n * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
  if (flags) NB_HIP(c, hipMemcpyAsync(flags, c->flags, (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
  if (frag_step) {
    if (c->fstep)
      NB_HIP(c, hipMemcpyAsync(frag_step, c->fstep, db, hipMemcpyDeviceToHost, c->stream));
    else
      memset(frag_step, 0, db);
  }
  NB_HIP(c, hipStreamSynchronize(c->stream));
  return NB_OK;
}

int nb_store_render(nb_ctx *c, float *xyz, uint8_t *exists) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_store_render: no bodies loaded");
  if (c->n == 0) return NB_OK;
  if (!xyz) return fail(c, NB_E_INVALID, "nb_store_render: NULL xyz");
  NB_HIP(c, hipSetDevice(c->device));
  float *tmp = reinterpret_cast<float *>(c->diag);  // 3 n floats of the context's scratch
  NB_HIP(c, launch_render(c->x, c->y, c->z, c->flags, (int)c->n, tmp, c->stream));
  NB_HIP(c, hipMemcpyAsync(xyz, tmp, (size_t)3 * c->n * sizeof(float), hipMemcpyDeviceToHost,
                           c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  if (exists) {
    std::vector<uint8_t> f((size_t)c->n);
    NB_HIP(c, hipMemcpy(f.data(), c->flags, (size_t)c->n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < c->n; i++) exists[i] = (f[(size_t)i] & NB_F_EXISTS) ? 1 : 0;
  }
  return NB_OK;
}

int nb_diagnostics(nb_ctx *c, nb_diag *out) {
  if (!c || !out) return c ? fail(c, NB_E_INVALID, "nb_diagnostics: NULL out") : NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_diagnostics: no bodies loaded");
  NB_HIP(c, hipSetDevice(c->device));
  NB_HIP(c, launch_diag(c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m, c->flags, (int)c->n, c->p.G,
                        c->p.eps, c->diag, c->diag_len, c->stream));
  double r[8];
  NB_HIP(c, hipMemcpyAsync(r, c->diag + (size_t)c->n * 8, sizeof r, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  out->kinetic = r[0];
  out->potential = r[1];
  out->momentum[0] = r[2];
  out->momentum[1] = r[3];
  out->momentum[2] = r[4];
  out->mass = r[5];
  out->n_active = (int64_t)r[6];
  return NB_OK;
}

int nb_set_timing(nb_ctx *c, int32_t enable) {
  if (!c) return NB_E_INVALID;
  c->timing = enable != 0;
  return NB_OK;
}

int nb_get_config(nb_ctx *c, int32_t *cfg) {
  if (!c || !cfg) return NB_E_INVALID;
  cfg[0] = c->shape.ipl;
  cfg[1] = c->shape.tile_j;
  cfg[2] = c->shape.jsplit;
  cfg[3] = c->shape.block;
  cfg[4] = c->shape.variant;
  return NB_OK;
}

int nb_autotune(nb_ctx *c, int32_t reps, int32_t *cfg) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded || c->n == 0) return fail(c, NB_E_STATE, "nb_autotune: no bodies loaded");
  if (c->p.precision == NB_PREC_FP64) {
    if (cfg) nb_get_config(c, cfg);
    return NB_OK;
  }
  if (reps < 1) reps = 1;
  NB_HIP(c, hipSetDevice(c->device));
  const ForceShape keep = c->shape;
  const int64_t npad = c->npad;
  std::vector<ForceShape> cand;
  for (int variant : {1, 2})
    for (int ipl : {1, 2, 4})
      for (int block : {128, 256})
        for (int tile : {256, 512}) {
          ForceShape s;
          s.variant = variant;
          s.ipl = ipl;
          s.block = block;
          s.tile_j = tile;
          if (npad % (block * ipl) || npad % tile) continue;
          const int64_t bx = npad / (block * ipl);
          const int64_t tiles = npad / tile;
          for (int64_t target : {1024, 2048, 4096}) {
            s.jsplit = (int)std::max<int64_t>(
                1, std::min<int64_t>({(target + bx - 1) / bx, (int64_t)kMaxSplit, tiles}));
            if (shape_valid(s, npad)) cand.push_back(s);
          }
        }
  hipEvent_t e0, e1;
  NB_HIP(c, hipEventCreate(&e0));
  NB_HIP(c, hipEventCreate(&e1));
  float best = 1e30f;
  ForceShape best_s = keep;
  for (const ForceShape &s : cand) {
    bool dup = false;
    for (const ForceShape &q : cand) {
      if (&q == &s) break;
      if (q.variant == s.variant && q.ipl == s.ipl && q.block == s.block &&
          q.tile_j == s.tile_j && q.jsplit == s.jsplit)
        dup = true;
    }
    if (dup) continue;
    c->shape = s;
    if (prepare_shape(c)) continue;
    ForceArgs a = make_force_args(c);
    if (launch_force_f32(s, a, c->stream) != hipSuccess) continue;  // warm-up
    NB_HIP(c, hipEventRecord(e0, c->stream));
    for (int k = 0; k < reps; k++) (void)launch_force_f32(s, a, c->stream);
    NB_HIP(c, hipEventRecord(e1, c->stream));
    NB_HIP(c, hipEventSynchronize(e1));
    float ms = 0;
    NB_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    if (ms < best) {
      best = ms;
      best_s = s;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->shape = best_s;
  c->shape_tuned = true;
  int rc = prepare_shape(c);
  if (rc) return rc;
  NB_HIP(c, hipStreamSynchronize(c->stream));
  if (cfg) nb_get_config(c, cfg);
  return NB_OK;
}

}  // extern "C"

// The largest radius over active bodies (the value nb_load computes on the host), from the device
// arrays after a removal, compaction or patch.
static int refresh_rmax(Ctx *c) {
  unsigned long long *d = c->small, h = 0;
  NB_HIP(c, hipMemsetAsync(d, 0, sizeof h, c->stream));
  NB_HIP(c, launch_rmax_active(c->r, c->flags, (int)c->n, d, c->stream));
  NB_HIP(c, hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  double v;
  memcpy(&v, &h, sizeof v);
  c->rmax = v;
  return NB_OK;
}

// ---------------------------------------------------------------------------------------------
// Collection lifecycle on the device
// ---------------------------------------------------------------------------------------------
extern "C" {

int nb_add_bodies(nb_ctx *c, int64_t k, const double *x, const double *y, const double *z,
                  const double *vx, const double *vy, const double *vz, const double *m,
                  const double *r, const int32_t *beh, const uint8_t *flags, const double *frag,
                  const double *frag_step) {
  if (!c) return NB_E_INVALID;
  if (k < 0 || k > (int64_t)1 << 30) return fail(c, NB_E_INVALID, "nb_add_bodies: bad count");
  if (k > 0 && (!x || !y || !z || !vx || !vy || !vz || !m || !r || !beh || !flags))
    return fail(c, NB_E_INVALID, "nb_add_bodies: NULL array");
  const double *src[10] = {x, y, z, vx, vy, vz, m, r, nullptr, frag};
  auto &A = c->adds;
  for (int q = 0; q < 10; q++)
    for (int64_t i = 0; i < k; i++) A.a[q].push_back(src[q] ? src[q][i] : (q == 8 ? 1.0 : 0.0));
  for (int64_t i = 0; i < k; i++) A.fstep.push_back(frag_step ? frag_step[i] : 0.0);
  A.beh.insert(A.beh.end(), beh, beh + k);
  A.flags.insert(A.flags.end(), flags, flags + k);
  return NB_OK;
}

int nb_cycle(nb_ctx *c, double R, int64_t *n_out) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_cycle: no bodies loaded");
  if (!std::isfinite(R)) return fail(c, NB_E_INVALID, "nb_cycle: R must be finite");
  NB_HIP(c, hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const int64_t n = c->n, nadd = c->adds.count();
  int64_t nk = 0;
  if (n > 0) {
    if (c->cyc_cap < n) {
      NB_HIP(c, dalloc(&c->cyc_keep, (size_t)c->cap));
      NB_HIP(c, dalloc(&c->cyc_pos, (size_t)c->cap));
      c->cyc_cap = c->cap;
    }
    int32_t *keep = c->cyc_keep, *pos = c->cyc_pos;
    NB_HIP(c, launch_keep_flags(c->flags, (int)n, keep, st));
    size_t tb = 0;
    NB_HIP(c, rocprim::exclusive_scan(nullptr, tb, keep, pos, 0, (size_t)n,
                                      rocprim::plus<int32_t>(), st));
    if (tb > c->cyc_tmp_bytes) {
      if (c->cyc_tmp) (void)hipFree(c->cyc_tmp);
      c->cyc_tmp = nullptr;
      NB_HIP(c, hipMalloc(&c->cyc_tmp, tb));
      c->cyc_tmp_bytes = tb;
    }
    NB_HIP(c, rocprim::exclusive_scan(c->cyc_tmp, tb, keep, pos, 0, (size_t)n,
                                      rocprim::plus<int32_t>(), st));
    int32_t lp = 0, lk = 0;
    NB_HIP(c, hipMemcpyAsync(&lp, pos + n - 1, 4, hipMemcpyDeviceToHost, st));
    NB_HIP(c, hipMemcpyAsync(&lk, keep + n - 1, 4, hipMemcpyDeviceToHost, st));
    NB_HIP(c, hipStreamSynchronize(st));
    nk = (int64_t)lp + lk;
  }
  int64_t nq = 0;  // fragments waiting in the device queue
  if (c->p.events) {
    DevCounters h;
    NB_HIP(c, hipMemcpyAsync(&h, c->ctr, sizeof h, hipMemcpyDeviceToHost, st));
    NB_HIP(c, hipStreamSynchronize(st));
    nq = (int64_t)h.frag_pending;
  }
  if (nk == n && nadd == 0 && nq == 0) {  // nothing removed, nothing added
    if (n_out) *n_out = n;
    return NB_OK;
  }
  const int64_t n_new = nk + nq + nadd;
  if (n_new > (int64_t)1 << 30) return fail(c, NB_E_INVALID, "nb_cycle: too many bodies");
  const int64_t npad_new = std::max<int64_t>(round_up(n_new, kPad), kPad);
  const int64_t cap_new = std::max<int64_t>(c->cap, npad_new);
  auto &S = c->spare;
  if (S.cap < cap_new || (c->p.events && !S.fst)) {  // (re)allocate the second set
    for (auto *&q : S.a)
      if (q) (void)hipFree(q), q = nullptr;
    void *ps[] = {S.beh, S.flags, S.collided, S.fstep, S.fst};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    S.beh = nullptr;
    S.flags = S.collided = nullptr;
    S.fstep = nullptr;
    S.fst = nullptr;
    S.cap = 0;
    bool ok = true;
    for (int q = 0; q < 10 && ok; q++)
      ok = hipMalloc((void **)&S.a[q], (size_t)cap_new * 8) == hipSuccess;
    ok = ok && hipMalloc((void **)&S.beh, (size_t)cap_new * 4) == hipSuccess &&
         hipMalloc((void **)&S.flags, (size_t)cap_new) == hipSuccess &&
         hipMalloc((void **)&S.collided, (size_t)cap_new) == hipSuccess;
    if (ok && c->p.events)
      ok = hipMalloc((void **)&S.fstep, (size_t)cap_new * 8) == hipSuccess &&
           hipMalloc((void **)&S.fst, (size_t)cap_new * sizeof(FragState)) == hipSuccess;
    if (!ok) {
      free_spare(c);
      return fail(c, NB_E_NOMEM, "nb_cycle: device allocation failed");
    }
    S.cap = cap_new;
  }
  double *d[10];
  for (int q = 0; q < 10; q++) d[q] = S.a[q];
  int32_t *beh = S.beh;
  uint8_t *fl = S.flags, *col = S.collided;
  double *fs = c->p.events ? S.fstep : nullptr;
  FragState *fst = c->p.events ? S.fst : nullptr;
  hipError_t e = hipSuccess;
  for (int q = 0; q < 10 && e == hipSuccess; q++)
    e = hipMemsetAsync(d[q], 0, (size_t)cap_new * 8, st);
  if (e == hipSuccess) e = hipMemsetAsync(beh, 0, (size_t)cap_new * 4, st);
  if (e == hipSuccess) e = hipMemsetAsync(fl, 0, (size_t)cap_new, st);
  if (e == hipSuccess) e = hipMemsetAsync(col, 0, (size_t)cap_new, st);
  if (e == hipSuccess && fs) e = hipMemsetAsync(fs, 0, (size_t)cap_new * 8, st);
  if (e == hipSuccess && fst) e = hipMemsetAsync(fst, 0, (size_t)cap_new * sizeof(FragState), st);
  CompactArgs ca;
  const double *olds[10] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m, c->r, c->rest, c->frag};
  for (int q = 0; q < 10; q++) {
    ca.src[q] = olds[q];
    ca.dst[q] = d[q];
  }
  ca.beh_src = c->beh;
  ca.beh_dst = beh;
  ca.flags_src = c->flags;
  ca.flags_dst = fl;
  ca.pos = c->cyc_pos;
  ca.fstep_src = c->fstep;
  ca.fstep_dst = fs;
  ca.fst_src = c->fst;
  ca.fst_dst = fst;
  ca.n = (int)n;
  if (e == hipSuccess) e = launch_compact(ca, st);
  // the device's fragments first (the newest events), then the host's queue
  if (e == hipSuccess && nq) e = launch_append_frags(c, nq, d, beh, fl, fs, fst, nk, R, st);
  // queued adds go after the survivors, newest first, each with restitution R: the reference
  // keeps its events with PushFront (body_collection.go:100) and Cycle walks them from the front
  // (body_collection.go:273-279, 286-290)
  auto &A = c->adds;
  std::fill(A.a[8].begin(), A.a[8].end(), R);
  for (int q = 0; q < 10; q++) std::reverse(A.a[q].begin(), A.a[q].end());
  std::reverse(A.beh.begin(), A.beh.end());
  std::reverse(A.flags.begin(), A.flags.end());
  std::reverse(A.fstep.begin(), A.fstep.end());
  const int64_t at = nk + nq;
  for (int q = 0; q < 10 && e == hipSuccess && nadd; q++)
    e = hipMemcpyAsync(d[q] + at, A.a[q].data(), (size_t)nadd * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && nadd)
    e = hipMemcpyAsync(beh + at, A.beh.data(), (size_t)nadd * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && nadd)
    e = hipMemcpyAsync(fl + at, A.flags.data(), (size_t)nadd, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && nadd && fs)
    e = hipMemcpyAsync(fs + at, A.fstep.data(), (size_t)nadd * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  NB_HIP(c, e);
  // swap the sets: the compacted arrays go live, the old ones become the second set
  const int64_t live_cap = S.cap;  // >= cap_new >= c->cap
  double **cur[10] = {&c->x, &c->y, &c->z, &c->vx, &c->vy, &c->vz, &c->m, &c->r, &c->rest,
                      &c->frag};
  for (int q = 0; q < 10; q++) std::swap(*cur[q], S.a[q]);
  std::swap(c->beh, S.beh);
  std::swap(c->flags, S.flags);
  std::swap(c->collided, S.collided);
  if (c->p.events) {
    std::swap(c->fstep, S.fstep);
    std::swap(c->fst, S.fst);
  }
  S.cap = c->cap;  // the old live set's capacity
  if (live_cap > c->cap) {
    c->cap = live_cap;
    NB_HIP(c, dalloc(&c->js, c->cap));
    c->diag_len = c->cap * 8 + 64;
    NB_HIP(c, dalloc(&c->diag, c->diag_len));
    c->acc_split = 0;   // per-slice buffers are re-sized by prepare_shape
    c->rmax_tile_j = 0;
  }
  {
    const int rc = size_pairs(c, n_new);
    if (rc) return rc;
  }
  c->n = n_new;
  c->npad = npad_new;
  NB_HIP(c, launch_write_stream(c->x, c->y, c->z, c->m, c->flags, c->js, (int)n_new,
                                (int)npad_new, st));
  choose_shape(c);
  int rc = prepare_shape(c);
  if (rc) return rc;
  for (int64_t i = 0; i < nadd; i++) {
    const uint8_t f = A.flags[(size_t)i];
    if ((f & NB_F_EXISTS) && !(f & NB_F_FRAGMENTING)) c->n_active++;
  }
  c->n_active += nq;  // fragments exist and are not fragmenting
  A.clear();
  if (c->p.events) {
    DevCounters h;
    NB_HIP(c, hipMemcpyAsync(&h, c->ctr, sizeof h, hipMemcpyDeviceToHost, st));
    NB_HIP(c, hipStreamSynchronize(st));
    events_refresh_host(c, h);
  }
  rc = refresh_rmax(c);
  if (rc) return rc;
  if (n_out) *n_out = n_new;
  return NB_OK;
}

int nb_remove_bodies(nb_ctx *c, int64_t count, int64_t *removed) {
  if (!c) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_remove_bodies: no bodies loaded");
  if (count < -1) return fail(c, NB_E_INVALID, "nb_remove_bodies: count < -1");
  if (removed) *removed = 0;
  const int64_t n = c->n;
  if (count == 0 || n == 0) return NB_OK;
  NB_HIP(c, hipSetDevice(c->device));
  const bool all = count == -1;
  const int64_t step = all ? 1 : (count > n ? 1 : n / count);
  const int64_t npoints = all ? n : (n + step - 1) / step;
  unsigned long long *cnt = c->small, h[2] = {0, 0};
  NB_HIP(c, hipMemsetAsync(cnt, 0, sizeof h, c->stream));
  NB_HIP(c, launch_remove_walk(c->flags, c->m, (int)n, (int)step, (int)npoints, all ? 1 : 0, cnt,
                               c->stream));
  NB_HIP(c, hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, launch_write_stream(c->x, c->y, c->z, c->m, c->flags, c->js, (int)n, (int)c->npad,
                                c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  c->n_active = std::max<int64_t>(0, c->n_active - (int64_t)h[1]);
  if (removed) *removed = (int64_t)h[0];
  return refresh_rmax(c);
}

int nb_get_body(nb_ctx *c, int64_t index, nb_body *out) {
  if (!c || !out) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_get_body: no bodies loaded");
  if (index < 0 || index >= c->n) return fail(c, NB_E_INVALID, "nb_get_body: index out of range");
  NB_HIP(c, hipSetDevice(c->device));
  memset(out, 0, sizeof *out);
  const double *src[10] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m, c->r, c->rest, c->frag};
  double *dst[10] = {&out->x, &out->y, &out->z, &out->vx, &out->vy, &out->vz, &out->m, &out->r,
                     &out->rest, &out->frag};
  for (int q = 0; q < 10; q++)
    NB_HIP(c, hipMemcpyAsync(dst[q], src[q] + index, 8, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipMemcpyAsync(&out->beh, c->beh + index, 4, hipMemcpyDeviceToHost, c->stream));
  NB_HIP(c, hipMemcpyAsync(&out->flags, c->flags + index, 1, hipMemcpyDeviceToHost, c->stream));
  if (c->fstep)
    NB_HIP(c, hipMemcpyAsync(&out->frag_step, c->fstep + index, 8, hipMemcpyDeviceToHost,
                             c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  return NB_OK;
}

int nb_set_body(nb_ctx *c, int64_t index, const nb_body *in) {
  if (!c || !in) return NB_E_INVALID;
  if (!c->loaded) return fail(c, NB_E_STATE, "nb_set_body: no bodies loaded");
  if (index < 0 || index >= c->n) return fail(c, NB_E_INVALID, "nb_set_body: index out of range");
  NB_HIP(c, hipSetDevice(c->device));
  uint8_t old = 0;
  NB_HIP(c, hipMemcpyAsync(&old, c->flags + index, 1, hipMemcpyDeviceToHost, c->stream));
  double *dst[10] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m, c->r, c->rest, c->frag};
  const double *src[10] = {&in->x, &in->y, &in->z, &in->vx, &in->vy, &in->vz, &in->m, &in->r,
                           &in->rest, &in->frag};
  for (int q = 0; q < 10; q++)
    NB_HIP(c, hipMemcpyAsync(dst[q] + index, src[q], 8, hipMemcpyHostToDevice, c->stream));
  NB_HIP(c, hipMemcpyAsync(c->beh + index, &in->beh, 4, hipMemcpyHostToDevice, c->stream));
  NB_HIP(c, hipMemcpyAsync(c->flags + index, &in->flags, 1, hipMemcpyHostToDevice, c->stream));
  if (c->fstep)
    NB_HIP(c, hipMemcpyAsync(c->fstep + index, &in->frag_step, 8, hipMemcpyHostToDevice,
                             c->stream));
  NB_HIP(c, launch_write_stream(c->x, c->y, c->z, c->m, c->flags, c->js, (int)c->n,
                                (int)c->npad, c->stream));
  NB_HIP(c, launch_rmax_tile(c->r, c->flags, (int)c->n, c->shape.tile_j, tiles_of(c),
                             c->rmax_tile, c->stream));
  NB_HIP(c, hipStreamSynchronize(c->stream));
  auto active = [](uint8_t f) { return (f & NB_F_EXISTS) && !(f & NB_F_FRAGMENTING); };
  c->n_active += (active(in->flags) ? 1 : 0) - (active(old) ? 1 : 0);
  return refresh_rmax(c);
}

}  // extern "C"
