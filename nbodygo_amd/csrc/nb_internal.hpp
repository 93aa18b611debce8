// nb_internal.hpp -- device data layout, context and launch helpers shared by the HIP sources of
// libnbgpu.so (the MI355X per-cycle compute path of aceeric/nbodygo).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nbody_gpu.h"

namespace nb {

// ---- fixed shapes ---------------------------------------------------------------------------
constexpr int kPad = 1024;           // body arrays are padded to a multiple of this
constexpr int kFlagSlots = 4;        // flagged tiles remembered per (slice, target)
constexpr int kMaxSplit = 64;        // max source-range slices
constexpr int kSortLdsKeys = 8192;   // pair keys sorted in LDS by one workgroup
constexpr float kSentinel = 1.0e18f; // position of padding / inactive sources in the FP32 stream

// Device counters, one struct per context (zeroed by a fill kernel, never by hipMemset, so the
// step sequence stays capturable).
struct DevCounters {
  uint32_t pair_count;       // pairs appended in the current force phase
  uint32_t rec_count;        // (slice, target) fix-up records appended in the current phase
  uint64_t pairs_total;
  uint64_t pairs_last;
  uint64_t resolved_total;
  uint64_t fragment_requests;
  uint64_t removed_total;
  uint64_t fixup_tiles;
  uint32_t pair_overflow;
  uint32_t pad;
  uint64_t sur_fallbacks;    // surrogate steps redone by direct summation
  // subsume and fragmentation events (nb_params.events)
  uint64_t subsumed_total;   // ResolveSubsume calls
  uint64_t frag_started;     // bodies that started fragmenting
  uint64_t frag_finished;    // fragmenting bodies that ran out of fragments
  uint64_t frag_emitted;     // fragments created
  uint64_t frag_seq;         // next fragmentation id (order of initiation)
  uint64_t inactive_events;  // active bodies made inactive by a subsume or a fragmentation
  uint32_t frag_pending;     // fragments in the queue, waiting for nb_cycle
  uint32_t frag_overflow;    // a cycle deferred its fragments (queue full)
};

// Fragmentation bookkeeping of one body (Body.fragInfo, fragcalc.go:75-83, and the force the body
// keeps while fragmenting).
struct FragState {
  double fx, fy, fz;    // force (or acceleration / G on the FP32 path) of the starting cycle
  double mass, radius;  // fragment mass; the body's radius when it started
  double px, py, pz;    // its position then (fragInfo.curPos)
  int64_t left;         // fragments still to create
  int64_t made;         // fragments created so far
  int64_t id;           // fragmentation id (seeds the fragment positions)
};

constexpr int kFragPerCycle = 101;  // fragment() creates until cnt > maxFragsPerCycle (100)
constexpr double kFragRadius = 1.0; // fragment radius: Pow(x, 1/3) with 1/3 == 0 (fragcalc.go:81)

__device__ __forceinline__ bool nb_collide_b(int32_t b) {
  return b == NB_CB_ELASTIC || b == NB_CB_FRAGMENT;
}
// Compute's dispatch of an overlapping pair (body.go:172-183): a collision event when both bodies
// are Elastic or Fragment; with events on, otherwise a subsume event when either body is Subsume
// and the larger one (radii unequal) reaches the other's centre (dist <= its radius).
__device__ __forceinline__ bool nb_pair_event(int32_t bi, int32_t bj, double ri, double rj,
                                              double dist, int events) {
  if (nb_collide_b(bi) && nb_collide_b(bj)) return true;
  if (!events || (bi != NB_CB_SUBSUME && bj != NB_CB_SUBSUME)) return false;
  return (ri > rj && dist <= ri) || (rj > ri && dist <= rj);
}
// Whether a body can be in any event pair.
__device__ __forceinline__ bool nb_may_pair(int32_t b, int events) {
  return events || nb_collide_b(b);
}

struct Sur;  // far-field surrogate state (nb_sur.hip)
struct Bh;   // Barnes-Hut octree state (nb_bh.hip)

// FP32 force kernel launch shape.
struct ForceShape {
  int ipl = 2;      // target bodies per lane
  int tile_j = 256; // sources per Kahan tile
  int jsplit = 1;   // source-range slices (grid.y)
  int block = 256;  // threads per workgroup
  int variant = 1;  // 1 = scalar-cache j stream, 2 = LDS-staged tiles
};

struct Ctx {
  nb_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  int64_t n = 0;     // bodies
  int64_t npad = 0;  // padded count (multiple of kPad)
  int64_t cap = 0;   // allocated capacity (bodies)
  int64_t n_active = 0;
  double rmax = 0;  // largest radius over active bodies (set by nb_load and the lifecycle calls)
  bool loaded = false;

  // FP64 state (SoA)
  double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *m = nullptr, *r = nullptr, *rest = nullptr, *frag = nullptr;
  int32_t *beh = nullptr;
  uint8_t *flags = nullptr, *collided = nullptr;

  // FP32 source stream (x, y, z, m) and per-tile max radius
  float4 *js = nullptr;
  float *rmax_tile = nullptr;
  int rmax_tile_j = 0;  // tile size rmax_tile was built for

  // per-slice accelerations (FP32 path) or force (FP64 path): [split][3][npad] doubles
  double *acc = nullptr;
  int acc_split = 0;
  // flagged tiles: counts [split][npad], lists [split][npad][kFlagSlots], records [split*npad]
  int32_t *fcnt = nullptr, *flist = nullptr;
  int2 *recs = nullptr;

  // collision pairs as 64-bit keys (i << 32 | j); the sort's merge buffer and the resolution's
  // swapped keys (same size); the resolution's per-pair words: 5 int32 per pair (two successors,
  // a predecessor count, two frontier lists)
  uint64_t *pairs = nullptr, *pairs_tmp = nullptr, *pairs_sw = nullptr;
  int32_t *res_words = nullptr;
  int64_t pair_cap = 0;
  int64_t pair_cap_pow2 = 0;

  DevCounters *ctr = nullptr;
  // scratch of cap * 8 + 64 doubles: the diagnostics' reduction, and the staging area of
  // nb_compute_forces, nb_store_render and nb_accel_fp64 (each call completes before it returns)
  double *diag = nullptr;
  int64_t diag_len = 0;
  unsigned long long *small = nullptr;  // 8 counters for single-value reductions (context-long)

  ForceShape shape;
  bool shape_tuned = false;
  bool timing = false;
  std::vector<hipEvent_t> ev;
  std::vector<hipEvent_t> ev_res;  // around each cycle's resolution (timing enabled)

  Sur *sur = nullptr;
  bool sur_on = false;
  Bh *bh = nullptr;
  bool bh_on = false;

  // events (nb_params.events): per-body FragStep and fragmentation state, the fragment queue
  double *fstep = nullptr;
  FragState *fst = nullptr;
  double *fq = nullptr;  // queue: 8 arrays of fq_cap (x y z vx vy vz m r)
  int64_t fq_cap = 0;
  nb_event_stats evs{};  // host copy, refreshed at the end of nb_step and nb_cycle
  void *ev_scan_tmp = nullptr;
  size_t ev_scan_bytes = 0;

  // nb_cycle compacts into a second set of per-body arrays and swaps it with the live one, so a
  // cycle allocates nothing once both sets have the capacity
  struct Spare {
    double *a[10] = {};  // x y z vx vy vz m r rest frag
    int32_t *beh = nullptr;
    uint8_t *flags = nullptr, *collided = nullptr;
    double *fstep = nullptr;
    FragState *fst = nullptr;
    int64_t cap = 0;
  } spare;
  int32_t *cyc_keep = nullptr, *cyc_pos = nullptr;  // keep flags and their prefix sum
  void *cyc_tmp = nullptr;
  size_t cyc_tmp_bytes = 0;
  int64_t cyc_cap = 0;

  // bodies queued by nb_add_bodies, appended by the next nb_cycle
  struct Adds {
    std::vector<double> a[10];  // x y z vx vy vz m r rest frag
    std::vector<double> fstep;  // FragStep
    std::vector<int32_t> beh;
    std::vector<uint8_t> flags;
    int64_t count() const { return (int64_t)beh.size(); }
    void clear() {
      for (auto &v : a) v.clear();
      fstep.clear();
      beh.clear();
      flags.clear();
    }
  } adds;
};

bool hip_ok(Ctx *c, hipError_t e, const char *what);
#define NB_HIP(c, call)                                  \
  do {                                                   \
    if (!::nb::hip_ok((c), (call), #call)) return NB_E_HIP; \
  } while (0)

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// A wave-uniform value copied into a VGPR.  On gfx950 a VALU op whose operand is an SGPR does not
// dual-issue (measured: ~61 vs ~90-110 FP32 lane-ops/clk/CU), so loop-invariant scalars used in
// the pairwise inner loops are moved to VGPRs once.
__device__ __forceinline__ float to_vgpr(float s) {
  float v;
  asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}

// ---- kernels launched from nb_api.cpp (defined in the .hip files) ----------------------------
struct ForceArgs {
  const float4 *js;
  const float *rmax_tile;
  const double *r;
  double *acc;
  int32_t *fcnt;
  int32_t *flist;
  int npad;
  int n;
  int tiles;            // npad / tile_j
  int tiles_per_slice;
  int tile_j;
  float eps2;
  int detect;           // track per-tile minimum distance and flag tiles
  int drop_flagged;     // flagged tiles' FP32 partials are discarded (recomputed in FP64)
  int2 *recs;           // (slice, target) of every target with flagged tiles
  int64_t rec_cap;      // entries in recs (slices * padded bodies)
  DevCounters *ctr;
  const uint32_t *run_flag;  // if set, the launch does nothing unless *run_flag != 0
};
hipError_t launch_force_f32(const ForceShape &s, const ForceArgs &a, hipStream_t st);
ForceArgs make_force_args(Ctx *c);
int fail_msg(Ctx *c, int code, const char *msg);
void set_global_error(const char *msg);  // the message nb_last_error(NULL) returns
void sur_free(Sur *s);
int sur_force_phase(Ctx *c, bool timed, int ev_slot);
void bh_free(Bh *b);
int bh_force_phase(Ctx *c, bool timed, int ev_slot);

struct FixupArgs {
  const double *x, *y, *z, *m, *r;
  const int32_t *beh;
  const uint8_t *flags;
  double *acc;
  const int32_t *fcnt;
  const int32_t *flist;
  const int2 *recs;
  int64_t rec_cap;
  uint64_t *pairs;
  DevCounters *ctr;
  int64_t pair_cap;
  int npad, n, tiles, tiles_per_slice, tile_j, split;
  double eps2;
  int recompute;  // add FP64 contributions of flagged tiles (the fast path dropped them)
  int cutoff;     // overlap gate
  int collisions; // emit pairs
  int events;     // subsume pairs too
};
hipError_t launch_fixup(const FixupArgs &a, hipStream_t st);

struct Force64Args {
  const double *x, *y, *z, *m, *r;
  const int32_t *beh;
  const uint8_t *flags;
  double *acc;  // force [3][npad]
  uint64_t *pairs;
  DevCounters *ctr;
  int64_t pair_cap;
  int n, npad;
  double G, eps;
  int cutoff, collisions, events;
};
hipError_t launch_force_f64(const Force64Args &a, hipStream_t st);

struct Accel64Args {
  const double *x, *y, *z, *m, *r;
  const uint8_t *flags;
  const int64_t *idx;
  int64_t nidx;
  double *out;  // [3][nidx]
  int n;
  double G, eps;
  int cutoff;
};
hipError_t launch_accel_f64(const Accel64Args &a, hipStream_t st);

hipError_t launch_sort_pairs(uint64_t *keys, uint64_t *tmp, int64_t cap, const DevCounters *ctr,
                             hipStream_t st);
int sort_pairs(Ctx *c);  // the context's pair list into canonical order

struct ResolveArgs {
  double *x, *y, *z, *vx, *vy, *vz, *m, *r, *rest, *frag;
  int32_t *beh;
  uint8_t *flags, *collided;
  const uint64_t *pairs;
  DevCounters *ctr;
  int64_t pair_cap;
  // events: subsume, and fragmentation started from shouldFragment
  int events;
  const double *fstep;
  FragState *fst;
  const double *acc;  // this cycle's force slices, kept by a body that starts fragmenting
  int split, npad;
  // events off: pairs that share no body resolve together.  pairs_sw / sort_tmp: pair_cap keys
  // each (the swapped keys and the sort's merge buffer); succ: 2 per pair; pending, list_a,
  // list_b: 1 per pair.  serial = one lane in canonical order.
  uint64_t *pairs_sw, *sort_tmp;
  int32_t *succ, *pending, *list_a, *list_b;
  int serial;
};
hipError_t launch_resolve(const ResolveArgs &a, hipStream_t st);

struct UpdateArgs {
  double *x, *y, *z, *vx, *vy, *vz;
  const double *m;
  double *rest;
  uint8_t *flags, *collided;
  const double *acc;
  int split, npad, n;
  double G, dt, R;
  int force_mode;  // 0: acc holds accelerations / G (FP32 path); 1: acc holds force (FP64 path)
  float4 *js;
  DevCounters *ctr;
  const FragState *fst;  // events: a fragmenting body is kicked by its kept force
};
hipError_t launch_update(const UpdateArgs &a, hipStream_t st);

hipError_t launch_write_stream(const double *x, const double *y, const double *z,
                               const double *m, const uint8_t *flags, float4 *js, int n,
                               int npad, hipStream_t st);
hipError_t launch_rmax_tile(const double *r, const uint8_t *flags, int n, int tile_j, int tiles,
                            float *out, hipStream_t st);
hipError_t launch_zero_counters(DevCounters *ctr, int full, hipStream_t st);
hipError_t launch_combine_force(const double *acc, int split, int npad, int n, const double *m,
                                double G, int force_mode, double *out, hipStream_t st);
hipError_t launch_diag(const double *x, const double *y, const double *z, const double *vx,
                       const double *vy, const double *vz, const double *m, const uint8_t *flags,
                       int n, double G, double eps, double *scratch, int64_t scratch_len,
                       hipStream_t st);
hipError_t launch_render(const double *x, const double *y, const double *z, const uint8_t *flags,
                         int n, float *xyz, hipStream_t st);

// collection lifecycle
struct CompactArgs {
  const double *src[10];  // x y z vx vy vz m r rest frag
  double *dst[10];
  const int32_t *beh_src;
  int32_t *beh_dst;
  const uint8_t *flags_src;
  uint8_t *flags_dst;
  const int32_t *pos;  // exclusive prefix sum of the keep flags
  const double *fstep_src;  // events: FragStep and fragmentation state (may be null)
  double *fstep_dst;
  const FragState *fst_src;
  FragState *fst_dst;
  int n;
};
hipError_t launch_keep_flags(const uint8_t *flags, int n, int32_t *keep, hipStream_t st);
hipError_t launch_compact(const CompactArgs &a, hipStream_t st);
hipError_t launch_rmax_active(const double *r, const uint8_t *flags, int n,
                              unsigned long long *out, hipStream_t st);
hipError_t launch_remove_walk(uint8_t *flags, double *m, int n, int step, int npoints, int all,
                              unsigned long long *removed, hipStream_t st);

// events (nb_events.hip)
int events_emit(Ctx *c);                 // fragment() of every fragmenting body, one cycle
int events_alloc(Ctx *c, int64_t cap);   // per-body arrays for cap bodies (zeroed)
int events_ensure_queue(Ctx *c);         // queue room for the next cycles' fragments
void events_free(Ctx *c);
void events_refresh_host(Ctx *c, const DevCounters &d);
// append the queue (newest first) at dst + at; returns fragments appended
hipError_t launch_append_frags(Ctx *c, int64_t nq, double *const dst[10], int32_t *beh,
                               uint8_t *flags, double *fstep, FragState *fst, int64_t at,
                               double R, hipStream_t st);

}  // namespace nb

// The opaque handle of the C-ABI is the context itself.
struct nb_ctx : public nb::Ctx {};
