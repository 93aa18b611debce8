/*
 * nbody_gpu.h -- C-ABI of the MI355X-native per-cycle compute path of aceeric/nbodygo.
 *
 * One call of nb_step() replaces, for every cycle it runs, the section of
 * ComputationRunner.runOneComputation from BuildTree to the Update loop
 * (cmd/runner/computation-runner.go:294-320):
 *   - the worker-pool fan-out of Body.Compute        (workpool.go:82-115, body.go:148-187)
 *     = O(N^2) gravity (calcForceFrom, body.go:214-225) + collision detection
 *       (Collided, body.go:192-208),
 *   - BodyCollection.ProcessMods                     (body_collection.go:212-233)
 *     = elastic collision resolution (ResolveCollision body.go:248-264,
 *       calcElasticCollision collisioncalc.go:42-186, doElastic collisioncalc.go:26-35),
 *   - the Body.Update loop                           (body.go:114-139).
 * The caller (the runner) keeps control polls, adds/deletes/mods, the result queue and
 * BodyCollection.Cycle.
 *
 * Conventions: plain C types and pointers only; every function returns NB_OK (0) or a negative
 * NB_E_* code, with a message from nb_last_error(); no exception crosses the boundary; the
 * caller keeps ownership of every pointer it passes and nothing is retained after return (so a
 * cgo caller can pass Go memory).  One context per caller thread; calls on one context are not
 * re-entrant.  Device memory is owned by the context.
 */
#ifndef NBODY_GPU_H
#define NBODY_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------------------- */
#define NB_OK 0
#define NB_E_INVALID (-1)       /* bad argument */
#define NB_E_HIP (-2)           /* HIP runtime error (message has the HIP error string) */
#define NB_E_NOMEM (-3)         /* device or host allocation failed */
#define NB_E_PAIR_OVERFLOW (-4) /* more collision pairs than pair_capacity in some step; the
                                   step completed, pairs beyond capacity were dropped (the
                                   reference drops events beyond its 1000-slot channel,
                                   body_collection.go:82-88) */
#define NB_E_STATE (-5)         /* call out of order (e.g. nb_step before nb_load) */
#define NB_E_UNSUPPORTED (-6)   /* option outside the ported path */

/* ---- body flags and collision behaviour (Body fields, body.go:34-52; globals.go:11-16) ---- */
#define NB_F_EXISTS 1u
#define NB_F_FRAGMENTING 2u
#define NB_F_SUN 4u
#define NB_F_PINNED 8u

#define NB_CB_NONE 0
#define NB_CB_SUBSUME 1
#define NB_CB_ELASTIC 2
#define NB_CB_FRAGMENT 3

/* ---- precision modes ---------------------------------------------------------------------- */
#define NB_PREC_FP32 0 /* FP32 pairwise force with per-tile Kahan summation, FP64 exact
                          overlap/collision tests, FP64 integration (production path) */
#define NB_PREC_FP64 1 /* FP64 force in the reference's operation order (bit-identical to the
                          reference's arithmetic for the force and update) */

typedef struct nb_params {
  double G;               /* gravitational constant; the reference hard-wires 6.673e-11
                             (body.go:19) */
  double eps;             /* Plummer softening length; 0 = the reference's unsoftened O(N^2)
                             force (body.go:214-225) */
  int32_t overlap_cutoff; /* 1 = pairs with d <= Ri+Rj contribute no force (body.go:219) */
  int32_t collisions;     /* 1 = detect (body.go:171-186) and resolve elastic collisions */
  int32_t precision;      /* NB_PREC_* */
  int32_t device;         /* HIP device ordinal */
  int64_t pair_capacity;  /* max collision pairs per step; 0 = default (max(4N, 65536)) */
  int32_t ipl;            /* FP32 force kernel: target bodies per lane (1,2,4); 0 = tuned */
  int32_t tile_j;         /* FP32 force kernel: j bodies per Kahan tile (multiple of 64);
                             0 = tuned */
  int32_t jsplit;         /* FP32 force kernel: source-range slices; 0 = tuned */
  int32_t block;          /* FP32 force kernel: threads per workgroup (64..256); 0 = tuned */
  int32_t variant;        /* FP32 force kernel: 0 = tuned, 1 = scalar-cache j stream,
                             2 = LDS-staged j tiles */
  int32_t events;         /* 1 = the reference's subsume and fragmentation events on the device
                             (Compute's dispatch, body.go:172-183; ResolveSubsume :228-244;
                             doFragment / fragment, fragcalc.go:54-117); 0 = elastic collisions
                             only (Subsume and None bodies make no pairs, shouldFragment is only
                             counted) */
  int32_t resolve;        /* collision resolution: 0 = pairs that share no body resolve together
                             (events off), 1 = one pair at a time on a single lane; both walk the
                             pairs in canonical order and give bit-identical states */
  int32_t reserved[5];
} nb_params;

typedef struct nb_stats {
  int64_t steps;             /* cycles run by this call */
  int64_t n;                 /* bodies in the context (existing or not) */
  int64_t n_active;          /* existing, non-fragmenting bodies after the call */
  double interactions;       /* ordered pairs evaluated by the force phase, summed over the
                                call's steps (the reference's nbody_computations counter,
                                body.go:166) */
  int64_t pairs_total;       /* detected unordered collision pairs, summed over steps */
  int64_t pairs_last;        /* detected unordered collision pairs in the last step */
  int64_t resolved_total;    /* pairs for which calcElasticCollision reported a collision */
  int64_t fragment_requests; /* shouldFragment() true (fragmentation is outside the path) */
  int64_t removed_total;     /* bodies set !Exists by the NaN guard (body.go:134-137) */
  int64_t fixup_tiles;       /* (target, tile) blocks sent to the FP64 exact pass */
  int32_t pair_overflow;     /* 1 if some step exceeded pair_capacity */
  int32_t pad0;
  double wall_ms;            /* host wall time of the call */
  double force_ms;           /* HIP-event time of the force kernels (timing enabled only) */
  int64_t force_launches;    /* force kernel launches timed */
  double resolve_ms;         /* HIP-event time of the collision resolution (timing enabled) */
} nb_stats;

typedef struct nb_diag {
  double kinetic;     /* sum 1/2 m v^2 over existing bodies */
  double potential;   /* -G sum_{i<j} mi mj / sqrt(d^2 + eps^2) */
  double momentum[3]; /* sum m v */
  double mass;        /* sum m */
  int64_t n_active;
} nb_diag;

typedef struct nb_ctx nb_ctx;

/* Defaults: G = 6.673e-11, eps = 0, overlap_cutoff = 1, collisions = 1, FP32, device 0 --
 * the reference's --no-barnes-hut semantics (cmd/server/main.go:221-222). */
int nb_params_init(nb_params *p);

/* Create a context on p->device.  On failure *out is NULL and nb_last_error(NULL) has the
 * message. */
int nb_create(const nb_params *p, nb_ctx **out);
void nb_destroy(nb_ctx *ctx);

/* Last error message of ctx (or of the last failed nb_create when ctx is NULL). */
const char *nb_last_error(const nb_ctx *ctx);

/* Replace the body array (BodyCollection.arr after Cycle, body_collection.go:253-296).
 * x..r: FP64 per body.  beh: CollisionBehavior per body.  flags: NB_F_* per body.
 * rest: per-body coefficient of restitution Body.r (NULL = 1.0, as NewBody sets, body.go:80).
 * frag: per-body Body.FragFactor (NULL = 0).  Copies into device SoA; nothing retained. */
int nb_load(nb_ctx *ctx, int64_t n, const double *x, const double *y, const double *z,
            const double *vx, const double *vy, const double *vz, const double *m,
            const double *r, const int32_t *beh, const uint8_t *flags, const double *rest,
            const double *frag);

/* Run nsteps cycles: force + detection, resolution in canonical pair order, FP64 update with
 * time step dt (Runner.timeScaling) and restitution R (Runner.R).  Synchronous.  out may be
 * NULL. */
int nb_step(nb_ctx *ctx, double dt, double R, int32_t nsteps, nb_stats *out);

/* Force phase only, on the current state (Body.fx/fy/fz after Compute): force components, not
 * acceleration.  Also records the detected pairs for nb_get_pairs.  No resolution, no update. */
int nb_compute_forces(nb_ctx *ctx, double *fx, double *fy, double *fz);

/* FP64 accelerations sum_j G mj (xj - xi) / (d^2 + eps^2)^(3/2) for the nidx targets in idx,
 * computed on the device in FP64 (with the context's overlap cutoff).  Used to state the FP32
 * path's force error against FP64. */
int nb_accel_fp64(nb_ctx *ctx, const int64_t *idx, int64_t nidx, double *ax, double *ay,
                  double *az);

/* Collision pairs detected by the last force phase, canonical order (i < j, ascending),
 * as 2*n int32 body indices.  *n receives the count; at most cap pairs are written. */
int nb_get_pairs(nb_ctx *ctx, int32_t *ij, int64_t cap, int64_t *n);

/* Copy the FP64 state back (any pointer may be NULL to skip it). */
int nb_store(nb_ctx *ctx, double *x, double *y, double *z, double *vx, double *vy, double *vz,
             uint8_t *flags, double *rest);
/* The whole body state, the inverse of nb_load (after nb_cycle the count is the one it
 * returned; any pointer may be NULL to skip it). */
int nb_store_bodies(nb_ctx *ctx, double *x, double *y, double *z, double *vx, double *vy,
                    double *vz, double *m, double *r, int32_t *beh, uint8_t *flags, double *rest,
                    double *frag, double *frag_step);

/* FP32 positions for rendering (NewRenderable, renderable.go:22-40): xyz is 3*n floats;
 * exists (n bytes, may be NULL) receives the Exists flag. */
int nb_store_render(nb_ctx *ctx, float *xyz, uint8_t *exists);

/* FP64 energy and momentum diagnostics (no reference counterpart; used for drift reports). */
int nb_diagnostics(nb_ctx *ctx, nb_diag *out);

/* Enable (1) or disable (0) HIP-event timing of the force kernels in nb_step. */
int nb_set_timing(nb_ctx *ctx, int32_t enable);

/* The FP32 force kernel shape in use: ipl, tile_j, jsplit, block, variant (5 ints). */
int nb_get_config(nb_ctx *ctx, int32_t *cfg5);

/* Measure the FP32 force kernel on the loaded state over a built-in candidate set of shapes
 * (ipl x tile_j x jsplit x block x variant), keep the fastest for this context, and write it to
 * cfg5 (may be NULL).  reps = timed launches per candidate. */
int nb_autotune(nb_ctx *ctx, int32_t reps, int32_t *cfg5);

/* ---- collection lifecycle on the device ---------------------------------------------------- */
/* The runner's collection bookkeeping without a round trip of the whole state through the host:
 * removals, queued adds and compaction (ComputationRunner.processDeletes,
 * computation-runner.go:176-216; BodyCollection.Cycle, body_collection.go:253-296), and the
 * GetBody / ModBody handshakes (body_collection.go:111-191) as single-body reads and patches. */
typedef struct nb_body {
  double x, y, z, vx, vy, vz, m, r, rest, frag, frag_step;
  int32_t beh;
  uint8_t flags;
  uint8_t pad[3];
} nb_body;

/* Queue k bodies (arrays as in nb_load; frag and frag_step may be NULL) to be appended by the
 * next nb_cycle. */
int nb_add_bodies(nb_ctx *ctx, int64_t k, const double *x, const double *y, const double *z,
                  const double *vx, const double *vy, const double *vz, const double *m,
                  const double *r, const int32_t *beh, const uint8_t *flags, const double *frag,
                  const double *frag_step);
/* Cycle: drop the bodies that no longer exist (survivors keep their order), then append the
 * queued adds, newest first as the reference's event list holds them (body_collection.go:100,
 * 273-279), each with restitution R: first the fragments the device created since the last
 * cycle (events on), then the bodies queued by nb_add_bodies.  *n_out = the new body count. */
int nb_cycle(nb_ctx *ctx, double R, int64_t *n_out);
/* processDeletes: count > 0 removes bodies at the reference's stride walk over the array
 * (existing, unpinned; the walk removes one more than asked in some cases, as there); count = -1
 * removes every existing body; 0 does nothing.  Removed bodies get Exists = 0 and mass 0
 * (SetNotExists) and stay in the array until nb_cycle.  *removed = bodies removed. */
int nb_remove_bodies(nb_ctx *ctx, int64_t count, int64_t *removed);
/* One body's state (GetBody) and a patch of it (ModBody / ApplyMods). */
int nb_get_body(nb_ctx *ctx, int64_t index, nb_body *out);
int nb_set_body(nb_ctx *ctx, int64_t index, const nb_body *in);

/* ---- subsume and fragmentation events (nb_params.events = 1) -------------------------------- */
/* Detection dispatches each overlapping pair as Compute does (body.go:172-183): a collision when
 * both bodies are Elastic or Fragment, else a subsume when either is Subsume and the larger body
 * (radii unequal) reaches the other's centre.  Pairs are resolved once each, in ascending (i, j)
 * order: ResolveCollision as before, except that a true shouldFragment now starts fragmentation
 * (doFragment / initiateFragmentation, fragcalc.go:54-83, with the reference's constants:
 * fragment radius 1, at most 2000 fragments); ResolveSubsume (body.go:228-244) moves the smaller
 * body's mass to the larger and removes it (Exists = 0, mass 0), with no existence check, as
 * there.  Each cycle a fragmenting body creates up to 101 fragments (fragment, fragcalc.go:
 * 97-117: position uniform in the ball of 0.9 times its radius, its velocity, mass / fragments,
 * Elastic) into a device queue that nb_cycle appends; when none are left it stops existing.  A
 * fragmenting body keeps the force of the cycle it started fragmenting in, as the reference's
 * stale Body.fx does.  Fragment positions come from a counter-based generator (the reference
 * seeds from the clock), so runs are reproducible. */
/* Body.FragStep of every body (fragments per unit of impact factor above FragFactor,
 * fragcalc.go:72); NULL sets 0.  nb_load sets 0; nb_add_bodies and nb_set_body carry it. */
int nb_set_frag_steps(nb_ctx *ctx, const double *frag_step);

typedef struct nb_event_stats {
  int64_t subsumed_total;     /* ResolveSubsume calls carried out */
  int64_t fragmentations;     /* bodies that started fragmenting */
  int64_t fragments_emitted;  /* fragment bodies created */
  int64_t fragments_pending;  /* created, waiting for nb_cycle */
  int64_t fragmenting;        /* bodies fragmenting now */
  int32_t queue_overflow;     /* 1 if a cycle deferred fragments for lack of queue space */
  int32_t pad;
} nb_event_stats;
/* Totals since nb_load, as of the end of the last nb_step or nb_cycle (no device access). */
int nb_get_event_stats(nb_ctx *ctx, nb_event_stats *out);

/* ---- Barnes-Hut octree (the reference's own far-field design) ------------------------------ */
/* cmd/body/octree.go with Build wired into BuildTree: the reference allocates an empty tree each
 * cycle and never builds it (body_collection.go:298-300, octree.go:248-250), so its default
 * --barnes-hut mode computes no gravity.  When enabled on an FP64 context, nb_step and
 * nb_compute_forces compute Compute's force as Tree.CalcForce (body.go:156-157): the octree of
 * the existing, non-fragmenting bodies (Build, octree.go:136-194: leaf capacity 1, the
 * reference's root cell and centre arithmetic), accumForce's walk (octree.go:276-309) for every
 * body and directForce sums for suns (octree.go:252-267).  G comes from nb_params; eps and
 * overlap_cutoff are not used (the tree's softening applies).  Collision detection keeps the
 * reference's exact test (body.go:192-208), pruned by the tree.  Node centres of mass are summed
 * bottom-up rather than folded in insertion order, so forces agree with the reference's to
 * rounding (bit for bit at theta = 0). */
typedef struct nb_bh_params {
  double theta;      /* opening test (2*half/dist) < theta; Theta = 0.3 (octree.go:9) */
  double softening;  /* Softening = 1e-4 (octree.go:14) */
  int32_t max_depth; /* 1..21: a cell at this depth no longer splits; bodies closer than the
                        root cell / 2^max_depth share it (the reference recurses without end
                        for coincident bodies).  Default 21. */
  int32_t group;     /* bodies per wavefront in the tree walk: 16, 32 or 64; 0 = 64 */
  int32_t mode;      /* 0 = exact: each body walks the tree with the reference's test and FP64
                        terms in the reference's order; 1 = fast: the 64 bodies of a wavefront
                        (neighbours in key order), cut at tree-cell boundaries into groups that
                        each lie in a cell of fewer than 1,088 bodies, share one walk per group
                        that opens every node any of them would open (theta against the distance
                        to the group's bounding box), and the terms are summed in FP32 (the
                        body's own term excluded); forces then differ from exact mode by the
                        FP32 and opening differences */
  int32_t term;      /* exact mode's term: 0 = chosen by occupancy, 1 = three divisions by the
                        distance, 2 = one shared reciprocal for them (the compiler's division
                        steps, those that depend on the distance alone done once; bit-identical
                        results, fewer instructions per term, a longer dependency chain) */
  int32_t reserved[4];
} nb_bh_params;

typedef struct nb_bh_stats {
  int64_t builds;       /* trees built (force phases) */
  int64_t bodies;       /* bodies in the last tree */
  int64_t nodes;        /* its cells holding two or more bodies (split cells and buckets) */
  int64_t buckets;      /* cells at max_depth holding two or more bodies */
  int64_t interactions; /* accepted cells + body-body terms of the last force phase */
  int32_t depth_max;    /* deepest leaf */
  int32_t overflow;     /* nonzero: an internal capacity was exceeded (forces not valid) */
} nb_bh_stats;

int nb_bh_params_init(nb_bh_params *p);
/* NB_E_UNSUPPORTED on an FP32 context or with the surrogate on.  p may be NULL (defaults). */
int nb_bh_enable(nb_ctx *ctx, const nb_bh_params *p);
int nb_bh_disable(nb_ctx *ctx);
int nb_bh_get_stats(nb_ctx *ctx, nb_bh_stats *out);
/* Each body's leaf depth in the last tree (root = 0), -1 for bodies outside it. */
int nb_bh_get_depths(nb_ctx *ctx, int32_t *depth);

/* ---- batched path: many independent simulations in one persistent launch ------------------ */
/* No counterpart in the reference (one BodyCollection per process); for callers that run many
 * small simulations (parameter sweeps, ensembles).  Each simulation follows the per-cycle
 * semantics above with the softened FP32 force (eps > 0) and FP64 update; collisions and the
 * overlap cutoff are not part of this path (nb_batch_create returns NB_E_UNSUPPORTED). */
typedef struct nb_batch nb_batch;

int nb_batch_create(const nb_params *p, nb_batch **out);
void nb_batch_destroy(nb_batch *b);
const char *nb_batch_last_error(const nb_batch *b);

/* Simulation k owns bodies [offsets[k], offsets[k+1]); offsets has nsims + 1 entries, offsets[0]
 * = 0.  flags may be NULL (all existing). */
int nb_batch_load(nb_batch *b, int32_t nsims, const int64_t *offsets, const double *x,
                  const double *y, const double *z, const double *vx, const double *vy,
                  const double *vz, const double *m, const uint8_t *flags);

/* nsteps cycles of every simulation as one hipGraph launch (captured on first use for a given
 * (nsteps, dt, R), replayed afterwards). */
int nb_batch_step(nb_batch *b, double dt, double R, int32_t nsteps, nb_stats *out);

/* 1 = hipGraph launch (default), 0 = eager kernel launches (same kernels). */
int nb_batch_set_graph(nb_batch *b, int32_t enable);

/* Force on every body (unpadded body order), for tests. */
int nb_batch_forces(nb_batch *b, double *fx, double *fy, double *fz);

int nb_batch_store(nb_batch *b, double *x, double *y, double *z, double *vx, double *vy,
                   double *vz, uint8_t *flags);

/* Kernel shape of the batched force.  Work items are blocks of 64 x IPL targets of one
 * simulation; `waves` (1, 2 or 4) wavefronts share an item, splitting its sources; each wave
 * consumes its sources in Kahan tiles of `tile` (64, 128 or 256) sources.  IPL and tile are chosen
 * per size bucket (N < 1024, < 4096, >= 4096), waves for the whole launch.  When every bucket
 * has the same (IPL, tile) the launch runs a kernel specialised to that shape (fewer registers
 * and LDS, so more waves per SIMD); otherwise one kernel reads each item's shape.  Both give the
 * same forces bit for bit for a given (waves, tile).
 * nb_batch_autotune, for every waves choice, finds the fastest (IPL, tile) of each bucket alone,
 * then times full launches of those per-bucket winners together and of every single (IPL,
 * tile) for all buckets, and keeps the fastest full launch; ipl3 (may be NULL) receives the
 * per-bucket IPL.
 * Tuning words: cfg7 = {waves, ipl0, tile0, ipl1, tile1, ipl2, tile2}. */
int nb_batch_get_shapes(nb_batch *b, int32_t *ipl3);
int nb_batch_set_shapes(nb_batch *b, const int32_t *ipl3);
int nb_batch_get_tuning(nb_batch *b, int32_t *cfg7);
int nb_batch_set_tuning(nb_batch *b, const int32_t *cfg7);
int nb_batch_autotune(nb_batch *b, int32_t reps, int32_t *ipl3);

/* ---- small MLP on the matrix cores (the learned part of the far-field surrogate) ----------- */
/* No counterpart in the reference.  d_in -> hidden (x layers, GELU) -> d_out; BF16 weights and
 * activations on MFMA with FP32 accumulation; FP32 master weights and Adam state.  Parameters
 * are exchanged in the unpadded layout: per linear layer, W [in x out] row-major, then b [out]. */
typedef struct nb_mlp nb_mlp;

int nb_mlp_create(int32_t device, int32_t d_in, int32_t hidden, int32_t layers, int32_t d_out,
                  uint64_t seed, nb_mlp **out);
void nb_mlp_destroy(nb_mlp *m);
const char *nb_mlp_last_error(const nb_mlp *m);
int nb_mlp_param_count(const nb_mlp *m, int64_t *count);
int nb_mlp_set_params(nb_mlp *m, const float *params, int64_t count);
int nb_mlp_get_params(nb_mlp *m, float *params, int64_t count);
/* y [rows x d_out] = MLP(x [rows x d_in]) */
int nb_mlp_forward(nb_mlp *m, const float *x, int64_t rows, float *y);
/* The same outputs, bit for bit, from the fused one-kernel inference path the surrogate uses
 * (padded d_in and hidden <= 128, hidden a multiple of 64, d_out <= 16). */
int nb_mlp_infer(nb_mlp *m, const float *x, int64_t rows, float *y);
/* Gradient of the mean squared error over rows (a multiple of 64) w.r.t. the parameters. */
int nb_mlp_grad(nb_mlp *m, const float *x, const float *t, int64_t rows, float *grad,
                float *loss);
/* steps Adam updates on consecutive minibatches of batch rows (a multiple of 64). */
int nb_mlp_train(nb_mlp *m, const float *x, const float *t, int64_t rows, int32_t steps,
                 int32_t batch, float lr, float *loss_out);

/* ---- learned far-field surrogate for the softened force ------------------------------------ */
/* No counterpart in the reference: its sub-quadratic design, the Barnes-Hut octree
 * (cmd/body/octree.go), is never built.  When enabled on a context (FP32 path, eps > 0, no
 * overlap cutoff), nb_step and nb_compute_forces replace the O(N^2) force with an exact near
 * field over adjacent cells of a 32^3 grid in stretched coordinates u = d / (|d| + r0) about the
 * centre of mass, per-body quadrupoles of the cells at Chebyshev distance 2..shell, a
 * third-order Taylor far expansion, and an MLP residual; every check_every steps a sampled
 * direct-sum error above tol makes the step fall back to the direct sum.  With collisions on,
 * detection runs in the near field (exact FP64 test, the same pairs as the direct path); this
 * needs every radius <= r0 / grid, so that each overlapping pair lies in adjacent cells.
 * nb_stats.interactions still counts the N(N-1) pairs the step stands for. */
typedef struct nb_sur_params {
  int32_t grid;          /* cells per axis, 4..64 (32) */
  int32_t shell;         /* per-body multipole shell depth K, 1..6 (2) */
  double r0;             /* stretch radius; 0 = 1.5 x the half-mass radius at enable time */
  int32_t use_mlp;       /* apply the learned residual once trained (1) */
  int32_t check_every;   /* steps between sampled error checks; 0 = never (1) */
  int32_t check_samples; /* bodies per check (256) */
  int32_t pad0;
  double tol;            /* sampled relative error ||da|| / ||a|| above which the step is redone
                            (1e-3) */
  uint64_t seed;         /* training shuffles, initial weights and the check's sample draws */
  double tol_body;       /* per-body bar: the step is also redone when more than a fraction
                            1 - body_quantile of the samples have |da_i| / |a_i| above it (0 =
                            off; 0) */
  double body_quantile;  /* (0.99) */
  int32_t reserved[4];
} nb_sur_params;

typedef struct nb_sur_stats {
  int64_t steps;           /* surrogate force phases run */
  int64_t checks;          /* sampled error checks */
  int64_t fallbacks;       /* checks that redid the step with the direct sum */
  double last_error;       /* relative error of the last check */
  double near_pairs_total; /* exact near-field pairs evaluated so far */
  int64_t nonempty_cells;  /* at the last step */
  int64_t train_rows;      /* collected training rows */
  double train_loss;       /* last training loss (normalised units) */
  double r0;               /* stretch radius in use */
  int32_t trained;
  int32_t pad1;
  double last_body_error;  /* the body_quantile quantile of the per-body errors at the last check */
} nb_sur_stats;

int nb_sur_params_init(nb_sur_params *p);
int nb_sur_enable(nb_ctx *ctx, const nb_sur_params *p);
int nb_sur_disable(nb_ctx *ctx);
/* Append `samples` training rows from the loaded state (labels from the FP32 direct sum). */
int nb_sur_collect(nb_ctx *ctx, int64_t samples, uint64_t seed);
/* Fit the MLP on the collected rows with the device Adam (steps minibatches of batch rows). */
int nb_sur_fit(nb_ctx *ctx, int32_t steps, int32_t batch, float lr, float *loss);
/* Sampled relative force error ||da|| / ||a|| of the surrogate on the loaded state, with (1) or
 * without (0) the learned residual: one body from each of `samples` equal strata of the
 * cell-sorted order, the same draw at every call (so settings compare on the same bodies; the
 * checks inside steps draw afresh each time). */
int nb_sur_error(nb_ctx *ctx, int32_t samples, int32_t use_mlp, double *err);
int nb_sur_get_stats(nb_ctx *ctx, nb_sur_stats *out);
int nb_sur_save(nb_ctx *ctx, const char *path);
int nb_sur_load(nb_ctx *ctx, const char *path);

/* ---- force phase without device allocation ---------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  The force phase of
 * nb_compute_forces on the softened FP32 path (eps > 0, no overlap cutoff, no collisions) for
 * bodies the caller already holds on the device, with the kernels and operation order of a
 * context created with the same G and eps and the shape {ipl 1, tile_j 512, jsplit min(4, tiles),
 * block 256, variant 2}: the forces are bit-identical to that context's.
 *   x, y, z, m: n FP64 device arrays; flags: n device bytes (NB_F_*; NULL = every body exists);
 *   fx, fy, fz: n FP64 device outputs, Body.fx/fy/fz = m_i * G * sum_j m_j d / (d^2+eps^2)^1.5;
 *   stream: a hipStream_t (NULL = the null stream).
 * Asynchronous: kernels are enqueued on `stream` and the call returns without synchronising; no
 * hipMalloc, hipFree or other allocating routine is called.  The intermediate source stream and
 * per-slice partials live in a fixed device array of 28 MiB per device (held from library load
 * until the process ends), which limits n to NB_NOALLOC_MAX_BODIES; larger n returns
 * NB_E_INVALID.  The first call on a device resolves that array's address (make it before any
 * stream capture).  Calls on one device share the array: calls that overlap on one device race.
 * Errors: NB_E_INVALID / NB_E_HIP with the message in nb_last_error(NULL). */
#define NB_NOALLOC_MAX_BODIES 262144
int nb_forces_noalloc(int32_t device, int64_t n, const double *x, const double *y,
                      const double *z, const double *m, const uint8_t *flags, double G,
                      double eps, double *fx, double *fy, double *fz, void *stream);

#ifdef __cplusplus
}
#endif
#endif
