/*
 * nb_oracle.h -- CPU FP64 restatement of the aceeric/nbodygo per-cycle compute path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and never as the thing
 * measured or shipped.  The product (nbodygo_amd/libnbgpu.so) does not link it and has no CPU
 * fallback.
 *
 * Every function follows the reference Go source op by op (citations are to cmd/... files of the
 * reference).  It is compiled with -ffp-contract=off -fno-fast-math on SSE2 so each + - * / sqrt
 * is one correctly rounded IEEE double operation, as in a default amd64 Go build.
 *
 * Parity pins (see tests/test_oracle.py):
 *   or_force_ref      closed form of the TestWpCompute geometry, Newton's third law, long-double
 *                     brute force, zero net force at the centre of symmetric lattices.
 *   or_force_soft     long-double brute force, Plummer-sphere analytic field (statistical).
 *   or_detect         brute-force pair enumeration with boundary cases (d == Ri+Rj, d == 0).
 *   or_elastic        head-on swap (equal masses), 1-D unequal-mass formula, impulse-along-the-
 *                     contact-normal closed form for oblique impacts, momentum and energy
 *                     conservation, v == 0 and separating no-ops, TestCollide NaN geometry.
 *                     The last-ulp results of Go's math.Acos/Atan2/Sin/Cos/Asin/Tan are NOT
 *                     pinned (glibc is used here): parity on those is tolerance based.
 *   or_update         free drift closed form, one kick with a known force, NaN removal,
 *                     collided-skips-kick.
 *   or_energy         closed forms for two-body and lattice configurations.
 */
#ifndef NB_ORACLE_H
#define NB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Body flag bits (Body fields Exists, fragmenting, IsSun, Pinned: cmd/body/body.go:34-52). */
#define OR_F_EXISTS 1u
#define OR_F_FRAGMENTING 2u
#define OR_F_SUN 4u
#define OR_F_PINNED 8u

/* CollisionBehavior enum values (cmd/globals/globals.go:11-16). */
#define OR_CB_NONE 0
#define OR_CB_SUBSUME 1
#define OR_CB_ELASTIC 2
#define OR_CB_FRAGMENT 3

/* SoA view over caller-owned FP64 body state. */
typedef struct or_state {
  int64_t n;
  double *x, *y, *z, *vx, *vy, *vz, *m, *r;
  double *rest;       /* per-body coefficient of restitution (Body.r) */
  double *frag;       /* Body.FragFactor, may be NULL (= 0 for every body) */
  int32_t *beh;       /* CollisionBehavior */
  uint8_t *flags;     /* OR_F_* */
  uint8_t *collided;  /* Body.collided */
} or_state;

typedef struct or_params {
  double G;           /* 6.673e-11 in the reference (body.go:19) */
  double eps;         /* 0 = reference O(N^2) semantics (no softening) */
  int32_t cutoff;     /* 1 = zero force for d <= Ri+Rj (body.go:219) */
  int32_t collisions; /* 1 = detect + resolve elastic collisions */
  int32_t nthreads;   /* worker threads for the O(N^2) loops */
} or_params;

typedef struct or_stats {
  int64_t pairs_total;        /* detected unordered pairs summed over steps */
  int64_t pairs_last;         /* detected unordered pairs of the last step */
  int64_t resolved_total;     /* pairs for which calcElasticCollision reported a collision */
  int64_t fragment_requests;  /* shouldFragment() == true (fragmentation is out of scope) */
  int64_t removed_total;      /* bodies set !Exists by the NaN guard */
} or_stats;

/* Force on each target i from all j, reference semantics (body.go:148-170, :214-225).
 * idx == NULL means all bodies, else the nidx targets listed.  Outputs are force components
 * (not acceleration), indexed by position in idx (or by body when idx == NULL). */
void or_force_ref(const or_state *s, double G, const int64_t *idx, int64_t nidx, double *fx,
                  double *fy, double *fz, int nthreads);

/* Softened force (Plummer form of directForce, cmd/body/octree.go:314-325) with an optional
 * reference overlap gate on the unsoftened distance. */
void or_force_soft(const or_state *s, double G, double eps, int cutoff, const int64_t *idx,
                   int64_t nidx, double *fx, double *fy, double *fz, int nthreads);

/* The benchmark's stronger CPU line (cpu_soa.c): the same softened sum over SoA arrays,
 * vectorised and OpenMP-threaded, for targets [t0, t1); G applied; not the reference's
 * algorithm and not in the reference's operation order. */
void or_soa_soft(int64_t n, const double *x, const double *y, const double *z, const double *m,
                 double G, double eps, int64_t t0, int64_t t1, double *ax, double *ay,
                 double *az, int nthreads);

/* Same sums in long double (80-bit x87) for pinning the double versions. */
void or_force_soft_ld(const or_state *s, double G, double eps, int cutoff, const int64_t *idx,
                      int64_t nidx, double *fx, double *fy, double *fz);

/* Collision-candidate pairs, canonical order (i < j, lexicographic).  A pair is reported when
 * both bodies exist, are not fragmenting, both CollisionBehavior are Elastic or Fragment, and
 * sqrt(dx*dx+dy*dy+dz*dz) <= Ri+Rj (Collided, body.go:192-208; dispatch body.go:172-177).
 * Returns the number of pairs found (may exceed cap; only cap are written). */
int64_t or_detect(const or_state *s, int32_t *pairs, int64_t cap);

/* The same canonical pair list by cell lists, for large N: cells at least twice the largest
 * radius wide, each candidate decided by the same Collided test.  Returns the pair count (only
 * cap pairs are written), -2 for a non-finite position of a collidable body, -3 on allocation
 * failure. */
int64_t or_detect_grid(const or_state *s, int32_t *pairs, int64_t cap, int nthreads);

/* calcElasticCollision (cmd/body/collisioncalc.go:42-186) for body b1 (index i) and b2 (j).
 * out[0..8] = vx1,vy1,vz1,vx2,vy2,vz2,vx_cm,vy_cm,vz_cm.  Returns 1 if "collided". */
int or_elastic(const or_state *s, int64_t i, int64_t j, double *out);

/* ProcessMods over a canonical pair list: ResolveCollision (body.go:248-264) with the
 * shouldFragment gate (fragcalc.go:24-49) and doElastic (collisioncalc.go:26-35). */
void or_resolve(or_state *s, const int32_t *pairs, int64_t npairs, or_stats *st);

/* Body.Update (body.go:114-139) for all bodies with force arrays fx,fy,fz. */
void or_update(or_state *s, const double *fx, const double *fy, const double *fz, double dt,
               double R, or_stats *st);

/* One or more full cycles in the runner's phase order (computation-runner.go:267-326):
 * force (+ detection), ProcessMods, Update.  pairs_last receives the last step's pairs. */
int or_step(or_state *s, const or_params *p, double dt, double R, int nsteps, int32_t *pairs_last,
            int64_t cap, or_stats *st);

/* FP64 diagnostics: KE = sum m v^2 / 2, PE = -G sum_{i<j} mi mj / sqrt(d^2 + eps^2), momentum.
 * Existing bodies only; long double accumulation. out[0]=KE out[1]=PE out[2..4]=P. */
void or_energy(const or_state *s, double G, double eps, double *out, int nthreads);

/* The reference's CPU algorithm as it is structured (AoS bodies reached through pointers, a
 * persistent pool of W worker threads each computing a contiguous slice of targets with the
 * force loop and then the collision-detection loop of Body.Compute, body.go:148-187;
 * computation-runner.go:285-309; workpool.go:82-115).  Computes targets [0, ntargets) against
 * all bodies for nsteps cycles of the Compute phase and returns the elapsed wall seconds.
 * Logging and channel plumbing are excluded. */
double or_bench_workpool(const or_state *s, double G, int64_t ntargets, int nsteps,
                         int nthreads, double *fx_out, int64_t *hits_out);

/* Whole cycles of the same CPU algorithm, reference semantics (eps 0, overlap cutoff, elastic
 * collisions): per cycle the worker pool's Compute phase on AoS bodies (force loop, then the
 * detection loop keeping each overlapping Elastic/Fragment pair once), then the serial
 * resolution of the pairs in canonical order (or_resolve's arithmetic) and the serial Update
 * loop (body.go:114-139); computation-runner.go:294-320.  The state is advanced in place and
 * is bit-identical to or_step with eps 0, cutoff and collisions on.  Returns the elapsed wall
 * seconds of the nsteps cycles (negative on allocation failure). */
double or_bench_cycles(or_state *s, double G, double dt, double R, int nsteps, int nthreads,
                       or_stats *st);

/* Barnes-Hut octree of cmd/body/octree.go with Build wired into BuildTree (nb_oracle_bh.c):
 * the reference's tree and traversal op by op, plus a depth limit (a leaf at max_depth keeps
 * further bodies as a bucket).  Pinned by: theta = 0 against the softened direct sum (every
 * leaf visited, long double brute force), hand-built trees of two to four bodies, the root's
 * mass and centre of mass, one body per leaf below the limit, and the far-field limit of a
 * distant cluster (monopole closed form). */
typedef struct or_bh_params {
  double G;          /* 6.673e-11 in the reference (body.go:19) */
  double theta;      /* Theta = 0.3 (octree.go:9) */
  double soft;       /* Softening = 1e-4 (octree.go:14) */
  int32_t max_depth; /* deepest node that may hold bodies; deeper splits become buckets */
  int32_t pad;
} or_bh_params;

typedef struct or_bh_info {
  int64_t nodes;          /* pool nodes allocated (Octree.used) */
  int64_t internal;       /* subdivided nodes */
  int64_t buckets;        /* leaves holding more than one body (depth limit) */
  int64_t interactions;   /* accepted nodes + direct terms over all targets */
  int32_t max_depth_seen; /* deepest leaf holding a body */
  int32_t pad;
  double root_mass, root_half, root_cm[3];
} or_bh_info;

/* Compute's force with the tree (body.go:148-157; CalcForce, octree.go:247-272) for every body:
 * force components, zero for bodies that do not exist or are fragmenting.  depth (may be NULL)
 * receives each body's leaf depth (-1 if not in the tree).  Returns 0, or -1 on allocation
 * failure. */
int or_bh_force(const or_state *s, const or_bh_params *p, double *fx, double *fy, double *fz,
                int32_t *depth, or_bh_info *info);

#ifdef __cplusplus
}
#endif
#endif
