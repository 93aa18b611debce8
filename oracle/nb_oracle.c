/*
 * nb_oracle.c -- CPU FP64 restatement of the aceeric/nbodygo per-cycle compute path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See nb_oracle.h for the list of
 * checks that pin each function.  Build: oracle/Makefile (-O2 -ffp-contract=off -fno-fast-math).
 */
#define _GNU_SOURCE
#include "nb_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static int active(const or_state *s, int64_t i) {
  return (s->flags[i] & OR_F_EXISTS) && !(s->flags[i] & OR_F_FRAGMENTING);
}

/* ------------------------------------------------------------------------------------------ */
/* Thread fan-out over target ranges (the result of each target is a sequential sum over j in */
/* index order, so the output does not depend on the number of threads).                     */
/* ------------------------------------------------------------------------------------------ */
typedef void (*range_fn)(void *ctx, int64_t lo, int64_t hi);
typedef struct {
  range_fn fn;
  void *ctx;
  int64_t lo, hi;
} range_job;

static void *range_thread(void *a) {
  range_job *j = (range_job *)a;
  j->fn(j->ctx, j->lo, j->hi);
  return NULL;
}

static void parallel_for(int64_t n, int nthreads, range_fn fn, void *ctx) {
  if (nthreads <= 1 || n < 64) {
    fn(ctx, 0, n);
    return;
  }
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  range_job jobs[256];
  int64_t chunk = (n + nthreads - 1) / nthreads;
  int used = 0;
  for (int t = 0; t < nthreads; t++) {
    int64_t lo = (int64_t)t * chunk, hi = lo + chunk;
    if (lo >= n) break;
    if (hi > n) hi = n;
    jobs[t].fn = fn;
    jobs[t].ctx = ctx;
    jobs[t].lo = lo;
    jobs[t].hi = hi;
    pthread_create(&th[t], NULL, range_thread, &jobs[t]);
    used++;
  }
  for (int t = 0; t < used; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------------------------ */
/* Force, reference semantics                                                                  */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  const or_state *s;
  double G, eps;
  int cutoff;
  const int64_t *idx;
  double *fx, *fy, *fz;
} force_ctx;

/* Body.Compute's O(N^2) branch (body.go:156-170) calling calcForceFrom (body.go:214-225):
 *   dx = other.X - b.X ...; dist = Sqrt(dx*dx + dy*dy + dz*dz)
 *   if dist > b.Radius + other.Radius { force = G*b.Mass*other.Mass/(dist*dist);
 *                                       fx += force*dx/dist ... }
 * The b.collided arm of the gate is always false inside Compute because Update clears it
 * (body.go:128) before the next cycle, so it is omitted. */
static void force_ref_range(void *c, int64_t lo, int64_t hi) {
  const force_ctx *fc = (const force_ctx *)c;
  const or_state *s = fc->s;
  const double G = fc->G;
  for (int64_t k = lo; k < hi; k++) {
    int64_t i = fc->idx ? fc->idx[k] : k;
    double fx = 0, fy = 0, fz = 0;
    if (active(s, i)) {
      const double xi = s->x[i], yi = s->y[i], zi = s->z[i], mi = s->m[i], ri = s->r[i];
      for (int64_t j = 0; j < s->n; j++) {
        if (j == i || !active(s, j)) continue; /* body.go:162-165 */
        double dx = s->x[j] - xi;
        double dy = s->y[j] - yi;
        double dz = s->z[j] - zi;
        double dist = sqrt(dx * dx + dy * dy + dz * dz);
        if (dist > ri + s->r[j]) {
          double force = G * mi * s->m[j] / (dist * dist);
          fx += force * dx / dist;
          fy += force * dy / dist;
          fz += force * dz / dist;
        }
      }
    }
    fc->fx[k] = fx;
    fc->fy[k] = fy;
    fc->fz[k] = fz;
  }
}

void or_force_ref(const or_state *s, double G, const int64_t *idx, int64_t nidx, double *fx,
                  double *fy, double *fz, int nthreads) {
  force_ctx c = {s, G, 0.0, 1, idx, fx, fy, fz};
  parallel_for(idx ? nidx : s->n, nthreads, force_ref_range, &c);
}

/* Softened force, directForce form (octree.go:314-325):
 *   distSq = dx*dx + dy*dy + dz*dz + eps*eps; dist = Sqrt(distSq)
 *   force = G*a.Mass*b.Mass/distSq; fx += force*dx/dist ...
 * With cutoff, pairs whose unsoftened distance is <= Ri+Rj contribute nothing (body.go:219). */
static void force_soft_range(void *c, int64_t lo, int64_t hi) {
  const force_ctx *fc = (const force_ctx *)c;
  const or_state *s = fc->s;
  const double G = fc->G, eps2 = fc->eps * fc->eps;
  for (int64_t k = lo; k < hi; k++) {
    int64_t i = fc->idx ? fc->idx[k] : k;
    double fx = 0, fy = 0, fz = 0;
    if (active(s, i)) {
      const double xi = s->x[i], yi = s->y[i], zi = s->z[i], mi = s->m[i], ri = s->r[i];
      for (int64_t j = 0; j < s->n; j++) {
        if (j == i || !active(s, j)) continue;
        double dx = s->x[j] - xi;
        double dy = s->y[j] - yi;
        double dz = s->z[j] - zi;
        double d2 = dx * dx + dy * dy + dz * dz;
        if (fc->cutoff && sqrt(d2) <= ri + s->r[j]) continue;
        double distSq = d2 + eps2;
        double dist = sqrt(distSq);
        double force = G * mi * s->m[j] / distSq;
        fx += force * dx / dist;
        fy += force * dy / dist;
        fz += force * dz / dist;
      }
    }
    fc->fx[k] = fx;
    fc->fy[k] = fy;
    fc->fz[k] = fz;
  }
}

void or_force_soft(const or_state *s, double G, double eps, int cutoff, const int64_t *idx,
                   int64_t nidx, double *fx, double *fy, double *fz, int nthreads) {
  force_ctx c = {s, G, eps, cutoff, idx, fx, fy, fz};
  parallel_for(idx ? nidx : s->n, nthreads, force_soft_range, &c);
}

void or_force_soft_ld(const or_state *s, double G, double eps, int cutoff, const int64_t *idx,
                      int64_t nidx, double *fxo, double *fyo, double *fzo) {
  int64_t cnt = idx ? nidx : s->n;
  const long double eps2 = (long double)eps * (long double)eps;
  for (int64_t k = 0; k < cnt; k++) {
    int64_t i = idx ? idx[k] : k;
    long double fx = 0, fy = 0, fz = 0;
    if (active(s, i)) {
      for (int64_t j = 0; j < s->n; j++) {
        if (j == i || !active(s, j)) continue;
        long double dx = (long double)s->x[j] - s->x[i];
        long double dy = (long double)s->y[j] - s->y[i];
        long double dz = (long double)s->z[j] - s->z[i];
        long double d2 = dx * dx + dy * dy + dz * dz;
        if (cutoff && sqrtl(d2) <= (long double)s->r[i] + s->r[j]) continue;
        long double ds = d2 + eps2;
        long double f = (long double)G * s->m[i] * s->m[j] / (ds * sqrtl(ds));
        fx += f * dx;
        fy += f * dy;
        fz += f * dz;
      }
    }
    fxo[k] = (double)fx;
    fyo[k] = (double)fy;
    fzo[k] = (double)fz;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Collision detection                                                                         */
/* ------------------------------------------------------------------------------------------ */
static int can_collide(int32_t b) { return b == OR_CB_ELASTIC || b == OR_CB_FRAGMENT; }

/* Collided (body.go:192-208): dx = other.X - b.X ...; dist = Sqrt(dx*dx+dy*dy+dz*dz);
 * hit iff !(dist > b.Radius + other.Radius).  Evaluated from the lower index's side; the
 * computation is symmetric (IEEE subtraction is sign symmetric, addition commutative). */
static int collided(const or_state *s, int64_t i, int64_t j) {
  double dx = s->x[j] - s->x[i];
  double dy = s->y[j] - s->y[i];
  double dz = s->z[j] - s->z[i];
  double dist = sqrt(dx * dx + dy * dy + dz * dz);
  return !(dist > s->r[i] + s->r[j]);
}

int64_t or_detect(const or_state *s, int32_t *pairs, int64_t cap) {
  int64_t np = 0;
  for (int64_t i = 0; i < s->n; i++) {
    if (!active(s, i) || !can_collide(s->beh[i])) continue;
    for (int64_t j = i + 1; j < s->n; j++) {
      if (!active(s, j) || !can_collide(s->beh[j])) continue;
      if (collided(s, i, j)) {
        if (np < cap) {
          pairs[2 * np] = (int32_t)i;
          pairs[2 * np + 1] = (int32_t)j;
        }
        np++;
      }
    }
  }
  return np;
}

/* Cell-list form of or_detect for large N: bodies are binned on a cubic grid whose cells are at
 * least as wide as the largest sum of two radii, so every overlapping pair lies in the same or
 * adjacent cells; each candidate is then decided by the same Collided test.  The pair list is
 * the same canonical list (pinned against or_detect on the CPU). */
typedef struct {
  uint64_t key;
  int64_t idx;
} cell_entry;

static int cell_cmp(const void *a, const void *b) {
  const cell_entry *x = (const cell_entry *)a, *y = (const cell_entry *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

typedef struct {
  const or_state *s;
  const cell_entry *ent;
  int64_t nent;
  double lo[3], h;
  int64_t *cnt;     /* per body: pairs with a higher index */
  int64_t *off;     /* per body: first slot in out */
  int32_t *out;     /* 2 ints per pair, or NULL for the counting pass */
} grid_ctx;

static uint64_t cell_key(int64_t ix, int64_t iy, int64_t iz) {
  return ((uint64_t)ix << 42) | ((uint64_t)iy << 21) | (uint64_t)iz;
}

static int64_t lower_key(const cell_entry *e, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (e[mid].key < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

static int pair_cmp_j(const void *a, const void *b) {
  const int32_t *x = (const int32_t *)a, *y = (const int32_t *)b;
  return (x[1] > y[1]) - (x[1] < y[1]);
}

static void grid_range(void *c, int64_t lo_i, int64_t hi_i) {
  grid_ctx *g = (grid_ctx *)c;
  const or_state *s = g->s;
  for (int64_t i = lo_i; i < hi_i; i++) {
    if (!active(s, i) || !can_collide(s->beh[i])) continue;
    const int64_t cx = (int64_t)floor((s->x[i] - g->lo[0]) / g->h);
    const int64_t cy = (int64_t)floor((s->y[i] - g->lo[1]) / g->h);
    const int64_t cz = (int64_t)floor((s->z[i] - g->lo[2]) / g->h);
    int64_t k = 0, base = g->out ? g->off[i] : 0;
    for (int64_t dx = -1; dx <= 1; dx++)
      for (int64_t dy = -1; dy <= 1; dy++)
        for (int64_t dz = -1; dz <= 1; dz++) {
          const int64_t ax = cx + dx, ay = cy + dy, az = cz + dz;
          if (ax < 0 || ay < 0 || az < 0 || ax >= (1 << 21) || ay >= (1 << 21) || az >= (1 << 21))
            continue;
          const uint64_t key = cell_key(ax, ay, az);
          for (int64_t q = lower_key(g->ent, g->nent, key); q < g->nent && g->ent[q].key == key;
               q++) {
            const int64_t j = g->ent[q].idx;
            if (j <= i || !collided(s, i, j)) continue;
            if (g->out) {
              g->out[2 * (base + k)] = (int32_t)i;
              g->out[2 * (base + k) + 1] = (int32_t)j;
            }
            k++;
          }
        }
    if (g->out) qsort(g->out + 2 * base, (size_t)k, 2 * sizeof(int32_t), pair_cmp_j);
    else g->cnt[i] = k;
  }
}

int64_t or_detect_grid(const or_state *s, int32_t *pairs, int64_t cap, int nthreads) {
  const int64_t n = s->n;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double rmax = 0;
  int64_t na = 0;
  for (int64_t i = 0; i < n; i++) {
    if (!active(s, i) || !can_collide(s->beh[i])) continue;
    const double p[3] = {s->x[i], s->y[i], s->z[i]};
    if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]))) return -2;
    for (int d = 0; d < 3; d++) {
      if (p[d] < lo[d]) lo[d] = p[d];
      if (p[d] > hi[d]) hi[d] = p[d];
    }
    if (s->r[i] > rmax) rmax = s->r[i];
    na++;
  }
  if (na < 2) return 0;
  /* cells at least 2 rmax wide (every overlap spans at most one cell boundary per axis), with
   * the 1.0000001 margin covering the rounding of the cell index computation, and few enough
   * that an index fits 21 bits */
  double span = 0;
  for (int d = 0; d < 3; d++)
    if (hi[d] - lo[d] > span) span = hi[d] - lo[d];
  double h = 2.0 * rmax * 1.0000001;
  if (h < span / (double)((1 << 20))) h = span / (double)((1 << 20));
  if (!(h > 0)) h = 1.0;
  cell_entry *ent = (cell_entry *)malloc(sizeof(cell_entry) * (size_t)na);
  int64_t *cnt = (int64_t *)calloc((size_t)n, sizeof(int64_t));
  int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  if (!ent || !cnt || !off) {
    free(ent);
    free(cnt);
    free(off);
    return -3;
  }
  int64_t e = 0;
  for (int64_t i = 0; i < n; i++) {
    if (!active(s, i) || !can_collide(s->beh[i])) continue;
    ent[e].key = cell_key((int64_t)floor((s->x[i] - lo[0]) / h),
                          (int64_t)floor((s->y[i] - lo[1]) / h),
                          (int64_t)floor((s->z[i] - lo[2]) / h));
    ent[e].idx = i;
    e++;
  }
  qsort(ent, (size_t)na, sizeof(cell_entry), cell_cmp);
  grid_ctx g = {s, ent, na, {lo[0], lo[1], lo[2]}, h, cnt, off, NULL};
  parallel_for(n, nthreads, grid_range, &g);
  off[0] = 0;
  for (int64_t i = 0; i < n; i++) off[i + 1] = off[i] + cnt[i];
  const int64_t np = off[n];
  int32_t *all = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(np > 0 ? np : 1));
  if (!all) {
    free(ent);
    free(cnt);
    free(off);
    return -3;
  }
  g.out = all;
  parallel_for(n, nthreads, grid_range, &g);
  const int64_t c = np < cap ? np : cap;
  if (c > 0) memcpy(pairs, all, sizeof(int32_t) * 2 * (size_t)c);
  free(all);
  free(ent);
  free(cnt);
  free(off);
  return np;
}

/* ------------------------------------------------------------------------------------------ */
/* Elastic collision: calcElasticCollision, collisioncalc.go:42-186 (coll3d)                  */
/* ------------------------------------------------------------------------------------------ */
int or_elastic(const or_state *s, int64_t i, int64_t j, double *out) {
  double r12, m21, d, v, theta2, phi2, st, ct, sp, cp, vx1r, vy1r, vz1r, fvz1r, thetav, phiv, dr,
      alpha, beta, sbeta, cbeta, a, dvz2, vx2r, vy2r, vz2r, x21, y21, z21, vx21, vy21, vz21,
      vx_cm, vy_cm, vz_cm;
  double m1 = s->m[i], m2 = s->m[j];
  double r1 = s->r[i], r2 = s->r[j];
  double x1 = s->x[i], y1 = s->y[i], z1 = s->z[i];
  double x2 = s->x[j], y2 = s->y[j], z2 = s->z[j];
  double vx1 = s->vx[i], vy1 = s->vy[i], vz1 = s->vz[i];
  double vx2 = s->vx[j], vy2 = s->vy[j], vz2 = s->vz[j];

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

  if (v == 0) return 0; /* collisioncalc.go:89-92 */

  x2 = x21;
  y2 = y21;
  z2 = z21;
  vx1 = -vx21;
  vy1 = -vy21;
  vz1 = -vz21;

  theta2 = acos(z2 / d);
  if (x2 == 0 && y2 == 0) {
    phi2 = 0;
  } else {
    phi2 = atan2(y2, x2);
  }
  st = sin(theta2);
  ct = cos(theta2);
  sp = sin(phi2);
  cp = cos(phi2);

  vx1r = ct * cp * vx1 + ct * sp * vy1 - st * vz1;
  vy1r = cp * vy1 - sp * vx1;
  vz1r = st * cp * vx1 + st * sp * vy1 + ct * vz1;
  fvz1r = vz1r / v;
  if (fvz1r > 1) {
    fvz1r = 1;
  } else if (fvz1r < -1) {
    fvz1r = -1;
  }
  thetav = acos(fvz1r);
  if (vx1r == 0 && vy1r == 0) {
    phiv = 0;
  } else {
    phiv = atan2(vy1r, vx1r);
  }

  dr = d * sin(thetav) / r12;

  if (thetav > M_PI / 2 || fabs(dr) > 1) return 0; /* collisioncalc.go:137-140 */

  alpha = asin(-dr);
  beta = phiv;
  sbeta = sin(beta);
  cbeta = cos(beta);

  a = tan(thetav + alpha);
  dvz2 = 2 * (vz1r + a * (cbeta * vx1r + sbeta * vy1r)) / ((1 + a * a) * (1 + m21));

  vz2r = dvz2;
  vx2r = a * cbeta * dvz2;
  vy2r = a * sbeta * dvz2;
  vz1r = vz1r - m21 * vz2r;
  vx1r = vx1r - m21 * vx2r;
  vy1r = vy1r - m21 * vy2r;

  out[0] = ct * cp * vx1r - sp * vy1r + st * cp * vz1r + vx2;
  out[1] = ct * sp * vx1r + cp * vy1r + st * sp * vz1r + vy2;
  out[2] = ct * vz1r - st * vx1r + vz2;
  out[3] = ct * cp * vx2r - sp * vy2r + st * cp * vz2r + vx2;
  out[4] = ct * sp * vx2r + cp * vy2r + st * sp * vz2r + vy2;
  out[5] = ct * vz2r - st * vx2r + vz2;
  out[6] = vx_cm;
  out[7] = vy_cm;
  out[8] = vz_cm;
  return 1;
}

/* shouldFragment (fragcalc.go:24-49); r = or_elastic output for (b, other).  Fragmentation
 * itself (doFragment) is outside the ported path: a true result is counted by the caller and the
 * velocities are left unchanged. */
static int should_fragment(const or_state *s, int64_t b, int64_t o, const double *r) {
  if (!(s->beh[b] == OR_CB_FRAGMENT || s->beh[o] == OR_CB_FRAGMENT)) return 0;
  double br = s->rest[b];
  double vThis = s->vx[b] + s->vy[b] + s->vz[b];
  double dvThis = fabs(s->vx[b] - ((r[0] - r[6]) * br + r[6])) +
                  fabs(s->vy[b] - ((r[1] - r[7]) * br + r[7])) +
                  fabs(s->vz[b] - ((r[2] - r[8]) * br + r[8]));
  double thisFactor = dvThis / fabs(vThis);
  double vOther = s->vx[o] + s->vy[o] + s->vz[o];
  double dvOther = fabs(s->vx[o] - ((r[3] - r[6]) * br + r[6])) +
                   fabs(s->vy[o] - ((r[4] - r[7]) * br + r[7])) +
                   fabs(s->vz[o] - ((r[5] - r[8]) * br + r[8]));
  double otherFactor = dvOther / fabs(vOther);
  double fb = s->frag ? s->frag[b] : 0.0, fo = s->frag ? s->frag[o] : 0.0;
  return (s->beh[b] == OR_CB_FRAGMENT && thisFactor > fb) ||
         (s->beh[o] == OR_CB_FRAGMENT && otherFactor > fo);
}

void or_resolve(or_state *s, const int32_t *pairs, int64_t npairs, or_stats *st) {
  double r[9];
  for (int64_t k = 0; k < npairs; k++) {
    int64_t lo = pairs[2 * k], hi = pairs[2 * k + 1];
    int64_t b1, b2;
    /* Only an Elastic first body resolves (body.go:252-253): the canonical pair is resolved
     * once, as lo.ResolveCollision(hi) when lo is Elastic, else hi.ResolveCollision(lo). */
    if (s->beh[lo] == OR_CB_ELASTIC && can_collide(s->beh[hi])) {
      b1 = lo;
      b2 = hi;
    } else if (s->beh[hi] == OR_CB_ELASTIC && can_collide(s->beh[lo])) {
      b1 = hi;
      b2 = lo;
    } else {
      continue;
    }
    if (!(s->flags[b1] & OR_F_EXISTS) || !(s->flags[b2] & OR_F_EXISTS)) continue; /* :249-251 */
    if (!or_elastic(s, b1, b2, r)) continue;
    if (st) st->resolved_total++;
    if (should_fragment(s, b1, b2, r)) {
      if (st) st->fragment_requests++;
      continue;
    }
    /* doElastic (collisioncalc.go:26-35): restitution of the first body. */
    double rr = s->rest[b1];
    s->vx[b1] = (r[0] - r[6]) * rr + r[6];
    s->vy[b1] = (r[1] - r[7]) * rr + r[7];
    s->vz[b1] = (r[2] - r[8]) * rr + r[8];
    s->vx[b2] = (r[3] - r[6]) * rr + r[6];
    s->vy[b2] = (r[4] - r[7]) * rr + r[7];
    s->vz[b2] = (r[5] - r[8]) * rr + r[8];
    s->collided[b1] = 1;
    s->collided[b2] = 1;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Update: Body.Update, body.go:114-139                                                        */
/* ------------------------------------------------------------------------------------------ */
void or_update(or_state *s, const double *fx, const double *fy, const double *fz, double dt,
               double R, or_stats *st) {
  for (int64_t i = 0; i < s->n; i++) {
    if (!(s->flags[i] & OR_F_EXISTS)) continue;
    if (!s->collided[i]) {
      s->vx[i] += dt * fx[i] / s->m[i];
      s->vy[i] += dt * fy[i] / s->m[i];
      s->vz[i] += dt * fz[i] / s->m[i];
    }
    s->x[i] += dt * s->vx[i];
    s->y[i] += dt * s->vy[i];
    s->z[i] += dt * s->vz[i];
    s->collided[i] = 0;
    s->rest[i] = R;
    if (isnan(s->x[i]) || isnan(s->y[i]) || isnan(s->z[i])) {
      s->flags[i] &= (uint8_t)~OR_F_EXISTS;
      if (st) st->removed_total++;
    }
  }
}

int or_step(or_state *s, const or_params *p, double dt, double R, int nsteps, int32_t *pairs_last,
            int64_t cap, or_stats *st) {
  int64_t n = s->n;
  double *f = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  if (!f) return -3;
  int64_t pcap = cap > 0 ? cap : 0;
  int32_t *pairs = NULL;
  int64_t work_cap = 0;
  int rc = 0;
  for (int step = 0; step < nsteps; step++) {
    /* Compute phase (Body.Compute for every body): force, then detection. */
    if (p->eps == 0.0 && p->cutoff)
      or_force_ref(s, p->G, NULL, 0, f, f + n, f + 2 * n, p->nthreads);
    else
      or_force_soft(s, p->G, p->eps, p->cutoff, NULL, 0, f, f + n, f + 2 * n, p->nthreads);
    int64_t np = 0;
    if (p->collisions) {
      np = or_detect(s, pairs, work_cap);
      if (np > work_cap) {
        free(pairs);
        work_cap = np;
        pairs = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)work_cap);
        if (!pairs) {
          rc = -3;
          break;
        }
        np = or_detect(s, pairs, work_cap);
      }
      /* ProcessMods */
      or_resolve(s, pairs, np, st);
    }
    if (st) {
      st->pairs_total += np;
      st->pairs_last = np;
    }
    if (step == nsteps - 1 && pairs_last) {
      int64_t c = np < pcap ? np : pcap;
      if (c) memcpy(pairs_last, pairs, sizeof(int32_t) * 2 * (size_t)c);
    }
    /* Update loop */
    or_update(s, f, f + n, f + 2 * n, dt, R, st);
  }
  free(pairs);
  free(f);
  return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* Diagnostics                                                                                 */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  const or_state *s;
  double G, eps;
  long double *pe; /* per-target partial */
} energy_ctx;

static void energy_range(void *c, int64_t lo, int64_t hi) {
  energy_ctx *ec = (energy_ctx *)c;
  const or_state *s = ec->s;
  const long double eps2 = (long double)ec->eps * ec->eps;
  for (int64_t i = lo; i < hi; i++) {
    long double pe = 0;
    if (s->flags[i] & OR_F_EXISTS) {
      for (int64_t j = i + 1; j < s->n; j++) {
        if (!(s->flags[j] & OR_F_EXISTS)) continue;
        long double dx = (long double)s->x[j] - s->x[i];
        long double dy = (long double)s->y[j] - s->y[i];
        long double dz = (long double)s->z[j] - s->z[i];
        long double d2 = dx * dx + dy * dy + dz * dz + eps2;
        if (d2 > 0) pe -= (long double)s->m[i] * s->m[j] / sqrtl(d2);
      }
    }
    ec->pe[i] = pe;
  }
}

void or_energy(const or_state *s, double G, double eps, double *out, int nthreads) {
  long double ke = 0, px = 0, py = 0, pz = 0, pe = 0;
  long double *pei = (long double *)calloc((size_t)(s->n > 0 ? s->n : 1), sizeof(long double));
  energy_ctx ec = {s, G, eps, pei};
  parallel_for(s->n, nthreads, energy_range, &ec);
  for (int64_t i = 0; i < s->n; i++) {
    pe += pei[i];
    if (!(s->flags[i] & OR_F_EXISTS)) continue;
    long double m = s->m[i];
    ke += 0.5L * m *
          ((long double)s->vx[i] * s->vx[i] + (long double)s->vy[i] * s->vy[i] +
           (long double)s->vz[i] * s->vz[i]);
    px += m * s->vx[i];
    py += m * s->vy[i];
    pz += m * s->vz[i];
  }
  free(pei);
  out[0] = (double)ke;
  out[1] = (double)(pe * (long double)G);
  out[2] = (double)px;
  out[3] = (double)py;
  out[4] = (double)pz;
}

/* ------------------------------------------------------------------------------------------ */
/* The reference's CPU work pool (timing baseline)                                             */
/* ------------------------------------------------------------------------------------------ */
/* AoS body as in cmd/body/body.go:34-52 (the fields the cycle touches). */
typedef struct wp_body {
  double X, Y, Z, Vx, Vy, Vz, Radius, Mass;
  double fx, fy, fz, r;
  int32_t beh;
  uint8_t exists, fragmenting, collided, pad;
  char name_class_padding[64]; /* Id/Name/Class/colour/... occupy the rest of the struct */
} wp_body;

typedef struct {
  wp_body **arr;
  int64_t n;
  double G;
  int64_t lo, hi;
  int64_t hits;
  int nsteps;
  pthread_barrier_t *start, *done;
} wp_worker;

/* Body.Compute (body.go:148-187) on one body: the force loop over IterateOnce, then the
 * collision-detection loop over IterateOnce (every body, self included). */
static int64_t wp_compute(wp_body *b, wp_body **arr, int64_t n, double G) {
  int64_t hits = 0;
  if (!b->exists) return 0;
  b->fx = b->fy = b->fz = 0;
  for (int64_t k = 0; k < n; k++) {
    wp_body *o = arr[k];
    if (!o->exists || b->fragmenting) continue;
    if (b != o && o->exists && !o->fragmenting) {
      double dx = o->X - b->X;
      double dy = o->Y - b->Y;
      double dz = o->Z - b->Z;
      double dist = sqrt(dx * dx + dy * dy + dz * dz);
      if (b->collided || dist > b->Radius + o->Radius) {
        double force = G * b->Mass * o->Mass / (dist * dist);
        b->fx += force * dx / dist;
        b->fy += force * dy / dist;
        b->fz += force * dz / dist;
      }
    }
  }
  for (int64_t k = 0; k < n; k++) {
    wp_body *o = arr[k];
    if (b->collided) continue;
    double dx = o->X - b->X;
    double dy = o->Y - b->Y;
    double dz = o->Z - b->Z;
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(dist > b->Radius + o->Radius)) {
      if ((b->beh == OR_CB_ELASTIC || b->beh == OR_CB_FRAGMENT) &&
          (o->beh == OR_CB_ELASTIC || o->beh == OR_CB_FRAGMENT))
        hits++; /* the reference enqueues newCollision(b, o) here */
    }
  }
  return hits;
}

static void *wp_thread(void *a) {
  wp_worker *w = (wp_worker *)a;
  for (int s = 0; s < w->nsteps; s++) {
    pthread_barrier_wait(w->start);
    int64_t h = 0;
    for (int64_t i = w->lo; i < w->hi; i++) h += wp_compute(w->arr[i], w->arr, w->n, w->G);
    w->hits += h;
    pthread_barrier_wait(w->done);
  }
  return NULL;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + ts.tv_nsec * 1e-9;
}

double or_bench_workpool(const or_state *s, double G, int64_t ntargets, int nsteps, int nthreads,
                         double *fx_out, int64_t *hits_out) {
  int64_t n = s->n;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (ntargets > n) ntargets = n;
  /* Heap-allocated bodies reached through a pointer array, like []*Body. */
  wp_body *store = (wp_body *)calloc((size_t)n, sizeof(wp_body));
  wp_body **arr = (wp_body **)malloc(sizeof(wp_body *) * (size_t)n);
  for (int64_t i = 0; i < n; i++) {
    wp_body *b = &store[i];
    b->X = s->x[i];
    b->Y = s->y[i];
    b->Z = s->z[i];
    b->Vx = s->vx[i];
    b->Vy = s->vy[i];
    b->Vz = s->vz[i];
    b->Radius = s->r[i];
    b->Mass = s->m[i];
    b->r = s->rest[i];
    b->beh = s->beh[i];
    b->exists = (s->flags[i] & OR_F_EXISTS) ? 1 : 0;
    b->fragmenting = (s->flags[i] & OR_F_FRAGMENTING) ? 1 : 0;
    arr[i] = b;
  }
  /* Slice sizing of runOneComputation (computation-runner.go:285-305): size = N/W, or one
   * slice when N < 100; the last slice absorbs the remainder. */
  int64_t size = ntargets / nthreads;
  if (ntargets < 100 || size == 0) {
    size = ntargets;
  }
  int nw = (int)((ntargets + size - 1) / size);
  if (nw < 1) nw = 1;
  pthread_barrier_t start, done;
  pthread_barrier_init(&start, NULL, (unsigned)nw + 1);
  pthread_barrier_init(&done, NULL, (unsigned)nw + 1);
  wp_worker *ws = (wp_worker *)calloc((size_t)nw, sizeof(wp_worker));
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nw);
  for (int t = 0; t < nw; t++) {
    ws[t].arr = arr;
    ws[t].n = n;
    ws[t].G = G;
    ws[t].lo = (int64_t)t * size;
    ws[t].hi = ws[t].lo + size > ntargets ? ntargets : ws[t].lo + size;
    ws[t].nsteps = nsteps;
    ws[t].start = &start;
    ws[t].done = &done;
    pthread_create(&th[t], NULL, wp_thread, &ws[t]);
  }
  double t0 = now_s();
  for (int st = 0; st < nsteps; st++) {
    pthread_barrier_wait(&start);
    pthread_barrier_wait(&done);
  }
  double el = now_s() - t0;
  int64_t hits = 0;
  for (int t = 0; t < nw; t++) {
    pthread_join(th[t], NULL);
    hits += ws[t].hits;
  }
  if (fx_out)
    for (int64_t i = 0; i < ntargets; i++) {
      fx_out[3 * i] = arr[i]->fx;
      fx_out[3 * i + 1] = arr[i]->fy;
      fx_out[3 * i + 2] = arr[i]->fz;
    }
  if (hits_out) *hits_out = hits;
  pthread_barrier_destroy(&start);
  pthread_barrier_destroy(&done);
  free(th);
  free(ws);
  free(arr);
  free(store);
  return el;
}

/* ------------------------------------------------------------------------------------------ */
/* Whole cycles of the reference's CPU algorithm (timing baseline for the 1,000-body scene)    */
/* ------------------------------------------------------------------------------------------ */
/* One worker of the cycle pool: Body.Compute (body.go:148-187) on its slice of targets, the
 * force loop then the detection loop over IterateOnce.  Where the reference enqueues
 * newCollision(b, o) for an overlapping Elastic/Fragment pair (body.go:172-177) the worker keeps
 * the unordered pair once, from its lower index, for the canonical resolution. */
typedef struct {
  wp_body **arr;
  int64_t n;
  double G;
  int64_t lo, hi;
  int32_t *pairs;
  int64_t np, cap;
  int nsteps, fail;
  pthread_barrier_t *start, *done;
} cy_worker;

static void cy_push(cy_worker *w, int64_t i, int64_t j) {
  if (w->np == w->cap) {
    int64_t c = w->cap ? 2 * w->cap : 64;
    int32_t *p = (int32_t *)realloc(w->pairs, sizeof(int32_t) * 2 * (size_t)c);
    if (!p) {
      w->fail = 1;
      return;
    }
    w->pairs = p;
    w->cap = c;
  }
  w->pairs[2 * w->np] = (int32_t)i;
  w->pairs[2 * w->np + 1] = (int32_t)j;
  w->np++;
}

static void cy_compute(cy_worker *w, int64_t i) {
  wp_body *b = w->arr[i];
  wp_body **arr = w->arr;
  const int64_t n = w->n;
  const double G = w->G;
  b->fx = b->fy = b->fz = 0;
  if (!b->exists || b->fragmenting) return;
  for (int64_t k = 0; k < n; k++) { /* calcForceFrom, body.go:214-225 */
    wp_body *o = arr[k];
    if (b == o || !o->exists || o->fragmenting) continue;
    double dx = o->X - b->X;
    double dy = o->Y - b->Y;
    double dz = o->Z - b->Z;
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (b->collided || dist > b->Radius + o->Radius) {
      double force = G * b->Mass * o->Mass / (dist * dist);
      b->fx += force * dx / dist;
      b->fy += force * dy / dist;
      b->fz += force * dz / dist;
    }
  }
  if (b->collided || !can_collide(b->beh)) return;
  for (int64_t k = i + 1; k < n; k++) { /* Collided, body.go:192-208 */
    wp_body *o = arr[k];
    if (!o->exists || o->fragmenting || !can_collide(o->beh)) continue;
    double dx = o->X - b->X;
    double dy = o->Y - b->Y;
    double dz = o->Z - b->Z;
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(dist > b->Radius + o->Radius)) cy_push(w, i, k);
  }
}

static void *cy_thread(void *a) {
  cy_worker *w = (cy_worker *)a;
  for (int s = 0; s < w->nsteps; s++) {
    pthread_barrier_wait(w->start);
    w->np = 0;
    for (int64_t i = w->lo; i < w->hi; i++) cy_compute(w, i);
    pthread_barrier_wait(w->done);
  }
  return NULL;
}

/* ProcessMods on the AoS bodies: each canonical pair resolved by or_resolve on a two-body view,
 * so the arithmetic is the oracle's own. */
static void cy_resolve(wp_body **arr, const int32_t *pairs, int64_t np, or_stats *st) {
  for (int64_t k = 0; k < np; k++) {
    wp_body *b[2] = {arr[pairs[2 * k]], arr[pairs[2 * k + 1]]};
    double x[2], y[2], z[2], vx[2], vy[2], vz[2], m[2], r[2], rest[2], frag[2] = {0, 0};
    int32_t beh[2];
    uint8_t flags[2], col[2];
    for (int q = 0; q < 2; q++) {
      x[q] = b[q]->X; y[q] = b[q]->Y; z[q] = b[q]->Z;
      vx[q] = b[q]->Vx; vy[q] = b[q]->Vy; vz[q] = b[q]->Vz;
      m[q] = b[q]->Mass; r[q] = b[q]->Radius; rest[q] = b[q]->r;
      beh[q] = b[q]->beh;
      flags[q] = (uint8_t)((b[q]->exists ? OR_F_EXISTS : 0) | (b[q]->fragmenting ? OR_F_FRAGMENTING : 0));
      col[q] = b[q]->collided;
    }
    or_state v = {2, x, y, z, vx, vy, vz, m, r, rest, frag, beh, flags, col};
    const int32_t p01[2] = {0, 1};
    or_resolve(&v, p01, 1, st);
    for (int q = 0; q < 2; q++) {
      b[q]->Vx = vx[q]; b[q]->Vy = vy[q]; b[q]->Vz = vz[q];
      b[q]->collided = col[q];
    }
  }
}

double or_bench_cycles(or_state *s, double G, double dt, double R, int nsteps, int nthreads,
                       or_stats *st) {
  int64_t n = s->n;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  wp_body *store = (wp_body *)calloc((size_t)(n > 0 ? n : 1), sizeof(wp_body));
  wp_body **arr = (wp_body **)malloc(sizeof(wp_body *) * (size_t)(n > 0 ? n : 1));
  if (!store || !arr) {
    free(store);
    free(arr);
    return -1.0;
  }
  for (int64_t i = 0; i < n; i++) {
    wp_body *b = &store[i];
    b->X = s->x[i]; b->Y = s->y[i]; b->Z = s->z[i];
    b->Vx = s->vx[i]; b->Vy = s->vy[i]; b->Vz = s->vz[i];
    b->Radius = s->r[i];
    b->Mass = s->m[i];
    b->r = s->rest[i];
    b->beh = s->beh[i];
    b->exists = (s->flags[i] & OR_F_EXISTS) ? 1 : 0;
    b->fragmenting = (s->flags[i] & OR_F_FRAGMENTING) ? 1 : 0;
    b->collided = s->collided ? s->collided[i] : 0;
    arr[i] = b;
  }
  /* slices as runOneComputation sizes them (computation-runner.go:285-305) */
  int64_t size = n / nthreads;
  if (n < 100 || size == 0) size = n > 0 ? n : 1;
  int nw = (int)((n + size - 1) / size);
  if (nw < 1) nw = 1;
  pthread_barrier_t start, done;
  pthread_barrier_init(&start, NULL, (unsigned)nw + 1);
  pthread_barrier_init(&done, NULL, (unsigned)nw + 1);
  cy_worker *ws = (cy_worker *)calloc((size_t)nw, sizeof(cy_worker));
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nw);
  for (int t = 0; t < nw; t++) {
    ws[t].arr = arr;
    ws[t].n = n;
    ws[t].G = G;
    ws[t].lo = (int64_t)t * size;
    ws[t].hi = ws[t].lo + size > n ? n : ws[t].lo + size;
    ws[t].nsteps = nsteps;
    ws[t].start = &start;
    ws[t].done = &done;
    pthread_create(&th[t], NULL, cy_thread, &ws[t]);
  }
  double t0 = now_s();
  for (int step = 0; step < nsteps; step++) {
    pthread_barrier_wait(&start); /* submitSlice for every slice */
    pthread_barrier_wait(&done);  /* wait */
    int64_t np = 0;
    for (int t = 0; t < nw; t++) { /* slices in order: canonical (i, j) order */
      cy_resolve(arr, ws[t].pairs, ws[t].np, st);
      np += ws[t].np;
    }
    if (st) {
      st->pairs_total += np;
      st->pairs_last = np;
    }
    for (int64_t i = 0; i < n; i++) { /* the Update loop, body.go:114-139 */
      wp_body *b = arr[i];
      if (!b->exists) continue;
      if (!b->collided) {
        b->Vx += dt * b->fx / b->Mass;
        b->Vy += dt * b->fy / b->Mass;
        b->Vz += dt * b->fz / b->Mass;
      }
      b->X += dt * b->Vx;
      b->Y += dt * b->Vy;
      b->Z += dt * b->Vz;
      b->collided = 0;
      b->r = R;
      if (isnan(b->X) || isnan(b->Y) || isnan(b->Z)) {
        b->exists = 0;
        if (st) st->removed_total++;
      }
    }
  }
  double el = now_s() - t0;
  int fail = 0;
  for (int t = 0; t < nw; t++) {
    pthread_join(th[t], NULL);
    fail |= ws[t].fail;
    free(ws[t].pairs);
  }
  for (int64_t i = 0; i < n; i++) {
    const wp_body *b = arr[i];
    s->x[i] = b->X; s->y[i] = b->Y; s->z[i] = b->Z;
    s->vx[i] = b->Vx; s->vy[i] = b->Vy; s->vz[i] = b->Vz;
    s->rest[i] = b->r;
    s->flags[i] = (uint8_t)((s->flags[i] & ~OR_F_EXISTS) | (b->exists ? OR_F_EXISTS : 0));
    if (s->collided) s->collided[i] = b->collided;
  }
  pthread_barrier_destroy(&start);
  pthread_barrier_destroy(&done);
  free(th);
  free(ws);
  free(arr);
  free(store);
  return fail ? -1.0 : el;
}
