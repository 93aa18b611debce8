/*
 * nb_oracle_bh.c -- CPU FP64 restatement of the reference's Barnes-Hut octree
 * (cmd/body/octree.go), TEST INFRASTRUCTURE ONLY (see nb_oracle.h).
 *
 * The reference never builds its tree (BuildTree, body_collection.go:298-300, only allocates an
 * empty Octree), so its default mode computes no gravity.  This file states the "fixed" wiring:
 * BuildTree = NewOctree() + Build(bc.arr) before the workers run, and Compute's force is
 * Tree.CalcForce (body.go:156-157).  Everything else follows octree.go op by op:
 *   Build          :136-194  tight box over existing, non-fragmenting bodies; square root cell of
 *                            half = max extent / 2 * 1.05; bodies inserted in array order
 *   insert         :198-217  updateMass on every node passed, leaf capacity 1, subdivide and
 *                            re-insert the resident body first
 *   updateMass     :91-97    incremental mass and centre of mass
 *   octantOf / childCenter / subdivide / insertIntoChild  :55-67, :70-88, :220-231, :234-237
 *   CalcForce      :247-272  suns: directForce from every other existing, non-fragmenting body
 *                            in array order; others: accumForce from the root
 *   accumForce     :276-309  softened distance; leaf or (2*half/dist) < Theta accepts the node
 *                            (a leaf holding b itself is skipped); children in index order,
 *                            those with totalMass > 0
 *   directForce    :314-325
 *
 * One addition the reference lacks: a depth limit.  A leaf at depth max_depth (the root is
 * depth 0) no longer subdivides; later bodies join it as a bucket (kept in insertion order).  An
 * accepted bucket acts as its bodies one by one (directForce arithmetic, b itself skipped).  The
 * reference recurses without limit and overflows its stack for two bodies at one position.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "nb_oracle.h"

typedef struct bh_node {
  double cx, cy, cz, half;
  double mass, cmx, cmy, cmz;
  int64_t body;        /* leaf body, -1 if none */
  int64_t bucket_tail; /* last body of the bucket list (max-depth leaf), -1 if none */
  int32_t child[8];    /* -1: no children (leaf) */
  int32_t depth;
} bh_node;

typedef struct bh_tree {
  bh_node *pool;
  int64_t used, cap;
  int64_t *next; /* bucket list links, per body */
  int32_t max_depth;
  const or_state *s;
} bh_tree;

static int32_t bh_alloc(bh_tree *t) {
  if (t->used >= t->cap) {
    t->cap = t->cap ? 2 * t->cap : 16384;
    t->pool = (bh_node *)realloc(t->pool, (size_t)t->cap * sizeof(bh_node));
  }
  bh_node *n = &t->pool[t->used];
  memset(n, 0, sizeof *n);
  n->body = -1;
  n->bucket_tail = -1;
  for (int k = 0; k < 8; k++) n->child[k] = -1;
  return (int32_t)t->used++;
}

/* updateMass (octree.go:91-97) */
static void bh_update_mass(bh_node *n, double mass, double x, double y, double z) {
  const double nt = n->mass + mass;
  n->cmx = (n->cmx * n->mass + x * mass) / nt;
  n->cmy = (n->cmy * n->mass + y * mass) / nt;
  n->cmz = (n->cmz * n->mass + z * mass) / nt;
  n->mass = nt;
}

/* octantOf (octree.go:55-67) */
static int bh_octant(const bh_node *n, double x, double y, double z) {
  int idx = 0;
  if (x >= n->cx) idx |= 1;
  if (y >= n->cy) idx |= 2;
  if (z >= n->cz) idx |= 4;
  return idx;
}

/* subdivide (octree.go:220-231) with childCenter (:70-88) */
static void bh_subdivide(bh_tree *t, int32_t ni) {
  for (int i = 0; i < 8; i++) {
    const int32_t c = bh_alloc(t); /* may move the pool: index, then re-read */
    bh_node *n = &t->pool[ni];
    bh_node *ch = &t->pool[c];
    const double q = n->half / 2;
    ch->cx = (i & 1) ? n->cx + q : n->cx - q;
    ch->cy = (i & 2) ? n->cy + q : n->cy - q;
    ch->cz = (i & 4) ? n->cz + q : n->cz - q;
    ch->half = n->half / 2;
    ch->depth = n->depth + 1;
    n->child[i] = c;
  }
}

static void bh_insert(bh_tree *t, int32_t ni, int64_t b);

static void bh_insert_into_child(bh_tree *t, int32_t ni, int64_t b) {
  const or_state *s = t->s;
  const bh_node *n = &t->pool[ni];
  const int32_t c = n->child[bh_octant(n, s->x[b], s->y[b], s->z[b])];
  bh_insert(t, c, b);
}

/* insert (octree.go:198-217), with the depth limit */
static void bh_insert(bh_tree *t, int32_t ni, int64_t b) {
  const or_state *s = t->s;
  bh_node *n = &t->pool[ni];
  bh_update_mass(n, s->m[b], s->x[b], s->y[b], s->z[b]);
  if (n->child[0] < 0) {
    if (n->body < 0) {
      n->body = b;
      return;
    }
    if (n->depth >= t->max_depth) { /* bucket: append in insertion order */
      t->next[b] = -1;
      if (n->bucket_tail < 0)
        t->next[n->body] = b;
      else
        t->next[n->bucket_tail] = b;
      n->bucket_tail = b;
      return;
    }
    const int64_t existing = n->body;
    n->body = -1;
    bh_subdivide(t, ni);
    bh_insert_into_child(t, ni, existing);
    bh_insert_into_child(t, ni, b);
    return;
  }
  bh_insert_into_child(t, ni, b);
}

static int bh_active(const or_state *s, int64_t i) {
  return (s->flags[i] & OR_F_EXISTS) && !(s->flags[i] & OR_F_FRAGMENTING);
}

/* Build (octree.go:136-194).  Returns the root index or -1. */
static int32_t bh_build(bh_tree *t) {
  const or_state *s = t->s;
  double minX = DBL_MAX, minY = DBL_MAX, minZ = DBL_MAX;
  double maxX = -DBL_MAX, maxY = -DBL_MAX, maxZ = -DBL_MAX;
  int64_t count = 0;
  for (int64_t i = 0; i < s->n; i++) {
    if (!bh_active(s, i)) continue;
    if (s->x[i] < minX) minX = s->x[i];
    if (s->y[i] < minY) minY = s->y[i];
    if (s->z[i] < minZ) minZ = s->z[i];
    if (s->x[i] > maxX) maxX = s->x[i];
    if (s->y[i] > maxY) maxY = s->y[i];
    if (s->z[i] > maxZ) maxZ = s->z[i];
    count++;
  }
  if (count == 0) return -1;
  const double cx = (minX + maxX) / 2, cy = (minY + maxY) / 2, cz = (minZ + maxZ) / 2;
  const double half = fmax(maxX - minX, fmax(maxY - minY, maxZ - minZ)) / 2 * 1.05;
  const int32_t root = bh_alloc(t);
  bh_node *r = &t->pool[root];
  r->cx = cx;
  r->cy = cy;
  r->cz = cz;
  r->half = half;
  r->depth = 0;
  for (int64_t i = 0; i < s->n; i++)
    if (bh_active(s, i)) bh_insert(t, root, i);
  return root;
}

/* directForce (octree.go:314-325): force on a from b */
static void bh_direct(const or_state *s, double G, double soft2, int64_t a, int64_t b, double *fx,
                      double *fy, double *fz) {
  const double dx = s->x[b] - s->x[a];
  const double dy = s->y[b] - s->y[a];
  const double dz = s->z[b] - s->z[a];
  const double distSq = dx * dx + dy * dy + dz * dz + soft2;
  const double dist = sqrt(distSq);
  const double force = G * s->m[a] * s->m[b] / distSq;
  *fx += force * dx / dist;
  *fy += force * dy / dist;
  *fz += force * dz / dist;
}

typedef struct bh_ctx {
  const bh_tree *t;
  double G, theta, soft2;
  int64_t interactions;
} bh_ctx;

/* accumForce (octree.go:276-309) */
static void bh_accum(bh_ctx *c, int32_t ni, int64_t b, double *fx, double *fy, double *fz) {
  const bh_tree *t = c->t;
  const or_state *s = t->s;
  const bh_node *n = &t->pool[ni];
  if (n->mass == 0) return;
  const double dx = n->cmx - s->x[b];
  const double dy = n->cmy - s->y[b];
  const double dz = n->cmz - s->z[b];
  const double distSq = dx * dx + dy * dy + dz * dz + c->soft2;
  const double dist = sqrt(distSq);
  const int leaf = n->child[0] < 0;
  if (leaf || (2 * n->half / dist) < c->theta) {
    if (leaf && n->bucket_tail >= 0) { /* a bucket: its bodies one by one */
      for (int64_t j = n->body;; j = t->next[j]) {
        if (j != b) {
          bh_direct(s, c->G, c->soft2, b, j, fx, fy, fz);
          c->interactions++;
        }
        if (j == n->bucket_tail) break;
      }
      return;
    }
    if (leaf && n->body == b) return;
    const double force = c->G * s->m[b] * n->mass / distSq;
    *fx += force * dx / dist;
    *fy += force * dy / dist;
    *fz += force * dz / dist;
    c->interactions++;
    return;
  }
  for (int k = 0; k < 8; k++) {
    const int32_t ch = n->child[k];
    if (ch >= 0 && t->pool[ch].mass > 0) bh_accum(c, ch, b, fx, fy, fz);
  }
}

int or_bh_force(const or_state *s, const or_bh_params *p, double *fx, double *fy, double *fz,
                int32_t *depth, or_bh_info *info) {
  bh_tree t;
  memset(&t, 0, sizeof t);
  t.s = s;
  t.max_depth = p->max_depth;
  t.next = (int64_t *)malloc((size_t)(s->n > 0 ? s->n : 1) * sizeof(int64_t));
  if (!t.next) return -1;
  for (int64_t i = 0; i < s->n; i++) fx[i] = fy[i] = fz[i] = 0.0;
  const int32_t root = bh_build(&t);
  bh_ctx c = {&t, p->G, p->theta, p->soft * p->soft, 0};
  if (root >= 0) {
    for (int64_t b = 0; b < s->n; b++) {
      if (!bh_active(s, b)) continue; /* Compute returns early (body.go:149-155) */
      double ax = 0, ay = 0, az = 0;
      if (s->flags[b] & OR_F_SUN) { /* CalcForce's sun branch (octree.go:252-267) */
        for (int64_t j = 0; j < s->n; j++) {
          if (j == b || !bh_active(s, j)) continue;
          bh_direct(s, p->G, c.soft2, b, j, &ax, &ay, &az);
          c.interactions++;
        }
      } else {
        bh_accum(&c, root, b, &ax, &ay, &az);
      }
      fx[b] = ax;
      fy[b] = ay;
      fz[b] = az;
    }
  }
  if (depth || info) {
    /* leaf depth of each body, and tree counts */
    int64_t internal = 0, buckets = 0;
    int32_t maxd = 0;
    if (depth)
      for (int64_t i = 0; i < s->n; i++) depth[i] = -1;
    for (int64_t k = 0; k < t.used; k++) {
      const bh_node *n = &t.pool[k];
      if (n->child[0] >= 0) {
        internal++;
        continue;
      }
      if (n->body < 0) continue;
      if (n->depth > maxd) maxd = n->depth;
      if (n->bucket_tail >= 0) buckets++;
      if (depth) {
        depth[n->body] = n->depth;
        if (n->bucket_tail >= 0)
          for (int64_t j = t.next[n->body];; j = t.next[j]) {
            depth[j] = n->depth;
            if (j == n->bucket_tail) break;
          }
      }
    }
    if (info) {
      memset(info, 0, sizeof *info);
      info->nodes = t.used;
      info->internal = internal;
      info->buckets = buckets;
      info->max_depth_seen = maxd;
      info->interactions = c.interactions;
      if (root >= 0) {
        info->root_mass = t.pool[root].mass;
        info->root_half = t.pool[root].half;
        info->root_cm[0] = t.pool[root].cmx;
        info->root_cm[1] = t.pool[root].cmy;
        info->root_cm[2] = t.pool[root].cmz;
      }
    }
  }
  free(t.pool);
  free(t.next);
  return 0;
}
