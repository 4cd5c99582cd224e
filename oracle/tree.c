/*
 * tree.c -- oracle (test infrastructure): Morton keys, the one-rank adaptive octree,
 * the box index, the U/V/W/X lists and the f64 FMM evaluation in the evaluator's
 * fixed order.
 *
 * Keys (octree/src/morton.rs:255-420): 16 levels of (x, y, z) bit triples, x highest,
 * shifted left by 15, the level in the low 15 bits.  Tree (octree/src/octree.rs:578-640
 * on one rank): the coarse tree is the root (implementation.rs:376-379), refined while a
 * box holds more than max_keys distinct level-16 keys (create_local_tree,
 * implementation.rs:682-716), then 2:1 balanced by the sibling/neighbour closure
 * (balance_closure, implementation.rs:1080-1120) and linearized (morton.rs:795-812).
 * Lists: the definitions of fmm-plan/src/interaction_manager.rs:10-28.  Order of calls
 * and rows: fmm-plan/src/operator.rs:34-48, 408-425.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "tree_internal.h"

#define DEEPEST 16
#define LEVEL_BITS 15
#define LEVEL_MASK 0x7FFFULL

static inline int nc(int p) { return (p + 1) * (p + 1); }

/* ---- keys ------------------------------------------------------------------------- */

uint64_t orc_key(int level, int64_t ix, int64_t iy, int64_t iz) {
  int shift = DEEPEST - level;
  uint64_t x = (uint64_t)ix << shift, y = (uint64_t)iy << shift, z = (uint64_t)iz << shift;
  uint64_t pos = 0;
  for (int b = 0; b < DEEPEST; ++b) {
    pos |= ((x >> b) & 1ULL) << (3 * b + 2);
    pos |= ((y >> b) & 1ULL) << (3 * b + 1);
    pos |= ((z >> b) & 1ULL) << (3 * b);
  }
  return (pos << LEVEL_BITS) | (uint64_t)level;
}

void orc_decode(uint64_t key, int64_t out[4]) {
  int level = (int)(key & LEVEL_MASK);
  uint64_t pos = key >> LEVEL_BITS, x = 0, y = 0, z = 0;
  for (int b = 0; b < DEEPEST; ++b) {
    x |= ((pos >> (3 * b + 2)) & 1ULL) << b;
    y |= ((pos >> (3 * b + 1)) & 1ULL) << b;
    z |= ((pos >> (3 * b)) & 1ULL) << b;
  }
  out[0] = level;
  out[1] = (int64_t)(x >> (DEEPEST - level));
  out[2] = (int64_t)(y >> (DEEPEST - level));
  out[3] = (int64_t)(z >> (DEEPEST - level));
}

static inline int key_level(uint64_t k) { return (int)(k & LEVEL_MASK); }

uint64_t orc_parent(uint64_t key) {
  int64_t d[4];
  orc_decode(key, d);
  return orc_key((int)d[0] - 1, d[1] >> 1, d[2] >> 1, d[3] >> 1);
}

static uint64_t child_key(uint64_t key, int o) {
  int64_t d[4];
  orc_decode(key, d);
  return orc_key((int)d[0] + 1, 2 * d[1] + ((o >> 2) & 1), 2 * d[2] + ((o >> 1) & 1),
                 2 * d[3] + (o & 1));
}

/* morton::is_ancestor (morton.rs:473-503): a key is its own ancestor. */
static int is_ancestor(uint64_t a, uint64_t b) {
  int la = key_level(a), lb = key_level(b);
  if (a == b) return 1;
  if (la > lb) return 0;
  int shift = LEVEL_BITS + 3 * (DEEPEST - la);
  return (a >> shift) == (b >> shift);
}

/* Closed cubes intersect and neither is an ancestor of the other
 * (fmm-plan/src/interaction_manager.rs:10-13). */
int orc_adjacent(uint64_t a, uint64_t b) {
  if (is_ancestor(a, b) || is_ancestor(b, a)) return 0;
  int64_t da[4], db[4];
  orc_decode(a, da);
  orc_decode(b, db);
  int L = (int)(da[0] > db[0] ? da[0] : db[0]);
  int64_t sa = 1LL << (L - da[0]), sb = 1LL << (L - db[0]);
  for (int k = 1; k <= 3; ++k) {
    int64_t alo = da[k] * sa, ahi = alo + sa, blo = db[k] * sb, bhi = blo + sb;
    if (alo > bhi || blo > ahi) return 0;
  }
  return 1;
}

/* compute_global_bounding_box (octree/src/octree.rs:1278-1369) on one rank. */
void orc_bounding_box(const double *pts, int64_t n, double out[6]) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], pts[3 * i + k]);
      hi[k] = fmax(hi[k], pts[3 * i + k]);
    }
  double diam[3], mean[3];
  for (int k = 0; k < 3; ++k) {
    diam[k] = hi[k] - lo[k];
    mean[k] = lo[k] + 0.5 * diam[k];
  }
  double maxd = fmax(fmax(diam[0], diam[1]), diam[2]);
  maxd = maxd * (1.0 + 1.0 / 65536.0);
  for (int k = 0; k < 3; ++k) {
    out[k] = mean[k] - 0.5 * maxd;
    out[k + 3] = mean[k] + 0.5 * maxd;
  }
}

/* morton::from_physical_point (morton.rs:255-273) with the box's per-axis sides. */
uint64_t orc_point_key(const double x[3], const double box[6], int level) {
  int64_t size = 1LL << level, idx[3];
  for (int k = 0; k < 3; ++k) {
    double ref = (x[k] - box[k]) / (box[k + 3] - box[k]);
    double v = ref * (double)size;
    int64_t i = v <= 0.0 ? 0 : (int64_t)v; /* Rust `as usize` saturates at 0 */
    idx[k] = i < size - 1 ? i : size - 1;
  }
  return orc_key(level, idx[0], idx[1], idx[2]);
}

/* ---- growable key vectors and sorting ---------------------------------------------- */

typedef struct {
  uint64_t *v;
  int64_t n, cap;
} kvec;

static void kpush(kvec *a, uint64_t k) {
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 64;
    a->v = realloc(a->v, sizeof(uint64_t) * a->cap);
  }
  a->v[a->n++] = k;
}

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* morton::linearize (morton.rs:795-812): sort, drop duplicates and ancestors. */
static void linearize(kvec *a) {
  if (a->n == 0) return;
  qsort(a->v, a->n, sizeof(uint64_t), cmp_u64);
  int64_t out = 0;
  for (int64_t i = 0; i + 1 < a->n; ++i) {
    if (a->v[i] == a->v[i + 1] || is_ancestor(a->v[i], a->v[i + 1])) continue;
    a->v[out++] = a->v[i];
  }
  a->v[out++] = a->v[a->n - 1];
  a->n = out;
}

int64_t orc_find_key(const uint64_t *v, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (v[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < n && v[lo] == key) ? lo : -1;
}

/* create_local_tree (implementation.rs:682-716) from one coarse key. */
static void refine(const uint64_t *keys, int64_t n, uint64_t box, int max_level,
                   int64_t max_keys, kvec *out) {
  int level = key_level(box);
  if (level < max_level && n > max_keys) {
    int shift = LEVEL_BITS + 3 * (DEEPEST - (level + 1));
    int64_t at = 0;
    for (int o = 0; o < 8; ++o) {
      uint64_t c = child_key(box, o);
      int64_t start = at;
      while (at < n && (keys[at] >> shift) == (c >> shift)) ++at;
      refine(keys + start, at - start, c, max_level, max_keys, out);
    }
  } else {
    kpush(out, box);
  }
}

/* balance_closure (implementation.rs:1080-1120) then linearize. */
static void balance(kvec *leaves) {
  int deepest = 0;
  for (int64_t i = 0; i < leaves->n; ++i)
    if (key_level(leaves->v[i]) > deepest) deepest = key_level(leaves->v[i]);
  if (deepest == 0) return;
  kvec work = {0}, result = {0};
  for (int64_t i = 0; i < leaves->n; ++i)
    if (key_level(leaves->v[i]) == deepest) kpush(&work, leaves->v[i]);
  for (int level = deepest; level >= 1; --level) {
    /* The distinct parents of the work list; each adds its eight children to the
     * result and its neighbours to the next work list (sets: order is immaterial, the
     * result is linearized). */
    kvec parents = {0}, next = {0};
    for (int64_t i = 0; i < work.n; ++i) kpush(&parents, orc_parent(work.v[i]));
    qsort(parents.v, parents.n, sizeof(uint64_t), cmp_u64);
    for (int64_t i = 0; i < parents.n; ++i) {
      uint64_t par = parents.v[i];
      if (i > 0 && par == parents.v[i - 1]) continue;
      for (int o = 0; o < 8; ++o) kpush(&result, child_key(par, o));
      int64_t d[4];
      orc_decode(par, d);
      int64_t size = 1LL << d[0];
      if (d[0] > 0)
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
              if (!dx && !dy && !dz) continue;
              int64_t x = d[1] + dx, y = d[2] + dy, z = d[3] + dz;
              if (x < 0 || y < 0 || z < 0 || x >= size || y >= size || z >= size) continue;
              kpush(&next, orc_key((int)d[0], x, y, z));
            }
    }
    for (int64_t i = 0; i < leaves->n; ++i)
      if (key_level(leaves->v[i]) == level - 1) kpush(&next, leaves->v[i]);
    free(parents.v);
    free(work.v);
    work = next;
  }
  free(work.v);
  linearize(&result);
  free(leaves->v);
  *leaves = result;
}

/* ---- the FMM ------------------------------------------------------------------------ */

static int64_t box_index(const orc_fmm *f, uint64_t key) {
  int l = key_level(key);
  if (l >= f->nlevels) return -1;
  return orc_find_key(f->keys[l], f->nbox[l], key);
}

static int is_leaf_key(const orc_fmm *f, uint64_t key) {
  int64_t b = box_index(f, key);
  return b >= 0 && f->box_leaf[key_level(key)][b] >= 0;
}

/* 26 same-level neighbour cells inside the domain. */
static int neighbours(uint64_t key, uint64_t out[26]) {
  int64_t d[4];
  orc_decode(key, d);
  int cnt = 0;
  if (d[0] == 0) return 0;
  int64_t size = 1LL << d[0];
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        if (!dx && !dy && !dz) continue;
        int64_t x = d[1] + dx, y = d[2] + dy, z = d[3] + dz;
        if (x < 0 || y < 0 || z < 0 || x >= size || y >= size || z >= size) continue;
        out[cnt++] = orc_key((int)d[0], x, y, z);
      }
  return cnt;
}

/* Leaves adjacent to b under cell c (c exists in the tree). */
static void collect_adjacent_leaves(const orc_fmm *f, uint64_t c, uint64_t b, kvec *out) {
  if (!orc_adjacent(c, b)) return;
  if (is_leaf_key(f, c)) {
    kpush(out, c);
    return;
  }
  for (int o = 0; o < 8; ++o) collect_adjacent_leaves(f, child_key(c, o), b, out);
}

/* U-list of leaf b: every leaf adjacent to b. */
static void u_list(const orc_fmm *f, uint64_t b, kvec *out) {
  uint64_t nb[26];
  int cnt = neighbours(b, nb);
  for (int i = 0; i < cnt; ++i) {
    uint64_t c = nb[i];
    if (box_index(f, c) >= 0) {
      collect_adjacent_leaves(f, c, b, out);
    } else {
      /* the cell is covered by a coarser leaf */
      uint64_t a = c;
      while (key_level(a) > 0 && box_index(f, a) < 0) a = orc_parent(a);
      if (is_leaf_key(f, a) && orc_adjacent(a, b)) kpush(out, a);
    }
  }
}

/* W-list of leaf b: finer boxes whose parent is adjacent to b, not adjacent to b. */
static void w_descend(const orc_fmm *f, uint64_t parent, uint64_t b, kvec *out) {
  for (int o = 0; o < 8; ++o) {
    uint64_t c = child_key(parent, o);
    if (orc_adjacent(c, b)) {
      if (!is_leaf_key(f, c)) w_descend(f, c, b, out);
    } else {
      kpush(out, c);
    }
  }
}

static void w_list(const orc_fmm *f, uint64_t b, kvec *out) {
  uint64_t nb[26];
  int cnt = neighbours(b, nb);
  for (int i = 0; i < cnt; ++i)
    if (box_index(f, nb[i]) >= 0 && !is_leaf_key(f, nb[i])) w_descend(f, nb[i], b, out);
}

/* V-list of b: boxes on b's level whose parent is adjacent to b's parent, not adjacent
 * to b.  X-list of b: leaves on the parent's level adjacent to the parent, not adjacent
 * to b. */
static void v_x_lists(const orc_fmm *f, uint64_t b, kvec *v, kvec *x) {
  if (key_level(b) == 0) return;
  uint64_t par = orc_parent(b), nb[26];
  int cnt = neighbours(par, nb);
  for (int i = 0; i < cnt; ++i) {
    if (box_index(f, nb[i]) < 0) continue;
    if (is_leaf_key(f, nb[i])) {
      if (!orc_adjacent(nb[i], b)) kpush(x, nb[i]);
    } else {
      for (int o = 0; o < 8; ++o) {
        uint64_t c = child_key(nb[i], o);
        if (!orc_adjacent(c, b)) kpush(v, c);
      }
    }
  }
}

static int cmp_level_key(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  int lx = key_level(x), ly = key_level(y);
  if (lx != ly) return lx < ly ? -1 : 1;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static int64_t leaf_index(const orc_fmm *f, uint64_t key) {
  int64_t b = box_index(f, key);
  return b < 0 ? -1 : f->box_leaf[key_level(key)][b];
}

/* leaf-scaled coordinates, fmm-exec/src/geometry.rs:291-298 */
static void leaf_coords(const orc_fmm *f, const double *x, uint64_t leaf, double *u) {
  int64_t d[4];
  orc_decode(leaf, d);
  double scale = ldexp(1.0, (int)d[0] + 1);
  for (int k = 0; k < 3; ++k) {
    double dd = x[k] - f->lower[k];
    u[k] = dd * scale / f->side - (double)(2 * d[k + 1] + 1);
  }
}

static int64_t leaf_of_point(const orc_fmm *f, uint64_t fine) {
  /* the leaf that is an ancestor of the level-16 key */
  for (int l = f->nlevels - 1; l >= 0; --l) {
    uint64_t a = fine;
    int64_t d[4];
    orc_decode(fine, d);
    a = orc_key(l, d[1] >> (DEEPEST - l), d[2] >> (DEEPEST - l), d[3] >> (DEEPEST - l));
    int64_t li = leaf_index(f, a);
    if (li >= 0) return li;
  }
  return -1;
}

static void csr_free(csr *c) {
  free(c->off);
  free(c->ent);
  free(c->aux);
}

static void csr_from_lists(csr *c, kvec *rows, int64_t nrows) {
  c->off = malloc(sizeof(int64_t) * (nrows + 1));
  c->off[0] = 0;
  for (int64_t r = 0; r < nrows; ++r) c->off[r + 1] = c->off[r] + rows[r].n;
  c->ent = malloc(sizeof(int64_t) * (c->off[nrows] + 1));
  c->aux = malloc(sizeof(int64_t) * (c->off[nrows] + 1));
}

static int64_t *points_to_leaves(const orc_fmm *f, const double *pts, int64_t n, int64_t **start,
                                 int64_t **order) {
  int64_t *leaf = malloc(sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; ++i) leaf[i] = leaf_of_point(f, orc_point_key(pts + 3 * i, f->box, DEEPEST));
  int64_t *cnt = calloc(f->nleaves + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) cnt[leaf[i] + 1]++;
  for (int64_t l = 0; l < f->nleaves; ++l) cnt[l + 1] += cnt[l];
  *start = cnt;
  int64_t *pos = malloc(sizeof(int64_t) * (f->nleaves + 1));
  memcpy(pos, cnt, sizeof(int64_t) * (f->nleaves + 1));
  *order = malloc(sizeof(int64_t) * (n + 1));
  for (int64_t i = 0; i < n; ++i) (*order)[pos[leaf[i]]++] = i; /* stable: caller order */
  free(pos);
  return leaf;
}

/* The near rule (fmm-plan/src/near.rs:15-40) with named crossovers: a W entry of a
 * leaf that is itself a leaf moves to the near row when it holds fewer than near_w
 * sources; the X entries of a leaf move to its near row when it holds fewer than
 * near_x targets.  Near rows stay in leaf (level, key) order.  Negative: off. */
static void apply_near_rule(orc_fmm *f, double near_w, double near_x) {
  for (int l = 0; l < f->nlevels; ++l) {
    int64_t l0 = f->leaf_off[l], nl = f->leaf_off[l + 1] - l0;
    kvec *near = calloc(nl + 1, sizeof(kvec)), *wk = calloc(nl + 1, sizeof(kvec));
    for (int64_t r = 0; r < nl; ++r) {
      int64_t lf = l0 + r;
      for (int64_t e = f->u[l].off[r]; e < f->u[l].off[r + 1]; ++e) kpush(&near[r], (uint64_t)f->u[l].ent[e]);
      for (int64_t e = f->w[l].off[r]; e < f->w[l].off[r + 1]; ++e) {
        int64_t sb = f->w[l].ent[e];
        int sl = (int)f->w[l].aux[e];
        int64_t leaf = f->box_leaf[sl][sb];
        if (near_w >= 0 && leaf >= 0 && (double)(f->src_start[leaf + 1] - f->src_start[leaf]) < near_w)
          kpush(&near[r], (uint64_t)leaf);
        else
          kpush(&wk[r], (uint64_t)e);
      }
      /* X row of this leaf's box */
      int64_t b = f->leaf_box[lf];
      if (near_x >= 0 && (double)(f->tgt_start[lf + 1] - f->tgt_start[lf]) < near_x) {
        for (int64_t e = f->x[l].off[b]; e < f->x[l].off[b + 1]; ++e) kpush(&near[r], (uint64_t)f->x[l].ent[e]);
      }
      qsort(near[r].v, near[r].n, sizeof(uint64_t), cmp_u64);
    }
    /* rebuild near and W rows */
    csr nu, nw;
    csr_from_lists(&nu, near, nl);
    csr_from_lists(&nw, wk, nl);
    for (int64_t r = 0; r < nl; ++r) {
      for (int64_t e = 0; e < near[r].n; ++e) nu.ent[nu.off[r] + e] = (int64_t)near[r].v[e];
      for (int64_t e = 0; e < wk[r].n; ++e) {
        nw.ent[nw.off[r] + e] = f->w[l].ent[wk[r].v[e]];
        nw.aux[nw.off[r] + e] = f->w[l].aux[wk[r].v[e]];
      }
      free(near[r].v);
      free(wk[r].v);
    }
    free(near);
    free(wk);
    csr_free(&f->u[l]);
    csr_free(&f->w[l]);
    f->u[l] = nu;
    f->w[l] = nw;
    /* X rows of converted leaves become empty */
    int64_t nb = f->nbox[l];
    kvec *xr = calloc(nb + 1, sizeof(kvec));
    for (int64_t b = 0; b < nb; ++b) {
      int64_t lf = f->box_leaf[l][b];
      int drop = near_x >= 0 && lf >= 0 && (double)(f->tgt_start[lf + 1] - f->tgt_start[lf]) < near_x;
      if (!drop)
        for (int64_t e = f->x[l].off[b]; e < f->x[l].off[b + 1]; ++e) kpush(&xr[b], (uint64_t)f->x[l].ent[e]);
    }
    csr nx;
    csr_from_lists(&nx, xr, nb);
    for (int64_t b = 0; b < nb; ++b) {
      for (int64_t e = 0; e < xr[b].n; ++e) nx.ent[nx.off[b] + e] = (int64_t)xr[b].v[e];
      free(xr[b].v);
    }
    free(xr);
    csr_free(&f->x[l]);
    f->x[l] = nx;
  }
}

orc_fmm *orc_fmm_new(const double *src, int64_t ns, const double *tgt, int64_t nt, int p,
                     int64_t leaf_points, int max_level, int gradients, double near_w, double near_x) {
  int64_t n = ns + nt;
  if (n == 0) return NULL;
  double *all = malloc(sizeof(double) * 3 * n);
  memcpy(all, src, sizeof(double) * 3 * ns);
  memcpy(all + 3 * ns, tgt, sizeof(double) * 3 * nt);
  orc_fmm *f = calloc(1, sizeof(orc_fmm));
  f->p = p;
  f->gradients = gradients;
  f->ns = ns;
  f->nt = nt;
  orc_bounding_box(all, n, f->box);
  double sides[3];
  for (int k = 0; k < 3; ++k) {
    f->lower[k] = f->box[k];
    sides[k] = f->box[k + 3] - f->box[k];
  }
  if (!(sides[0] > 0 || sides[1] > 0 || sides[2] > 0)) {
    free(all);
    free(f);
    return NULL;
  }
  f->side = fmax(fmax(sides[0], sides[1]), sides[2]);
  /* distinct level-16 keys */
  kvec fine = {0};
  for (int64_t i = 0; i < n; ++i) kpush(&fine, orc_point_key(all + 3 * i, f->box, DEEPEST));
  qsort(fine.v, fine.n, sizeof(uint64_t), cmp_u64);
  int64_t distinct = 0;
  for (int64_t i = 0; i < fine.n; ++i)
    if (i == 0 || fine.v[i] != fine.v[i - 1]) fine.v[distinct++] = fine.v[i];
  fine.n = distinct;
  kvec leaves = {0};
  if (max_level > DEEPEST) max_level = DEEPEST;
  refine(fine.v, fine.n, 0ULL /* root */, max_level, leaf_points, &leaves);
  free(fine.v);
  balance(&leaves);
  /* all keys = leaves and their ancestors */
  kvec all_keys = {0};
  int deepest = 0;
  for (int64_t i = 0; i < leaves.n; ++i) {
    uint64_t k = leaves.v[i];
    if (key_level(k) > deepest) deepest = key_level(k);
    kpush(&all_keys, k);
    while (key_level(k) > 0) {
      k = orc_parent(k);
      kpush(&all_keys, k);
    }
  }
  qsort(all_keys.v, all_keys.n, sizeof(uint64_t), cmp_u64);
  f->nlevels = deepest + 1;
  f->nbox = calloc(f->nlevels, sizeof(int64_t));
  f->keys = calloc(f->nlevels, sizeof(uint64_t *));
  f->box_leaf = calloc(f->nlevels, sizeof(int64_t *));
  f->first_child = calloc(f->nlevels, sizeof(int64_t *));
  for (int l = 0; l < f->nlevels; ++l) f->keys[l] = malloc(sizeof(uint64_t) * (all_keys.n + 1));
  for (int64_t i = 0; i < all_keys.n; ++i) {
    if (i > 0 && all_keys.v[i] == all_keys.v[i - 1]) continue;
    int l = key_level(all_keys.v[i]);
    f->keys[l][f->nbox[l]++] = all_keys.v[i];
  }
  free(all_keys.v);
  /* leaves by (level, key) */
  qsort(leaves.v, leaves.n, sizeof(uint64_t), cmp_level_key);
  f->nleaves = leaves.n;
  f->leaf_key = leaves.v;
  f->leaf_box = malloc(sizeof(int64_t) * f->nleaves);
  f->leaf_off = calloc(f->nlevels + 1, sizeof(int64_t));
  for (int l = 0; l < f->nlevels; ++l) {
    f->box_leaf[l] = malloc(sizeof(int64_t) * (f->nbox[l] + 1));
    f->first_child[l] = malloc(sizeof(int64_t) * (f->nbox[l] + 1));
    for (int64_t b = 0; b < f->nbox[l]; ++b) f->box_leaf[l][b] = -1;
  }
  for (int64_t i = 0; i < f->nleaves; ++i) {
    int l = key_level(f->leaf_key[i]);
    int64_t b = orc_find_key(f->keys[l], f->nbox[l], f->leaf_key[i]);
    f->leaf_box[i] = b;
    f->box_leaf[l][b] = i;
    f->leaf_off[l + 1]++;
  }
  for (int l = 0; l < f->nlevels; ++l) f->leaf_off[l + 1] += f->leaf_off[l];
  for (int l = 0; l < f->nlevels; ++l)
    for (int64_t b = 0; b < f->nbox[l]; ++b)
      f->first_child[l][b] =
          (f->box_leaf[l][b] < 0 && l + 1 < f->nlevels)
              ? orc_find_key(f->keys[l + 1], f->nbox[l + 1], child_key(f->keys[l][b], 0))
              : -1;
  /* lists */
  f->v = calloc(f->nlevels, sizeof(csr));
  f->x = calloc(f->nlevels, sizeof(csr));
  f->u = calloc(f->nlevels, sizeof(csr));
  f->w = calloc(f->nlevels, sizeof(csr));
  for (int l = 0; l < f->nlevels; ++l) {
    int64_t nb = f->nbox[l];
    kvec *vr = calloc(nb + 1, sizeof(kvec)), *xr = calloc(nb + 1, sizeof(kvec));
    for (int64_t b = 0; b < nb; ++b) v_x_lists(f, f->keys[l][b], &vr[b], &xr[b]);
    csr_from_lists(&f->v[l], vr, nb);
    csr_from_lists(&f->x[l], xr, nb);
    for (int64_t b = 0; b < nb; ++b) {
      /* V by offset index, X by key */
      int64_t dt[4], ds[4], base = f->v[l].off[b];
      orc_decode(f->keys[l][b], dt);
      int64_t cntv = vr[b].n;
      int64_t *tmp = malloc(sizeof(int64_t) * 2 * (cntv + 1));
      for (int64_t e = 0; e < cntv; ++e) {
        orc_decode(vr[b].v[e], ds);
        tmp[2 * e] = orc_offset_index((int)(dt[1] - ds[1]), (int)(dt[2] - ds[2]), (int)(dt[3] - ds[3]));
        tmp[2 * e + 1] = orc_find_key(f->keys[l], nb, vr[b].v[e]);
      }
      /* insertion sort by offset index */
      for (int64_t i = 1; i < cntv; ++i)
        for (int64_t j = i; j > 0 && tmp[2 * j - 2] > tmp[2 * j]; --j) {
          int64_t a0 = tmp[2 * j], a1 = tmp[2 * j + 1];
          tmp[2 * j] = tmp[2 * j - 2];
          tmp[2 * j + 1] = tmp[2 * j - 1];
          tmp[2 * j - 2] = a0;
          tmp[2 * j - 1] = a1;
        }
      for (int64_t e = 0; e < cntv; ++e) {
        f->v[l].aux[base + e] = tmp[2 * e];
        f->v[l].ent[base + e] = tmp[2 * e + 1];
      }
      free(tmp);
      qsort(xr[b].v, xr[b].n, sizeof(uint64_t), cmp_u64);
      for (int64_t e = 0; e < xr[b].n; ++e) f->x[l].ent[f->x[l].off[b] + e] = leaf_index(f, xr[b].v[e]);
      free(vr[b].v);
      free(xr[b].v);
    }
    free(vr);
    free(xr);
    /* leaves of the level: near (U + self) and W */
    int64_t l0 = f->leaf_off[l], l1 = f->leaf_off[l + 1], nl = l1 - l0;
    kvec *ur = calloc(nl + 1, sizeof(kvec)), *wr = calloc(nl + 1, sizeof(kvec));
    for (int64_t r = 0; r < nl; ++r) {
      uint64_t k = f->leaf_key[l0 + r];
      u_list(f, k, &ur[r]);
      kpush(&ur[r], k);
      qsort(ur[r].v, ur[r].n, sizeof(uint64_t), cmp_level_key);
      int64_t out = 0;
      for (int64_t e = 0; e < ur[r].n; ++e)
        if (e == 0 || ur[r].v[e] != ur[r].v[e - 1]) ur[r].v[out++] = ur[r].v[e];
      ur[r].n = out;
      w_list(f, k, &wr[r]);
      qsort(wr[r].v, wr[r].n, sizeof(uint64_t), cmp_u64);
    }
    csr_from_lists(&f->u[l], ur, nl);
    csr_from_lists(&f->w[l], wr, nl);
    for (int64_t r = 0; r < nl; ++r) {
      for (int64_t e = 0; e < ur[r].n; ++e) f->u[l].ent[f->u[l].off[r] + e] = leaf_index(f, ur[r].v[e]);
      for (int64_t e = 0; e < wr[r].n; ++e) {
        /* W entries lie on level l + 1 under 2:1 balance */
        f->w[l].ent[f->w[l].off[r] + e] = box_index(f, wr[r].v[e]);
        f->w[l].aux[f->w[l].off[r] + e] = key_level(wr[r].v[e]);
      }
      free(ur[r].v);
      free(wr[r].v);
    }
    free(ur);
    free(wr);
  }
  /* points into leaves, leaf-scaled coordinates */
  f->src_leaf = points_to_leaves(f, src, ns, &f->src_start, &f->src_order);
  f->tgt_leaf = points_to_leaves(f, tgt, nt, &f->tgt_start, &f->tgt_order);
  f->us = malloc(sizeof(double) * 3 * (ns + 1));
  f->ut = malloc(sizeof(double) * 3 * (nt + 1));
  for (int64_t lf = 0; lf < f->nleaves; ++lf) {
    for (int64_t i = f->src_start[lf]; i < f->src_start[lf + 1]; ++i)
      leaf_coords(f, src + 3 * f->src_order[i], f->leaf_key[lf], f->us + 3 * i);
    for (int64_t i = f->tgt_start[lf]; i < f->tgt_start[lf + 1]; ++i)
      leaf_coords(f, tgt + 3 * f->tgt_order[i], f->leaf_key[lf], f->ut + 3 * i);
  }
  if (near_w >= 0 || near_x >= 0) apply_near_rule(f, near_w, near_x);
  /* tables */
  size_t n2 = (size_t)nc(p) * nc(p);
  f->m2m = malloc(sizeof(double) * 8 * n2);
  f->l2l = malloc(sizeof(double) * 8 * n2);
  f->m2l = malloc(sizeof(double) * 316 * n2);
  orc_table_m2m(p, f->m2m);
  orc_table_l2l(p, f->l2l);
  if (f->nlevels > 2) orc_table_m2l(p, f->m2l);
  free(all);
  return f;
}

void orc_fmm_free(orc_fmm *f) {
  if (!f) return;
  if (f->mult) {
    for (int l = 0; l < f->nlevels; ++l) {
      free(f->mult[l]);
      free(f->loc[l]);
    }
    free(f->mult);
    free(f->loc);
  }
  for (int l = 0; l < f->nlevels; ++l) {
    free(f->keys[l]);
    free(f->box_leaf[l]);
    free(f->first_child[l]);
    csr_free(&f->v[l]);
    csr_free(&f->x[l]);
    csr_free(&f->u[l]);
    csr_free(&f->w[l]);
  }
  free(f->keys);
  free(f->box_leaf);
  free(f->first_child);
  free(f->nbox);
  free(f->v);
  free(f->x);
  free(f->u);
  free(f->w);
  free(f->leaf_key);
  free(f->leaf_box);
  free(f->leaf_off);
  free(f->src_start);
  free(f->tgt_start);
  free(f->src_order);
  free(f->tgt_order);
  free(f->src_leaf);
  free(f->tgt_leaf);
  free(f->us);
  free(f->ut);
  free(f->m2m);
  free(f->l2l);
  free(f->m2l);
  free(f);
}

/* ĉ(s|t), r̂(s|t): CONVENTIONS §3.13 "Relative frames", exact from integer centres. */
void orc_relative_frame(uint64_t s, uint64_t t, double c[3], double *r) {
  int64_t ds[4], dt[4];
  orc_decode(s, ds);
  orc_decode(t, dt);
  int L = (int)(ds[0] > dt[0] ? ds[0] : dt[0]);
  for (int k = 1; k <= 3; ++k) {
    int64_t cs = (2 * ds[k] + 1) << (L - ds[0]), ct = (2 * dt[k] + 1) << (L - dt[0]);
    c[k - 1] = (double)(cs - ct) * ldexp(1.0, (int)dt[0] - L);
  }
  *r = ldexp(1.0, (int)(dt[0] - ds[0]));
}

/* output += A x, A column-major n x n */
static void matvec_add(int n, const double *a, const double *x, double *y) {
  for (int k = 0; k < n; ++k) {
    double xk = x[k];
    if (xk == 0.0) continue;
    const double *col = a + (size_t)k * n;
    for (int i = 0; i < n; ++i) y[i] += col[i] * xk;
  }
}

void orc_fmm_evaluate(orc_fmm *f, const double *q_caller, double *phi, double *grad) {
  int p = f->p, n = nc(p), L = f->nlevels - 1;
  size_t n2 = (size_t)n * n;
  if (!f->mult) {
    f->mult = calloc(f->nlevels, sizeof(double *));
    f->loc = calloc(f->nlevels, sizeof(double *));
    for (int l = 0; l < f->nlevels; ++l) {
      f->mult[l] = malloc(((size_t)f->nbox[l] * n + 1) * sizeof(double));
      f->loc[l] = malloc(((size_t)f->nbox[l] * n + 1) * sizeof(double));
    }
  }
  double **mult = f->mult, **loc = f->loc;
  for (int l = 0; l < f->nlevels; ++l) {
    memset(mult[l], 0, ((size_t)f->nbox[l] * n + 1) * sizeof(double));
    memset(loc[l], 0, ((size_t)f->nbox[l] * n + 1) * sizeof(double));
  }
  double *q = malloc(sizeof(double) * (f->ns + 1));
  for (int64_t i = 0; i < f->ns; ++i) q[i] = q_caller[f->src_order[i]];
  double *ph = calloc(f->nt + 1, sizeof(double)), *gr = calloc(3 * f->nt + 3, sizeof(double));
  double *gp = f->gradients ? gr : NULL;
  const double zero3[3] = {0, 0, 0};
  /* upward */
  for (int l = L; l >= 0; --l) {
    for (int64_t lf = f->leaf_off[l]; lf < f->leaf_off[l + 1]; ++lf) {
      int64_t b = f->leaf_box[lf], s0 = f->src_start[lf], s1 = f->src_start[lf + 1];
      orc_p2m(p, zero3, 1.0, f->us + 3 * s0, q + s0, s1 - s0, mult[l] + (size_t)b * n);
    }
    if (l >= 1)
      for (int64_t b = 0; b < f->nbox[l - 1]; ++b) {
        int64_t c0 = f->first_child[l - 1][b];
        if (c0 < 0) continue;
        for (int o = 0; o < 8; ++o)
          matvec_add(n, f->m2m + o * n2, mult[l] + (size_t)(c0 + o) * n, mult[l - 1] + (size_t)b * n);
      }
  }
  /* downward */
  for (int l = 1; l <= L; ++l) {
    for (int64_t b = 0; b < f->nbox[l]; ++b) {
      uint64_t k = f->keys[l][b];
      int64_t pb = orc_find_key(f->keys[l - 1], f->nbox[l - 1], orc_parent(k));
      int64_t d[4];
      orc_decode(k, d);
      int o = (int)(((d[1] & 1) << 2) | ((d[2] & 1) << 1) | (d[3] & 1));
      matvec_add(n, f->l2l + o * n2, loc[l - 1] + (size_t)pb * n, loc[l] + (size_t)b * n);
    }
    for (int64_t b = 0; b < f->nbox[l]; ++b)
      for (int64_t e = f->v[l].off[b]; e < f->v[l].off[b + 1]; ++e)
        matvec_add(n, f->m2l + f->v[l].aux[e] * n2, mult[l] + (size_t)f->v[l].ent[e] * n,
                   loc[l] + (size_t)b * n);
    for (int64_t b = 0; b < f->nbox[l]; ++b)
      for (int64_t e = f->x[l].off[b]; e < f->x[l].off[b + 1]; ++e) {
        int64_t sl = f->x[l].ent[e], s0 = f->src_start[sl], s1 = f->src_start[sl + 1];
        double c[3], r;
        orc_relative_frame(f->keys[l][b], f->leaf_key[sl], c, &r); /* ĉ(t|s), r̂(t|s) */
        orc_p2l(p, c, r, f->us + 3 * s0, q + s0, s1 - s0, loc[l] + (size_t)b * n);
      }
  }
  /* leaves */
  for (int l = 0; l <= L; ++l) {
    for (int64_t lf = f->leaf_off[l]; lf < f->leaf_off[l + 1]; ++lf) {
      int64_t t0 = f->tgt_start[lf], t1 = f->tgt_start[lf + 1], b = f->leaf_box[lf];
      orc_l2p(p, zero3, 1.0, loc[l] + (size_t)b * n, f->ut + 3 * t0, t1 - t0, ph + t0,
              gp ? gp + 3 * t0 : NULL);
    }
    for (int64_t lf = f->leaf_off[l]; lf < f->leaf_off[l + 1]; ++lf) {
      int64_t r = lf - f->leaf_off[l], t0 = f->tgt_start[lf], t1 = f->tgt_start[lf + 1];
      for (int64_t e = f->w[l].off[r]; e < f->w[l].off[r + 1]; ++e) {
        int64_t sb = f->w[l].ent[e];
        int sl = (int)f->w[l].aux[e];
        double c[3], rr;
        orc_relative_frame(f->keys[sl][sb], f->leaf_key[lf], c, &rr); /* ĉ(s|t), r̂(s|t) */
        orc_m2p(p, c, rr, mult[sl] + (size_t)sb * n, f->ut + 3 * t0, t1 - t0, ph + t0,
                gp ? gp + 3 * t0 : NULL);
      }
    }
    for (int64_t lf = f->leaf_off[l]; lf < f->leaf_off[l + 1]; ++lf) {
      int64_t r = lf - f->leaf_off[l], t0 = f->tgt_start[lf], t1 = f->tgt_start[lf + 1];
      for (int64_t e = f->u[l].off[r]; e < f->u[l].off[r + 1]; ++e) {
        int64_t sl = f->u[l].ent[e], s0 = f->src_start[sl], s1 = f->src_start[sl + 1];
        double *y = malloc(sizeof(double) * 3 * (s1 - s0 + 1));
        if (sl == lf) {
          memcpy(y, f->us + 3 * s0, sizeof(double) * 3 * (s1 - s0));
        } else {
          double c[3], rr;
          orc_relative_frame(f->leaf_key[sl], f->leaf_key[lf], c, &rr); /* ĉ(s|t), r̂(s|t) */
          for (int64_t j = 0; j < s1 - s0; ++j)
            for (int k = 0; k < 3; ++k) y[3 * j + k] = c[k] + rr * f->us[3 * (s0 + j) + k];
        }
        orc_p2p(y, q + s0, s1 - s0, f->ut + 3 * t0, t1 - t0, ph + t0, gp ? gp + 3 * t0 : NULL);
        free(y);
      }
    }
  }
  /* output pass: phi = phi^ / (4 pi r_t), grad = g^ / (4 pi r_t^2) */
  for (int64_t lf = 0; lf < f->nleaves; ++lf) {
    int lvl = key_level(f->leaf_key[lf]);
    double rt = f->side / ldexp(1.0, lvl + 1);
    double s1 = 4.0 * 3.141592653589793 * rt, s2 = 4.0 * 3.141592653589793 * rt * rt;
    for (int64_t i = f->tgt_start[lf]; i < f->tgt_start[lf + 1]; ++i) {
      int64_t c = f->tgt_order[i];
      phi[c] = ph[i] / s1;
      if (grad && f->gradients)
        for (int k = 0; k < 3; ++k) grad[3 * c + k] = gr[3 * i + k] / s2;
    }
  }
  free(q);
  free(ph);
  free(gr);
}

/* Helmholtz FMM over the same tree and lists (CONVENTIONS §3.15; nd_fmm_exec::helmholtz):
 * complex charges q[2 ns] (re, im interleaved, caller order), outputs phi[2 nt] and
 * grad[6 nt].  Level l works in its own units, wavenumber kh_l = k r_l; stored expansions
 * as hexpansion.c; P2P sums e^{i kh r}/r in the target leaf's units; the output pass divides
 * by 4 pi r_t (potential) and 4 pi r_t^2 (gradient) as for Laplace. */
void orc_hfmm_evaluate(orc_fmm *f, double kre, double kim, const double *q_caller, double *phi, double *grad) {
  const int p = f->p, n = nc(p), L = f->nlevels - 1;
  const size_t n2 = (size_t)n * n, w = 2 * (size_t)n;
  double **mult = calloc(f->nlevels, sizeof(double *)), **loc = calloc(f->nlevels, sizeof(double *));
  double *kh = malloc(sizeof(double) * 2 * f->nlevels);
  for (int l = 0; l <= L; ++l) {
    mult[l] = calloc((size_t)f->nbox[l] * w + 1, sizeof(double));
    loc[l] = calloc((size_t)f->nbox[l] * w + 1, sizeof(double));
    const double r = f->side / ldexp(1.0, l + 1);
    kh[2 * l] = kre * r;
    kh[2 * l + 1] = kim * r;
  }
  double *q = malloc(sizeof(double) * 2 * (f->ns + 1));
  for (int64_t i = 0; i < f->ns; ++i) {
    q[2 * i] = q_caller[2 * f->src_order[i]];
    q[2 * i + 1] = q_caller[2 * f->src_order[i] + 1];
  }
  double *ph = calloc(2 * f->nt + 2, sizeof(double)), *gr = calloc(6 * f->nt + 6, sizeof(double));
  double *gp = f->gradients ? gr : NULL;
  double *T = malloc(sizeof(double) * 2 * n2);
  const double zero3[3] = {0, 0, 0};
  /* upward */
  for (int l = L; l >= 0; --l) {
    for (int64_t lf = f->leaf_off[l]; lf < f->leaf_off[l + 1]; ++lf) {
      int64_t b = f->leaf_box[lf], s0 = f->src_start[lf], s1 = f->src_start[lf + 1];
      orc_h_p2m(p, kh[2 * l], kh[2 * l + 1], zero3, 1.0, f->us + 3 * s0, q + 2 * s0, s1 - s0, mult[l] + b * w);
    }
    if (l >= 1)
      for (int o = 0; o < 8; ++o) {
        orc_h_m2m(p, kh[2 * (l - 1)], kh[2 * (l - 1) + 1], o, T);
        for (int64_t b = 0; b < f->nbox[l - 1]; ++b) {
          int64_t c0 = f->first_child[l - 1][b];
          if (c0 >= 0) orc_h_apply(p, T, mult[l] + (c0 + o) * w, mult[l - 1] + b * w);
        }
      }
  }
  /* downward */
  char *have = malloc(316);
  double *m2l = malloc(sizeof(double) * 2 * n2 * 316);
  for (int l = 1; l <= L; ++l) {
    for (int o = 0; o < 8; ++o) {
      orc_h_l2l(p, kh[2 * (l - 1)], kh[2 * (l - 1) + 1], o, T);
      for (int64_t b = 0; b < f->nbox[l]; ++b) {
        int64_t d[4];
        orc_decode(f->keys[l][b], d);
        if ((int)(((d[1] & 1) << 2) | ((d[2] & 1) << 1) | (d[3] & 1)) != o) continue;
        int64_t pb = orc_find_key(f->keys[l - 1], f->nbox[l - 1], orc_parent(f->keys[l][b]));
        orc_h_apply(p, T, loc[l - 1] + pb * w, loc[l] + b * w);
      }
    }
    memset(have, 0, 316);
    for (int64_t b = 0; b < f->nbox[l]; ++b)
      for (int64_t e = f->v[l].off[b]; e < f->v[l].off[b + 1]; ++e) {
        int t = (int)f->v[l].aux[e];
        if (!have[t]) {
          int dv[3];
          orc_offset_of_index(t, dv);
          orc_h_m2l(p, kh[2 * l], kh[2 * l + 1], dv, m2l + 2 * n2 * t);
          have[t] = 1;
        }
        orc_h_apply(p, m2l + 2 * n2 * t, mult[l] + f->v[l].ent[e] * w, loc[l] + b * w);
      }
    for (int64_t b = 0; b < f->nbox[l]; ++b)
      for (int64_t e = f->x[l].off[b]; e < f->x[l].off[b + 1]; ++e) {
        int64_t sl = f->x[l].ent[e], s0 = f->src_start[sl], s1 = f->src_start[sl + 1];
        double c[3], r;
        orc_relative_frame(f->keys[l][b], f->leaf_key[sl], c, &r); /* ĉ(t|s), r̂(t|s) */
        orc_h_p2l(p, kh[2 * l], kh[2 * l + 1], c, r, f->us + 3 * s0, q + 2 * s0, s1 - s0, loc[l] + b * w);
      }
  }
  free(have);
  free(m2l);
  /* leaves */
  for (int l = 0; l <= L; ++l) {
    for (int64_t lf = f->leaf_off[l]; lf < f->leaf_off[l + 1]; ++lf) {
      int64_t t0 = f->tgt_start[lf], t1 = f->tgt_start[lf + 1], b = f->leaf_box[lf];
      orc_h_l2p(p, kh[2 * l], kh[2 * l + 1], zero3, 1.0, loc[l] + b * w, f->ut + 3 * t0, t1 - t0, ph + 2 * t0,
                gp ? gp + 6 * t0 : NULL);
    }
    for (int64_t lf = f->leaf_off[l]; lf < f->leaf_off[l + 1]; ++lf) {
      int64_t r = lf - f->leaf_off[l], t0 = f->tgt_start[lf], t1 = f->tgt_start[lf + 1];
      for (int64_t e = f->w[l].off[r]; e < f->w[l].off[r + 1]; ++e) {
        int64_t sb = f->w[l].ent[e];
        int sl = (int)f->w[l].aux[e];
        double c[3], rr;
        orc_relative_frame(f->keys[sl][sb], f->leaf_key[lf], c, &rr); /* ĉ(s|t), r̂(s|t) */
        orc_h_m2p(p, kh[2 * sl], kh[2 * sl + 1], c, rr, mult[sl] + sb * w, f->ut + 3 * t0, t1 - t0, ph + 2 * t0,
                  gp ? gp + 6 * t0 : NULL);
      }
    }
    for (int64_t lf = f->leaf_off[l]; lf < f->leaf_off[l + 1]; ++lf) {
      int64_t r = lf - f->leaf_off[l], t0 = f->tgt_start[lf], t1 = f->tgt_start[lf + 1];
      for (int64_t e = f->u[l].off[r]; e < f->u[l].off[r + 1]; ++e) {
        int64_t sl = f->u[l].ent[e], s0 = f->src_start[sl], s1 = f->src_start[sl + 1];
        double *y = malloc(sizeof(double) * 3 * (s1 - s0 + 1));
        if (sl == lf) {
          memcpy(y, f->us + 3 * s0, sizeof(double) * 3 * (s1 - s0));
        } else {
          double c[3], rr;
          orc_relative_frame(f->leaf_key[sl], f->leaf_key[lf], c, &rr); /* ĉ(s|t), r̂(s|t) */
          for (int64_t j = 0; j < s1 - s0; ++j)
            for (int k = 0; k < 3; ++k) y[3 * j + k] = c[k] + rr * f->us[3 * (s0 + j) + k];
        }
        orc_helmholtz_direct_sum(kh[2 * l], kh[2 * l + 1], y, q + 2 * s0, s1 - s0, f->ut + 3 * t0, t1 - t0,
                                 ph + 2 * t0, gp ? gp + 6 * t0 : NULL);
        free(y);
      }
    }
  }
  /* output pass */
  for (int64_t lf = 0; lf < f->nleaves; ++lf) {
    int lvl = key_level(f->leaf_key[lf]);
    double rt = f->side / ldexp(1.0, lvl + 1);
    double s1 = 4.0 * 3.141592653589793 * rt, s2 = 4.0 * 3.141592653589793 * rt * rt;
    for (int64_t i = f->tgt_start[lf]; i < f->tgt_start[lf + 1]; ++i) {
      int64_t c = f->tgt_order[i];
      phi[2 * c] = ph[2 * i] / s1;
      phi[2 * c + 1] = ph[2 * i + 1] / s1;
      if (grad && f->gradients)
        for (int k = 0; k < 6; ++k) grad[6 * c + k] = gr[6 * i + k] / s2;
    }
  }
  for (int l = 0; l <= L; ++l) {
    free(mult[l]);
    free(loc[l]);
  }
  free(mult);
  free(loc);
  free(kh);
  free(q);
  free(ph);
  free(gr);
  free(T);
}

/* ---- inspection --------------------------------------------------------------------- */

int orc_fmm_nlevels(const orc_fmm *f) { return f->nlevels; }

int orc_fmm_expansions(const orc_fmm *f, int locals, int level, double *out) {
  if (!f->mult || level < 0 || level >= f->nlevels) return -1;
  memcpy(out, locals ? f->loc[level] : f->mult[level], sizeof(double) * (size_t)f->nbox[level] * nc(f->p));
  return 0;
}

double orc_fmm_domain(const orc_fmm *f, double box[6]) {
  memcpy(box, f->box, sizeof(double) * 6);
  return f->side;
}

int64_t orc_fmm_nboxes(const orc_fmm *f, int level) {
  return level < f->nlevels ? f->nbox[level] : 0;
}

void orc_fmm_box_keys(const orc_fmm *f, int level, uint64_t *out) {
  memcpy(out, f->keys[level], sizeof(uint64_t) * f->nbox[level]);
}

int64_t orc_fmm_nleaves(const orc_fmm *f) { return f->nleaves; }

void orc_fmm_leaf_keys(const orc_fmm *f, uint64_t *out) {
  memcpy(out, f->leaf_key, sizeof(uint64_t) * f->nleaves);
}

int64_t orc_fmm_list(const orc_fmm *f, int kind, int level, int64_t *offsets, int64_t *entries,
                     int64_t *aux) {
  if (level >= f->nlevels) return 0;
  const csr *c = kind == 0 ? &f->u[level] : kind == 1 ? &f->v[level] : kind == 2 ? &f->w[level] : &f->x[level];
  int64_t rows = (kind == 0 || kind == 2) ? f->leaf_off[level + 1] - f->leaf_off[level] : f->nbox[level];
  int64_t total = c->off[rows];
  if (offsets) memcpy(offsets, c->off, sizeof(int64_t) * (rows + 1));
  if (entries) memcpy(entries, c->ent, sizeof(int64_t) * total);
  if (aux && (kind == 1 || kind == 2)) memcpy(aux, c->aux, sizeof(int64_t) * total);
  return total;
}

void orc_fmm_leaf_coordinates(const orc_fmm *f, int targets, double *u, int64_t *order) {
  int64_t n = targets ? f->nt : f->ns;
  memcpy(u, targets ? f->ut : f->us, sizeof(double) * 3 * n);
  memcpy(order, targets ? f->tgt_order : f->src_order, sizeof(int64_t) * n);
}

void orc_fmm_point_leaves(const orc_fmm *f, int targets, int64_t *out) {
  memcpy(out, targets ? f->tgt_leaf : f->src_leaf, sizeof(int64_t) * (targets ? f->nt : f->ns));
}
