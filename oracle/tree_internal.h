/*
 * tree_internal.h -- oracle (test infrastructure): the one-rank tree, lists and points of
 * an orc_fmm, shared by tree.c (build, serial evaluation) and cpu_path.c (the
 * multi-threaded restatement of the reference's host evaluation used as the CPU
 * baseline).  Not part of the library.
 */
#ifndef ORACLE_TREE_INTERNAL_H
#define ORACLE_TREE_INTERNAL_H

#include <stdint.h>

typedef struct {
  int64_t *off, *ent, *aux; /* CSR rows; aux = offset index for V */
} csr;

struct orc_fmm {
  int p, gradients, nlevels;
  double box[6], lower[3], side;
  int64_t ns, nt;
  /* box index per level */
  int64_t *nbox;
  uint64_t **keys;
  int64_t **box_leaf;  /* leaf index or -1 */
  int64_t **first_child; /* box index of child 0 on level+1 or -1 */
  /* leaves by (level, key) */
  int64_t nleaves;
  uint64_t *leaf_key;
  int64_t *leaf_box, *leaf_off; /* leaf_off[nlevels+1] */
  /* lists */
  csr *v, *x;  /* per level, rows = boxes of the level */
  csr *u, *w;  /* per level, rows = leaves of the level (near = U + self) */
  /* points */
  int64_t *src_start, *tgt_start; /* per leaf, CSR */
  int64_t *src_order, *tgt_order; /* leaf order -> caller index */
  int64_t *src_leaf, *tgt_leaf;   /* caller index -> leaf */
  double *us, *ut;                /* leaf-scaled coordinates, leaf order */
  double *m2m, *l2l, *m2l;        /* tables */
  double **mult, **loc;           /* expansions of the last evaluation, per level */
};

static inline int orc_key_level(uint64_t k) { return (int)(k & 0x7FFFULL); }
/* binary search of key in the ascending v[0..n): its index, or -1 */
int64_t orc_find_key(const uint64_t *v, int64_t n, uint64_t key);
/* ĉ(s|t), r̂(s|t) of CONVENTIONS §3.13, exact from the integer cells */
void orc_relative_frame(uint64_t s, uint64_t t, double c[3], double *r);

#endif
