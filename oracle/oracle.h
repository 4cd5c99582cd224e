/*
 * oracle.h -- CPU restatement of the reference FMM (tbetcke/fmm) for the evaluation
 * hot path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the GPU path
 * is compared against.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library never links or calls it.
 *
 * Everything is f64 and serial, written from the reference's conventions
 * (docs/CONVENTIONS.md §3.1-§3.13) and its reference crates:
 *   rng / points      fmm-validate/src/rng.rs, fmm-validate/src/points.rs
 *   direct sum        fmm-ref/src/p2p.rs:147-205 (Neumaier :407-431)
 *   harmonics         fmm-math/src/harmonics.rs:118-420, CONVENTIONS §3.5-§3.6
 *   leaf operators    fmm-ref/src/leaf.rs:122-409, CONVENTIONS §3.7, §3.11
 *   translations      fmm-ref/src/direct.rs:75-210, CONVENTIONS §3.11
 *   tables            fmm-tables/src/m2l.rs:114, octant.rs, CONVENTIONS §3.12
 *   morton / tree     octree/src/morton.rs, octree/src/octree.rs:1278-1420,
 *                     octree/src/octree/implementation.rs:682-716,1014-1120
 *   lists / index     fmm-plan/src/interaction_manager.rs:1-60, index.rs:1-80
 *   evaluation order  fmm-plan/src/operator.rs:408-425, evaluator.rs:873-881
 *   leaf geometry     fmm-exec/src/geometry.rs:253-317, CONVENTIONS §3.13
 *   output pass       fmm-exec/src/fmm.rs:117-145, CONVENTIONS §3.13 "Output"
 *   metrics           fmm-validate/src/metrics.rs:51-170
 */
#ifndef FMM_ORACLE_H
#define FMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- rng.c: SplitMix64 and the point distributions ---------------------- */
uint64_t orc_next_u64(uint64_t *state);
double orc_uniform(uint64_t *state);
double orc_range(uint64_t *state, double lo, double hi);
int64_t orc_below(uint64_t *state, int64_t n);
double orc_normal(uint64_t *state);
void orc_cube(uint64_t *state, int64_t n, const double centre[3], double half, double *out);
void orc_ball(uint64_t *state, int64_t n, const double centre[3], double radius, double *out);
void orc_sphere(uint64_t *state, int64_t n, const double centre[3], double radius, double *out);
void orc_plummer(uint64_t *state, int64_t n, const double centre[3], double scale, double *out);
void orc_gaussian_clusters(uint64_t *state, int64_t n, const double *centres, int64_t k,
                           double width, double *out);
void orc_charges(uint64_t *state, int64_t n, double *out);
void orc_normals(uint64_t *state, int64_t n, double *out);
void orc_sample(uint64_t *state, int64_t n, int64_t sampled, int64_t *out);

/* ---- direct.c: P2P reference terms and the compensated direct sum -------- */
void orc_p2p(const double *src, const double *q, int64_t ns, const double *tgt, int64_t nt,
             double *phi, double *grad);
void orc_direct_sum(const double *src, const double *q, int64_t ns, const double *tgt,
                    int64_t nt, double *phi, double *grad);
double orc_neumaier_sum(const double *terms, int64_t n);
void orc_error_norms(const double *approx, const double *exact, int64_t n, int vec3,
                     double out[2]);

/* ---- harmonics.c: solid harmonics, leaf operators, translations, tables --- */
void orc_regular(int p, const double x[3], double *out);
void orc_irregular(int p, const double x[3], double *out);
void orc_regular_grad(int p, const double x[3], double *gx, double *gy, double *gz);
void orc_irregular_grad(int p, const double x[3], double *gx, double *gy, double *gz);
void orc_p2m(int p, const double c[3], double r, const double *src, const double *q,
             int64_t n, double *m);
void orc_p2l(int p, const double c[3], double r, const double *src, const double *q,
             int64_t n, double *l);
void orc_l2p(int p, const double c[3], double r, const double *l, const double *tgt,
             int64_t n, double *phi, double *grad);
void orc_m2p(int p, const double c[3], double r, const double *m, const double *tgt,
             int64_t n, double *phi, double *grad);
void orc_m2m(int p, const double b[3], double rho, const double *in, double *out);
void orc_l2l(int p, const double t[3], double sigma, const double *in, double *out);
void orc_m2l(int p, const double b[3], double sigma, const double *in, double *out);
int orc_offset_index(int dx, int dy, int dz);
void orc_offset_of_index(int index, int d[3]);
void orc_table_m2m(int p, double *out);
void orc_table_l2l(int p, double *out);
void orc_table_m2l(int p, double *out);

/* ---- tree.c: Morton keys, octree, lists ---------------------------------- */
uint64_t orc_key(int level, int64_t ix, int64_t iy, int64_t iz);
void orc_decode(uint64_t key, int64_t out[4]);
uint64_t orc_parent(uint64_t key);
int orc_adjacent(uint64_t a, uint64_t b);
void orc_bounding_box(const double *pts, int64_t n, double out[6]);
uint64_t orc_point_key(const double x[3], const double box[6], int level);

typedef struct orc_fmm orc_fmm;

/* Build the tree, lists and tables of an FMM over sources and targets
 * (FmmBuilder::build, fmm-exec/src/fmm.rs:1853, on one rank).  near_w / near_x: the
 * near rule's named crossovers (NearRuleChoice::Crossovers), negative for off.
 * Returns NULL on a degenerate extent. */
orc_fmm *orc_fmm_new(const double *src, int64_t ns, const double *tgt, int64_t nt, int p,
                     int64_t leaf_points, int max_level, int gradients, double near_w,
                     double near_x);
void orc_fmm_free(orc_fmm *f);
/* Evaluate with one charge vector; output in caller order: phi[nt], grad[3 nt]. */
void orc_fmm_evaluate(orc_fmm *f, const double *q, double *phi, double *grad);
/* The CPU baseline (cpu_path.c): the reference's multi-threaded host evaluation restated
 * (threads over targets, SIMD P2P, batched M2L column kernel) on the same tree; threads
 * <= 0 keeps OpenMP's default.  orc_cpu_isa: 2 = AVX-512F kernels, 0 = scalar loops. */
void orc_fmm_evaluate_cpu(orc_fmm *f, const double *q, double *phi, double *grad, int threads);
int orc_cpu_isa(void);
int orc_cpu_threads(void);

/* Inspection for the parity tests. */
/* Multipoles (locals = 0) or locals of `level` after the last evaluation, box-major. */
int orc_fmm_expansions(const orc_fmm *f, int locals, int level, double *out);
int orc_fmm_nlevels(const orc_fmm *f);
double orc_fmm_domain(const orc_fmm *f, double box[6]);
int64_t orc_fmm_nboxes(const orc_fmm *f, int level);
void orc_fmm_box_keys(const orc_fmm *f, int level, uint64_t *out);
int64_t orc_fmm_nleaves(const orc_fmm *f);
void orc_fmm_leaf_keys(const orc_fmm *f, uint64_t *out);
/* kind: 0 = near (U + self, leaf numbering), 1 = V (box, by offset), 2 = W (box on
 * level+1), 3 = X (leaf numbering); rows of boxes of `level` (V, X) or of the leaves
 * of `level` (near, W).  Returns the total entry count; fills offsets[nrows+1] and
 * entries (and, for V, offsets indices in aux) when the pointers are non-null. */
int64_t orc_fmm_list(const orc_fmm *f, int kind, int level, int64_t *offsets,
                     int64_t *entries, int64_t *aux);
/* Leaf of every source / target point (leaf numbering), in caller order. */
void orc_fmm_point_leaves(const orc_fmm *f, int targets, int64_t *out);
/* Leaf-scaled coordinates in leaf order and the caller index of each position. */
void orc_fmm_leaf_coordinates(const orc_fmm *f, int targets, double *u, int64_t *order);

#ifdef __cplusplus
}
#endif
/* Helmholtz: compensated f64 direct sum of e^{ikr}/r (no 1/(4 pi)); complex values
 * interleaved (re, im): q[2 ns], pot[2 nt], grad[6 nt] (x re, x im, y re, ...) or NULL. */
void orc_helmholtz_direct_sum(double kre, double kim, const double *src, const double *q, int64_t ns,
                              const double *tgt, int64_t nt, double *pot, double *grad);

/* Helmholtz expansions (hexpansion.c): level units, stored form; complex expansions and
 * tables as two real planes (real, then imaginary). */
void orc_h_solid(int p, const double x[3], double *out);
void orc_h_solid_grad(int p, const double x[3], double *gx, double *gy, double *gz);
void orc_h_radial(int P, double zre, double zim, double *j, double *h);
void orc_h_p2m(int p, double kre, double kim, const double c[3], double r, const double *src, const double *q,
               int64_t n, double *out);
void orc_h_l2p(int p, double kre, double kim, const double c[3], double r, const double *lc, const double *tgt,
               int64_t n, double *phi, double *grad);
void orc_h_p2l(int p, double kre, double kim, const double c[3], double r, const double *src, const double *q,
               int64_t n, double *out);
void orc_h_m2p(int p, double kre, double kim, const double c[3], double r, const double *mc, const double *tgt,
               int64_t n, double *phi, double *grad);
void orc_h_rotation_blocks(int p, const double Q[9], double *out);
void orc_h_coaxial(int p, int singular_start, double kre, double kim, double t, double *out);
void orc_h_m2m(int p, double kre, double kim, int o, double *out);
void orc_h_l2l(int p, double kre, double kim, int o, double *out);
void orc_h_m2l(int p, double kre, double kim, const int d[3], double *out);
void orc_h_apply(int p, const double *T, const double *x, double *y);
void orc_hfmm_evaluate(orc_fmm *f, double kre, double kim, const double *q, double *phi, double *grad);

#endif
