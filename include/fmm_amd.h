/*
 * fmm_amd.h -- C ABI of the MI355X FMM evaluation path.
 *
 * The boundary replaces, for one rank and one GPU, the reference's device operator and
 * its kernel wrappers (tbetcke/fmm):
 *   - the level-batched operator interface FmmOperator (fmm-plan/src/operator.rs:434-497,
 *     implemented for the device by DeviceOperator, fmm-exec/src/device.rs:5212-5446):
 *     fmmh_p2m ... fmmh_p2p below, one call per level and kind, every call adding (+=)
 *     into its output, with HostData::Reset (operator.rs:326-336) as fmmh_reset;
 *   - the evaluation entry points of fmm-exec (Fmm::evaluate, fmm-exec/src/fmm.rs:3691;
 *     DeviceOperator::begin_evaluation, device.rs:4386; read_output, device.rs:4687):
 *     fmmh_load_charges, fmmh_evaluate, fmmh_read_output;
 *   - the host build of the tree and lists that stays on the CPU (FmmBuilder::build,
 *     fmm-exec/src/fmm.rs:1853-1935, on one rank): fmmh_plan_build and its accessors.
 *
 * Plain pointers and sizes only.  Every function returns an fmmh_status (0 = success);
 * fmmh_last_error() describes the last failure of the calling thread.  Status codes
 * mirror KernelError (fmm-kernels/src/error.rs:12-72).  A failed launch surfaces at the
 * next synchronising call (fmmh_read_output), as in the reference (device.rs:414-425).
 * Everything is allocated at fmmh_operator_create; evaluations allocate nothing.
 * One operator = one HIP stream; calls on one operator must not overlap.
 */
#ifndef FMM_AMD_H
#define FMM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  FMMH_OK = 0,
  FMMH_NOT_COMPILED = 1,
  FMMH_NO_DEVICE = 2,
  FMMH_UNSUPPORTED_PRECISION = 3,
  FMMH_OUT_OF_MEMORY = 4,
  FMMH_TOO_LARGE = 5,
  FMMH_WRONG_DEVICE = 6,
  FMMH_UNSUPPORTED_LAYOUT = 7,
  FMMH_DEVICE = 8,
  FMMH_INVALID_ARGUMENT = 9
} fmmh_status;

typedef enum {
  FMMH_PLAIN_F64 = 0, /* every operator in f64 */
  FMMH_MIXED = 1      /* f32 P2P (double-float separations), bf16x3 M2L on MFMA; f64 elsewhere */
} fmmh_precision;

/* The M2L formulation (M2lMethod): dense tables (fmm-kernels/src/translate.rs:1703),
 * rotation, coaxial shift and rotation back (fmm-kernels/src/rotation.rs:441; in the mixed
 * precision in f32), or plane waves in six directions by merge and shift
 * (fmm-kernels/src/plane_wave.rs:1-60; f64 in both precisions, on FMM3D's quadrature of
 * the degree, whose error is about a tenth of the truncation's). */
typedef enum { FMMH_M2L_DENSE = 0, FMMH_M2L_ROTATION = 1, FMMH_M2L_PLANE_WAVE = 2 } fmmh_m2l_method;

typedef enum { FMMH_LIST_NEAR = 0, FMMH_LIST_V = 1, FMMH_LIST_W = 2, FMMH_LIST_X = 3 } fmmh_list_kind;

/* Operator kinds for timings: the order of fmm-bench's kind table. */
typedef enum {
  FMMH_KIND_P2M = 0,
  FMMH_KIND_M2M = 1,
  FMMH_KIND_L2L = 2,
  FMMH_KIND_M2L = 3,
  FMMH_KIND_P2L = 4,
  FMMH_KIND_L2P = 5,
  FMMH_KIND_M2P = 6,
  FMMH_KIND_P2P = 7,
  FMMH_KIND_OTHER = 8,
  FMMH_KIND_M2L_CONTRACT = 9, /* the contraction kernel alone, inside FMMH_KIND_M2L */
  FMMH_NKINDS = 10
} fmmh_kind;

typedef struct fmmh_plan fmmh_plan;
typedef struct fmmh_operator fmmh_operator;

/* FmmBuilder settings of the tree (fmm-exec/src/fmm.rs:1302-1760). */
typedef struct {
  int64_t leaf_points; /* max_points_per_leaf: refine while a box holds more keys */
  int32_t max_level;   /* deepest level, at most 16 */
  double near_w;       /* near rule crossovers (NearRuleChoice::Crossovers); < 0: off */
  double near_x;
} fmmh_plan_options;

typedef struct {
  int32_t p;         /* expansion degree, 0..15 on the device */
  int32_t precision; /* fmmh_precision */
  int32_t gradients; /* 1: potential and gradient, 0: potential only */
  int32_t device;    /* HIP device ordinal */
  int32_t m2l;       /* fmmh_m2l_method */
} fmmh_operator_options;

typedef struct {
  int32_t nlevels;
  int64_t nleaves;
  int64_t nboxes;
  uint64_t p2p_pairs;  /* n_s n_t over near pairs, coincident included */
  uint64_t m2l_pairs;  /* V-list pairs */
  uint64_t m2l_slots;  /* column x group slots the M2L contraction issues */
  uint64_t device_bytes;
} fmmh_stats;

const char *fmmh_last_error(void);
int32_t fmmh_abi_version(void);

/* ---- plan (host) --------------------------------------------------------------- */
/* sources/targets: 3 doubles per point (x, y, z); targets may equal sources. */
int32_t fmmh_plan_build(const double *sources, int64_t nsources, const double *targets, int64_t ntargets,
                        const fmmh_plan_options *options, fmmh_plan **out);
/* fmmh_plan_build with its per-point stages on GPU `device` (octree/src/octree/
 * implementation.rs:682-716,1014): the bounding box, the points' level-16 Morton keys, their
 * sort and distinct set (device radix sort and unique), then, from the tree and lists the
 * host derives from those keys, each point's leaf, the points in leaf order and their
 * leaf-scaled coordinates.  Host point arrays; the same plan bit for bit. */
int32_t fmmh_plan_build_device(int32_t device, const double *sources, int64_t nsources, const double *targets,
                               int64_t ntargets, const fmmh_plan_options *options, fmmh_plan **out);
/* Batched independent clouds as one plan (a forest): nclouds clouds of counts[c] points,
 * concatenated in points (3 doubles each), sources = targets within each cloud.  Each
 * cloud is refined and listed as if alone (its own domain); one operator on the forest
 * evaluates every cloud in the same launches; outputs and charges follow the points'
 * order.  fmmh_plan_domain reports the first cloud's domain.  No counterpart in the
 * reference (its batched-clouds configuration is a north-star addition); it stands where
 * FmmBuilder::build (fmm-exec/src/fmm.rs:1853-1935) would be called once per cloud. */
int32_t fmmh_forest_build(const double *points, const int64_t *counts, int32_t nclouds,
                          const fmmh_plan_options *options, fmmh_plan **out);
void fmmh_plan_free(fmmh_plan *plan);
int32_t fmmh_plan_nlevels(const fmmh_plan *plan);
int64_t fmmh_plan_nboxes(const fmmh_plan *plan, int32_t level);
int64_t fmmh_plan_nleaves(const fmmh_plan *plan);
/* box[6] = [xmin, ymin, zmin, xmax, ymax, zmax]; *side = w */
int32_t fmmh_plan_domain(const fmmh_plan *plan, double *box, double *side);
int32_t fmmh_plan_box_keys(const fmmh_plan *plan, int32_t level, uint64_t *out);
int32_t fmmh_plan_leaf_keys(const fmmh_plan *plan, uint64_t *out);
/* Rows of `kind` on `level` (near, W: leaves of the level; V, X: boxes of the level).
 * Returns the entry count (or -1); fills offsets[rows + 1], entries and, for V, the
 * offset indices in aux when the pointers are not NULL. */
int64_t fmmh_plan_list(const fmmh_plan *plan, int32_t kind, int32_t level, uint32_t *offsets, uint32_t *entries,
                       uint16_t *aux);
/* Leaf of every point in caller order. */
int32_t fmmh_plan_point_leaves(const fmmh_plan *plan, int32_t targets, uint32_t *out);
/* Leaf-scaled coordinates (3 per point, leaf order) and the caller index of each. */
int32_t fmmh_plan_leaf_coordinates(const fmmh_plan *plan, int32_t targets, double *coords, uint32_t *order);

/* Cell of every leaf (4 per leaf: level, ix, iy, iz; fmm-kernels/src/p2p.rs:320-337) and the
 * points of every leaf (CSR over the leaves, nleaves + 1 offsets into leaf order). */
int32_t fmmh_plan_leaf_cells(const fmmh_plan *plan, uint32_t *out);
/* Cell of every box of one level (4 per box, the box index's order). */
int32_t fmmh_plan_box_cells(const fmmh_plan *plan, int32_t level, uint32_t *out);
int32_t fmmh_plan_point_starts(const fmmh_plan *plan, int32_t targets, uint32_t *out);

/* Operator tables of degree p (column-major, 8 / 8 / 316 matrices of (p+1)^4 doubles),
 * through the table cache when one is set. */
int32_t fmmh_tables(int32_t p, double *m2m, double *l2l, double *m2l);
/* The on-disk table cache of this process (fmm-tables/src/cache.rs, its file format:
 * "<kind>-p<pp>-f64-c4-v1.bin", a 44-byte header with magic NDFMMTAB, versions, key,
 * payload length and FNV-1a checksum, then the matrices): operators and fmmh_tables load
 * the M2M, L2L and dense M2L tables of degree p from `directory` when its files are valid,
 * else build and store them (temporary file, then rename).  NULL disables the cache. */
int32_t fmmh_set_table_cache(const char *directory);
/* Loads that hit (every file of a degree valid) or missed since fmmh_set_table_cache. */
int32_t fmmh_table_cache_stats(int64_t *hits, int64_t *misses);
/* The M2L of offset index `offset` composed from the rotation-M2L tables, G(Q^T) A(|d|)
 * F(Q) (fmm-tables/src/rotation.rs:307 ShiftTables::apply), as a dense column-major
 * (p+1)^2 x (p+1)^2 matrix: equal to the dense table to rounding. */
int32_t fmmh_m2l_rotation_table(int32_t p, int32_t offset, double *table);

/* Launch-shape decisions of an operator (0 everywhere: the static rule), as the
 * autotuner takes them.  Per level index 0..16. */
typedef struct {
  int32_t p2p_uniform_min; /* mixed P2P: chunks of at least this many targets read sources by
                              uniform loads, smaller ones through LDS tiles (0: 33) */
  int32_t m2l_stage;       /* mixed dense M2L: 0 or 1 one k-step per pipeline stage, 2 two k-steps
                              in a ring of two LDS buffers (where the table fits) */
  int32_t m2l_mode[17];    /* dense M2L of the level: 0 static rule, 1 box mode, 2 clustered */
  int32_t l2l_width[17];   /* L2L into the level: 0 static rule, or 16, 64, 128 columns per task */
  int32_t m2m_form[17];    /* M2M into the parents of the level: 0 static, 1 per-tile, 2 split */
} fmmh_tuning;

/* ---- device operator -------------------------------------------------------------- */
int32_t fmmh_device_count(int32_t *count);
/* The device's architecture name and compute-unit count, as "<gcnArchName>/<CUs>cu"
 * (for example "gfx950:sramecc+:xnack-/256cu"), NUL-terminated in name[capacity]: the key
 * under which tuned decisions are cached.  FMMH_INVALID_ARGUMENT for a device out of
 * range or a buffer too small. */
int32_t fmmh_device_describe(int32_t device, char *name, int32_t capacity);
int32_t fmmh_operator_create(const fmmh_plan *plan, const fmmh_operator_options *options, fmmh_operator **out);
/* fmmh_operator_create with launch-shape decisions (fmmh_autotune's, or a cache's). */
int32_t fmmh_operator_create_tuned(const fmmh_plan *plan, const fmmh_operator_options *options,
                                   const fmmh_tuning *tuning, fmmh_operator **out);
/* The autotune of the device path (fmm-exec/src/tune.rs:1-110): times each decision's
 * candidates on operators built over the plan (the static rule's candidate first), on its
 * own levels, within budget_seconds, and writes the winners (0 where the static rule was
 * kept).  report (may be NULL): 5 x 17 x 5 doubles, ms per call of every candidate timed
 * ([kind: P2P, M2L, L2L, M2M, M2L stage][level][candidate]), NaN where not timed.  The cache file of
 * decisions is the caller's (fmm_amd/tune.py). */
int32_t fmmh_autotune(const fmmh_plan *plan, const fmmh_operator_options *options, double budget_seconds,
                      fmmh_tuning *out, double *report);
void fmmh_operator_free(fmmh_operator *op);
int32_t fmmh_operator_stats(const fmmh_operator *op, fmmh_stats *out);
/* Charges of the sources in caller order, on the host (on_device = 0) or on the GPU;
 * count must equal the plan's source count (begin_evaluation asserts charges.len(),
 * fmm-exec/src/device.rs:4386-4390), else FMMH_INVALID_ARGUMENT and nothing is read. */
int32_t fmmh_load_charges(fmmh_operator *op, const double *charges, int64_t count, int32_t on_device);
/* HostData::Reset: zero multipoles, locals and target output. */
int32_t fmmh_reset(fmmh_operator *op);
/* FmmOperator level calls (fmm-plan/src/operator.rs:434-485). */
int32_t fmmh_p2m(fmmh_operator *op, int32_t level);
int32_t fmmh_m2m(fmmh_operator *op, int32_t level);
int32_t fmmh_m2l(fmmh_operator *op, int32_t level);
int32_t fmmh_p2l(fmmh_operator *op, int32_t level);
int32_t fmmh_l2l(fmmh_operator *op, int32_t level);
int32_t fmmh_l2p(fmmh_operator *op, int32_t level);
int32_t fmmh_m2p(fmmh_operator *op, int32_t level);
int32_t fmmh_p2p(fmmh_operator *op, int32_t level);
/* Every level call in the evaluator's order (fmm-plan/src/evaluator.rs:873-881). */
int32_t fmmh_evaluate(fmmh_operator *op);
/* fmmh_reset + fmmh_evaluate as one hipGraph launch, captured on the first call and
 * replayed after (kind timing must be off).  No counterpart in the reference: it replaces
 * the per-launch host work of an evaluation (DeviceOperator's launches,
 * fmm-exec/src/device.rs:5212-5446) by one graph launch. */
int32_t fmmh_evaluate_graph(fmmh_operator *op);
/* Output pass and download: phi[nt], grad[3 nt] (may be NULL) in caller order, to the
 * host (on_device = 0) or to device buffers; count must equal nt (read_output,
 * fmm-exec/src/device.rs:4687).  Synchronises. */
int32_t fmmh_read_output(fmmh_operator *op, double *phi, double *grad, int64_t count, int32_t on_device);
/* Adjoint (vector-Jacobian product) of the potentials, for a plan built with one array
 * for sources and targets.  Inputs in caller order: charges q[n], the cotangent
 * g[n] = dL/dphi (e.g. 2 phi / n for L = sum phi^2 / n) and the forward gradient
 * field[3 n] = grad phi.  Outputs: grad_charges[n] = dL/dq_j = psi(z_j) and, when
 * grad_points is not NULL, grad_points[3 n] = dL/dz_j = q_j grad psi(z_j) + g_j grad phi(z_j),
 * psi the potential of the charges g.  One more evaluation (the operator's charges and
 * expansions are overwritten); the operator must compute gradients.  No counterpart in
 * the reference's Rust API (its north-star addition: the adjoint of Fmm::evaluate,
 * fmm-exec/src/fmm.rs:3691).  Synchronises. */
int32_t fmmh_vjp(fmmh_operator *op, const double *charges, const double *cotangent, const double *field,
                 int64_t count, int32_t on_device, double *grad_charges, double *grad_points);
/* Order the operator's stream after the work queued on `stream` (a hipStream_t; NULL is
 * the null stream) so far: call it before passing device pointers that another stream
 * writes (charges, cotangents).  Outputs are complete when fmmh_read_output or fmmh_vjp
 * returns.  No counterpart in the reference (its Device owns every buffer it reads,
 * fmm-kernels/src/error.rs:56-57); it is the ordering half of accepting a caller's
 * device buffers. */
int32_t fmmh_stream_wait(fmmh_operator *op, void *stream);
int32_t fmmh_sync(fmmh_operator *op);
/* The compensated f64 direct sum of nd_fmm_ref::p2p::direct_sum (fmm-ref/src/p2p.rs:147-205):
 * per target phi = sum q / r and grad = -sum q d / r^3 (d = target - source, no 1/(4 pi)),
 * Neumaier-compensated (:407-431), a source equal to the target skipped.  Points 3
 * doubles each; grad may be NULL; host arrays (on_device = 0) or device pointers on
 * `device`.  Synchronous; allocates its own scratch. */
int32_t fmmh_direct_sum(int32_t device, const double *sources, const double *charges, int64_t nsources,
                        const double *targets, int64_t ntargets, double *phi, double *grad, int32_t on_device);
/* fmmh_direct_sum without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Device pointers only; queued on `stream` (a
 * hipStream_t, NULL = the null stream) and not synchronised: the outputs are ready when
 * the caller's stream reaches them.  The partial sums use a fixed 32 MiB array per GPU,
 * held until the process ends: at most 262,144 targets per call (FMMH_TOO_LARGE above).
 * Calls that overlap on one GPU share the array and race: order them (one stream, or
 * events).  The same source chunks and the same bits as fmmh_direct_sum. */
int32_t fmmh_direct_sum_noalloc(int32_t device, const double *sources, const double *charges, int64_t nsources,
                                const double *targets, int64_t ntargets, double *phi, double *grad, void *stream);
/* Multipoles (locals = 0) or locals of one level, box-major (p+1)^2 doubles per box in
 * the box index's order: the values an operator hands to the host store at the "send"
 * events of HostData (fmm-plan/src/operator.rs:344-404).  Synchronises. */
int32_t fmmh_read_expansions(fmmh_operator *op, int32_t locals, int32_t level, double *out);
/* Per-kind device times (HIP events on the operator's stream), summed since the last
 * call with enable: ms[FMMH_NKINDS]. */
int32_t fmmh_kind_timing(fmmh_operator *op, int32_t enable);
int32_t fmmh_kind_times(fmmh_operator *op, double *ms, int64_t *launches);
/* Errors of an evaluation (fmm-exec/src/device.rs:414-425): the operator keeps the first
 * device failure (FMMH_DEVICE) of any call that touches the device (charges, reset, the
 * level calls, evaluate, output, VJP, sync, expansions, kind times, stream wait) and
 * returns it, with its message, from every such call after it without device work; a new
 * operator evaluates again.  Argument errors are returned and not kept.
 * fmmh_inject_error is the test hook (DeviceOperator::inject_error, device.rs:4621, and
 * Fmm::inject_device_error, fmm-exec/src/fmm.rs:4715): the next such call fails with a
 * device error whose message says "injected", which the operator then keeps. */
int32_t fmmh_inject_error(fmmh_operator *op);

/* The plane-wave quadrature of degree p (fmm-tables/src/plane_wave/quadrature.rs:48-104,
 * Quadrature::for_degree / with_nodes): FMM3D's table with `nodes` lambda nodes (2..39), or
 * with the reference's count for degree p (0..20) when nodes is 0.  Returns the node count,
 * the half-node count H (sum of the alpha counts rounded up to even, halved), and per node
 * lambda, weight, FMM3D's alpha count, the Fourier modes kept at degree p (min(F_k, p + 1))
 * and FMM3D's stated maximum error (arrays of >= 39; any output may be NULL but nodes). */
int32_t fmmh_plane_wave_quadrature(int32_t p, int32_t nodes, int32_t *out_nodes, int32_t *half_nodes, double *lambda,
                                   double *weight, int32_t *nalpha, int32_t *fourier, double *max_error);
/* One V pair's plane-wave M2L on the host (PlaneWaveTables::m2l, fmm-tables/src/plane_wave/
 * mod.rs:52-58: one shift from the source to the target): out += M2L_pw(offset) in, scaled
 * storage, in direction `direction` (0..5: up, down, north, south, east, west; -1: the first
 * axis of z, y, x along which the offset is at least two boxes). */
int32_t fmmh_plane_wave_m2l(int32_t p, int32_t nodes, int32_t offset, int32_t direction, const double *in,
                            double *out);
/* One parent's merge and shift on the host (fmm-kernels/src/plane_wave.rs:12-24, the device
 * steps outgoing, merge, gather, local): slot_multipoles holds 216 multipoles (slot a =
 * 36 a_x + 6 a_y + a_z is the box 2P - 2 + a one level down), present[a] says which are held;
 * out (8 x (p+1)^2, child 4 b_x + 2 b_y + b_z) += the plane-wave M2L of every pair slot ->
 * child that is not adjacent. */
int32_t fmmh_plane_wave_merge(int32_t p, int32_t nodes, const double *slot_multipoles, const uint8_t *present,
                              double *out);
/* The direction (0..5) of the merge-and-shift pair slot -> child (geometry.rs
 * Direction::for_slot), -1 if adjacent, -2 for an index out of range. */
int32_t fmmh_plane_wave_slot_direction(int32_t slot, int32_t child);

/* ---- kernels (the per-level launch wrappers of fmm-kernels: p2p.rs:1244, leaf.rs:1851-1951,
 * movement.rs:411,482) ------------------------------------------------------------- */
/* Device pointers, f64, queued on `stream` (NULL: the default stream); nothing is allocated
 * and nothing synchronises.  Every call adds (+=) into its outputs, each output written by
 * one thread in a fixed order (the same bits on every call).  Preconditions, as the
 * reference's kernels take them (fmm-kernels/src/error.rs:7-9): offsets ascending, indices
 * within their arrays, leaf cells of the points' leaves; they are not checked on the device. */
typedef struct {
  uint32_t leaf_first, count;   /* target leaves leaf_first .. leaf_first + count (global indices) */
  const uint32_t *near_off;     /* near CSR over all leaves (global): source leaves of each target leaf */
  const uint32_t *near_ent;
  const uint32_t *leaf_cell;    /* 4 per leaf: level, ix, iy, iz */
  const uint32_t *src_start;    /* CSR of each leaf's sources (leaf order) */
  const uint32_t *tgt_start;    /* CSR of each leaf's targets (leaf order) */
  const double *sx, *sy, *sz;   /* sources in leaf order: leaf-scaled coordinates */
  const double *q;              /* charges in leaf order */
  const double *tx, *ty, *tz;   /* targets in leaf order: leaf-scaled coordinates */
} fmmh_p2p_inputs;
/* P2P (p2p::p2p, fmm-kernels/src/p2p.rs:1244): phi_hat[t] += sum q / |u_t - y^|, grad_hat[3t..]
 * -= sum q (u_t - y^) / |u_t - y^|^3 over each target leaf's near row in row order, in the
 * target leaf's units (relative frames from the cells); grad_hat may be NULL without gradients. */
int32_t fmmh_kernel_p2p(int32_t device, const fmmh_p2p_inputs *in, int32_t gradients, double *phi_hat,
                        double *grad_hat, void *stream);
typedef struct {
  uint32_t rows;                /* output expansions */
  const uint32_t *row_box;      /* output slot of each row: coeffs[row_box[r] (p+1)^2 ..] */
  const uint32_t *row_cell;     /* 4 per row: the output box's level, ix, iy, iz */
  const uint32_t *row_off;      /* CSR of each row's source leaves */
  const uint32_t *row_ent;
  const uint32_t *leaf_cell, *src_start;
  const double *sx, *sy, *sz, *q;
} fmmh_source_inputs;
/* P2M (leaf::p2m) and P2L (leaf::p2l, fmm-kernels/src/leaf.rs:1851,1873): coeffs += the
 * multipole at each row's box of its leaves' sources, or the local of its X-list leaves'. */
int32_t fmmh_kernel_p2m(int32_t device, int32_t p, const fmmh_source_inputs *in, double *coeffs, void *stream);
int32_t fmmh_kernel_p2l(int32_t device, int32_t p, const fmmh_source_inputs *in, double *coeffs, void *stream);
typedef struct {
  uint32_t rows;                /* target leaves with work */
  const uint32_t *row_leaf;     /* the target leaf of each row */
  const uint32_t *row_off;      /* CSR of each row's boxes */
  const uint32_t *row_box;      /* box slots (coefficients at coeffs[box (p+1)^2 ..]) */
  const uint32_t *box_cell;     /* 4 per box slot: level, ix, iy, iz */
  const uint32_t *leaf_cell, *tgt_start;
  const double *tx, *ty, *tz;
} fmmh_target_inputs;
/* L2P (leaf::l2p; one box per row, the leaf's own) and M2P (leaf::m2p, the W-list boxes),
 * fmm-kernels/src/leaf.rs:1902,1933: phi_hat, grad_hat += the expansions at the row's targets. */
int32_t fmmh_kernel_l2p(int32_t device, int32_t p, int32_t gradients, const fmmh_target_inputs *in,
                        const double *coeffs, double *phi_hat, double *grad_hat, void *stream);
int32_t fmmh_kernel_m2p(int32_t device, int32_t p, int32_t gradients, const fmmh_target_inputs *in,
                        const double *coeffs, double *phi_hat, double *grad_hat, void *stream);
/* movement::zero (movement.rs:482): values[0 .. count) = 0. */
int32_t fmmh_kernel_zero(int32_t device, double *values, int64_t count, void *stream);
/* movement::gather_output (movement.rs:411): phi[order[i]] = phi_hat[i] / (4 pi r), grad[3 order[i]
 * ..] = grad_hat[3 i ..] / (4 pi r^2) for the targets of each of the nleaves leaves, r the
 * leaf's radius in the caller's units (leaf_r[leaf]). */
int32_t fmmh_kernel_gather_output(int32_t device, uint32_t nleaves, const uint32_t *tgt_start, const double *leaf_r,
                                  const uint32_t *order, const double *phi_hat, const double *grad_hat,
                                  int32_t gradients, double *phi, double *grad, void *stream);

/* ---- Helmholtz (kernel e^{ikr} / (4 pi r), complex charges) ---------------------- */
/* The reference's Helmholtz FMM (nd_fmm_exec::helmholtz) runs on the host only
 * (fmm-exec/src/helmholtz/fmm.rs:71); these entry points are its device path over the same
 * plan, keeping its stored forms (docs/CONVENTIONS.md:1435-1700) and its evaluator's order.
 * Complex values are interleaved (re, im): charges[2 ns], phi[2 nt], grad[6 nt] (x re,
 * x im, y re, y im, z re, z im).  The wavenumber is in the units of the points. */
typedef struct fmmh_helmholtz fmmh_helmholtz;

typedef struct {
  int32_t p;         /* expansion degree of every level, 0..15 */
  double k_re;       /* wavenumber k: k != 0, Re k >= 0, Im k >= 0 (Wavenumber::new, */
  double k_im;       /*   fmm-exec/src/helmholtz/mod.rs) */
  int32_t precision; /* fmmh_precision: plain f64, or f32 P2P and bf16x3 M2L */
  int32_t gradients;
  int32_t device;
  int32_t m2l;       /* fmmh_m2l_method (HelmholtzM2l, fmm-exec/src/helmholtz/operator.rs:448-479):
                        0 dense: one complex table per offset (MFMA contraction) on the levels
                        whose tables fit m2l_dense_budget, from the finest level up
                        (DenseM2l::select), per-pair rotation on the others; 1 rotation:
                        per pair rotation, coaxial shift, rotation back in f64 (RCR) on every
                        level, tables O(p^3) per distance instead of O(p^4) per offset; 2 plane
                        wave: the exponential M2L by merge and shift in f64 on the levels the
                        reference's rule gives a quadrature (iprec 2 to degree 15, levels >= 2
                        whose kappa = k w lies in FMM3D's tables), RCR on the others */
  int64_t m2l_dense_budget; /* bytes of dense tables, counted as 316 x 2 (p+1)^4 f64 per level: 0
                               the reference's default (1 GiB), negative no limit */
} fmmh_helmholtz_options;

/* The degree rule at tolerance eps for boxes of half-width radius (fmm-exec/src/helmholtz/
 * degree.rs:8-15): max(ceil(0.30 L^2 + 0.25 L + 2.60), ceil(x + 0.80 (ln 1/eps)^(2/3) x^(1/3))),
 * x = Re(k) sqrt(3) radius, L = log10(1/eps). */
int32_t fmmh_helmholtz_degree(double k_re, double radius, double eps);
/* HelmholtzFmm built over a plan (fmm-exec/src/helmholtz/fmm.rs; operator.rs:448-480):
 * per-level translation tables (dense complex, rotation-coaxial-rotation) uploaded, every
 * buffer allocated.  A plan of one domain (not a forest). */
int32_t fmmh_helmholtz_create(const fmmh_plan *plan, const fmmh_helmholtz_options *options, fmmh_helmholtz **out);
void fmmh_helmholtz_free(fmmh_helmholtz *op);
int32_t fmmh_helmholtz_stats(const fmmh_helmholtz *op, fmmh_stats *out);
/* count: the number of complex charges (the plan's source count). */
int32_t fmmh_helmholtz_load_charges(fmmh_helmholtz *op, const double *charges, int64_t count, int32_t on_device);
int32_t fmmh_helmholtz_reset(fmmh_helmholtz *op);
/* P2M, M2M, M2L, P2L, L2L, L2P, M2P, P2P over every level (the order of
 * fmm-exec/src/helmholtz/fmm.rs's evaluation). */
int32_t fmmh_helmholtz_evaluate(fmmh_helmholtz *op);
/* count: the number of targets. */
int32_t fmmh_helmholtz_read_output(fmmh_helmholtz *op, double *phi, double *grad, int64_t count,
                                   int32_t on_device);
/* As fmmh_stream_wait, for a Helmholtz operator. */
int32_t fmmh_helmholtz_stream_wait(fmmh_helmholtz *op, void *stream);
/* Stored multipoles or locals of one level: per box the real plane then the imaginary
 * plane of (p+1)^2 values. */
int32_t fmmh_helmholtz_read_expansions(fmmh_helmholtz *op, int32_t locals, int32_t level, double *out);
int32_t fmmh_helmholtz_kind_timing(fmmh_helmholtz *op, int32_t enable);
int32_t fmmh_helmholtz_kind_times(fmmh_helmholtz *op, double *ms, int64_t *launches);
/* One translation table of a level whose wavenumber in its own units is kh (host):
 * kind 0 M2M from the child of octant `index` (kh of the parent level), 1 L2L into that
 * child, 2 M2L of offset index `index`, 3 the same M2L composed from the pieces of the
 * per-pair rotation path (rotation blocks, scaled coaxial blocks, rotation back; the
 * RCR of fmm-exec/src/helmholtz/operator.rs:448-461), for tests; out[2 (p+1)^4] real
 * plane then imaginary plane, column-major (the operators of
 * fmm-tables/src/helmholtz/dense.rs). */
int32_t fmmh_helmholtz_table(int32_t p, double kh_re, double kh_im, int32_t kind, int32_t index, double *out);
/* Compensated f64 direct sum of q e^{ikr}/r and its gradient (ik r - 1) e^{ikr} d / r^3
 * (no 1/(4 pi)), nd_fmm_ref::helmholtz::p2p::direct_sum (fmm-ref/src/helmholtz/p2p.rs:155);
 * a source equal to the target skipped.  Host or device pointers; synchronous. */
int32_t fmmh_helmholtz_direct_sum(int32_t device, double k_re, double k_im, const double *sources,
                                  const double *charges, int64_t nsources, const double *targets, int64_t ntargets,
                                  double *phi, double *grad, int32_t on_device);

/* The Helmholtz exponential (plane-wave) M2L on the host (fmm-tables/src/helmholtz/
 * plane_wave/{quadrature,pair,merge}.rs).  Quadrature: FMM3D's iprec-2 table of the domain
 * of kappa = k w (quadrature.rs Quadrature::new), its domain (0 outside the tables), node
 * count, stated error, per node lambda, t and the weight (complex, interleaved re/im), the
 * alpha and Fourier counts (at most `capacity` nodes). */
int32_t fmmh_helmholtz_plane_wave_quadrature(double kappa_re, double kappa_im, int32_t capacity, int32_t *domain,
                                             int32_t *nodes, double *stated_error, double *lambda, double *t,
                                             double *weight, int32_t *alpha, int32_t *fourier);
/* |J_0(z)| .. |J_top(z)| of complex z (quadrature.rs bessel_j_norms): top + 1 values. */
int32_t fmmh_bessel_j_norms(double z_re, double z_im, int32_t top, double *out);
/* One V pair of a level with kh = k r_l (pair.rs PlaneWaveLevel::m2l): out (two planes) +=
 * the exponential M2L of the stored multipole in (two planes) along `direction` (0..5, -1:
 * the offset's own), the level's quadrature at iprec 2. */
int32_t fmmh_helmholtz_plane_wave_m2l(int32_t p, double kh_re, double kh_im, int32_t offset, int32_t direction,
                                      const double *in, double *out);
/* One parent's merge and shift (merge.rs): 216 slot multipoles of two planes each, which
 * are held, and out (8 x two planes) += the M2L of every pair slot -> child. */
int32_t fmmh_helmholtz_plane_wave_merge(int32_t p, double kh_re, double kh_im, const double *slot_multipoles,
                                        const uint8_t *present, double *out);

#ifdef __cplusplus
}
#endif
#endif
