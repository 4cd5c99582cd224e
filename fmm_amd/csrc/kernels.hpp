// kernels.hpp -- launch wrappers of the HIP kernels (kernels.hip) for the operator.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fmmh {
namespace dev {

constexpr int kMaxDegree = 15;  // device path: p <= 15, (p+1)^2 <= 256 coefficients
constexpr int kMaxCoef = (kMaxDegree + 1) * (kMaxDegree + 1);
constexpr int kMaxCoefGrad = (kMaxDegree + 2) * (kMaxDegree + 2);
constexpr int COLS = 128;  // output columns (boxes) per contraction task: 8 waves x 16
constexpr int KC = 4;      // MFMA k-steps per LDS chunk of a table
// scratch rows of the per-pair rotation M2L (one or two rows per pair): at most this many
// bytes at once; the pairs run in chunks, each followed by its ordered reduction
constexpr size_t kRotScratchBytes = size_t(512) << 20;

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2: block b
// runs on XCD b % 8.  With a chunk C > 0, xcd_index deals chunks of C consecutive items
// (leaves, boxes, tasks in Morton order) to the XCDs in turn, so neighbouring items share
// one L2 while every XCD still sees the whole range (the items' costs vary along it).
// C = 0 keeps block b on item b.  Launch xcd_grid(n) blocks; an index >= n is idle.
#ifndef FMMH_XCD_CHUNK
#define FMMH_XCD_CHUNK 0
#endif
constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdChunk = FMMH_XCD_CHUNK;
inline uint32_t xcd_grid(uint32_t n) {
  return kXcdChunk == 0 ? n : (n + kXcds * kXcdChunk - 1) / (kXcds * kXcdChunk) * (kXcds * kXcdChunk);
}
__device__ __forceinline__ uint32_t xcd_index(uint32_t n) {
  (void)n;
  if (kXcdChunk == 0) return blockIdx.x;
  const uint32_t x = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
  return ((k / kXcdChunk) * kXcds + x) * kXcdChunk + k % kXcdChunk;
}
// The same with a chunk chosen at launch (0: block b on item b).
inline uint32_t xcd_grid(uint32_t n, uint32_t chunk) {
  return chunk == 0 ? n : (n + kXcds * chunk - 1) / (kXcds * chunk) * (kXcds * chunk);
}
__device__ __forceinline__ uint32_t xcd_index(uint32_t n, uint32_t chunk) {
  (void)n;
  if (chunk == 0) return blockIdx.x;
  const uint32_t x = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
  return ((k / chunk) * kXcds + x) * chunk + k % chunk;
}

hipError_t upload_constants();

void fill(hipStream_t s, double *p, size_t n, double v);
// q64[j] = q[order[j]]; with rec set, also rec[8 j + 3] = f32 of it (the charge slot of
// the mixed-precision P2P source records)
void gather_charges(hipStream_t s, const double *q, const uint32_t *order, size_t n, double *q64, float *rec);
// Fixed-point scale of the mixed P2P's source records: 2^28 per leaf half-width.
constexpr double kP2pFixed = 268435456.0;
// r^2 (integer units) below which a pair's quantised separation would cost more than 1e-7
// of its terms (|d| < 2^23.5 units, 0.044 half-widths): 2^47
constexpr float kP2pClose = 140737488355328.0f;
// Mixed-precision P2P source records, 4 words per source in leaf order: the source's
// own-leaf coordinates u (|u| <= 1) as round(u 2^28) in int32, then the charge's f32
// bits, which gather_charges writes.
void build_source_records(const double *sx, const double *sy, const double *sz, size_t n, int32_t *host_rec);
// Rotation M2L (rotation.hip): blocks of at most rotation_pairs_per_block pairs of one
// offset (blk_offset), pairs blk_first[b] .. blk_first[b + 1] with source boxes pair_src;
// tables per offset (fwd, bwd: rb values) and per distance (coax: cb values); pair q's
// result goes to scratch row q - q_base of S (nc values; blk_first and blk_offset start
// at the launch's first block, its pairs from q_base on), added to the locals by
// reduce_slots.
template <typename T>
int rotation_pairs_per_block(int p, int rb, int cb);
template <typename T>
hipError_t rotation_allow_lds(int p, int rb, int cb);
template <typename T>
void m2l_rotation(hipStream_t s, int p, uint32_t nblocks, uint32_t q_base, const uint32_t *blk_first,
                  const uint16_t *blk_offset, const uint32_t *pair_src, const uint16_t *dist_of, const T *fwd,
                  const T *bwd, const T *coax, int rb, int cb, const T *M, T *S);
// b[i] = f32(a[i]), i < n
void to_f32(hipStream_t s, size_t n, const double *a, float *b);
// Helmholtz rotation M2L (rotation.hip): blocks of at most h_rotation_pairs_per_block pairs
// of one (level, offset); fwd: 316 offsets' rotation blocks (rb each), coax: per (level,
// distance) entry blk_coax[b] the coaxial blocks, real then imaginary plane (2 cb each);
// M: three planes per box (re, im, -im); S: two scratch rows per pair (re, im), pair q at
// rows 2 (q - q_base), 2 (q - q_base) + 1 (blk_first, blk_offset, blk_coax start at the
// launch's first block; its pairs from q_base on).  h_rotation_allow_lds once per
// process before the first launch.
int h_rotation_pairs_per_block(int p, int rb, int cb);
hipError_t h_rotation_allow_lds(int p, int rb, int cb);
void h_m2l_rotation(hipStream_t s, int p, uint32_t nblocks, uint32_t q_base, const uint32_t *blk_first,
                    const uint16_t *blk_offset, const uint32_t *blk_coax, const uint32_t *pair_src, const double *fwd,
                    const double *coax, int rb, int cb, const double *M, double *S);

// Compensated f64 direct sum (unscaled): chunks of sources per block, partials in
// part[nchunks][nt][8], then added per target in chunk order.
int direct_chunks(int64_t ns, int64_t nt);
void direct_sum(hipStream_t s, const double *src, const double *q, int64_t ns, const double *tgt, int64_t nt,
                int gradients, double *part, int nchunks, double *phi, double *grad);
// grad_q = psi, grad_pos = q dpsi + g field (n points, caller order; grad_pos may be null)
void vjp_combine(hipStream_t s, size_t n, const double *q, const double *g, const double *psi, const double *dpsi,
                 const double *field, double *grad_q, double *grad_pos);
void output_pass(hipStream_t s, uint32_t nleaves, const uint32_t *tgt_start, const double *leaf_r,
                 const uint32_t *order, const double *phi_hat, const double *grad_hat, double *phi, double *grad,
                 int gradients);

// P2P over the leaves leaf_first .. leaf_first + count (global leaf indices), computed in
// T (double, or float with fixed-point or double-float separations); f64 coordinates and
// charges.  T = float also reads the source records of build_source_records (4 words per
// source; rec is unused for T = double and may be null).  uniform_min: for T = float,
// chunks of at least this many targets (and at most 128) take the uniform-source form;
// near_close: per near entry 1 where p2p_close_flags found a close pair (T = float; may
// be null for T = double).
template <typename T>
void p2p(hipStream_t s, uint32_t leaf_first, uint32_t count, int gradients, uint32_t uniform_min,
         const uint32_t *near_off, const uint32_t *near_ent, const uint32_t *leaf_cell, const uint32_t *src_start,
         const uint32_t *tgt_start, const double *sx, const double *sy, const double *sz, const double *sq,
         const float *rec, const uint8_t *near_close, const double *tx, const double *ty, const double *tz,
         double *phi, double *grad);
// Flags the near entries (global CSR over all leaves) that hold a pair closer than
// kP2pClose in the mixed P2P's integer units, and every own-leaf entry.
void p2p_close_flags(hipStream_t s, uint32_t nleaves, const uint32_t *near_off, const uint32_t *near_ent,
                     const uint32_t *leaf_cell, const uint32_t *src_start, const uint32_t *tgt_start,
                     const float *rec, const double *tx, const double *ty, const double *tz, uint8_t *flags);

// P2M (irregular = 0, row_ent the leaf of each row, unit frame) or P2L (irregular = 1,
// row_ent the X-list leaves).  rows: output box gid and cell per row, CSR of entries.
void expand(hipStream_t s, int p, int irregular, uint32_t rows, const uint32_t *row_box,
            const uint32_t *row_cell, const uint32_t *row_off, const uint32_t *row_ent,
            const uint32_t *leaf_cell, const uint32_t *src_start, const double *sx, const double *sy,
            const double *sz, const double *sq, double *coeffs);

// L2P (irregular = 0: one entry per row, the leaf's own box, unit frame) or M2P
// (irregular = 1: the W-list boxes with their cells) at the targets of each row's leaf.
void evaluate(hipStream_t s, int p, int irregular, int gradients, uint32_t rows, const uint32_t *row_leaf,
              const uint32_t *row_off, const uint32_t *row_box, const uint32_t *box_cell,
              const uint32_t *leaf_cell, const uint32_t *tgt_start, const double *tx, const double *ty,
              const double *tz, const double *coeffs, double *phi, double *grad);

// Grouped dense contraction on MFMA (contract.hip): for every task (COLS output columns)
// and every group of the task, Y[out[c]] += T_group X[src[g, c]]; f64, or bf16x3 split
// with f32 accumulation per group and f64 across groups.  An output id >= first_slot is
// a scratch slot (written by exactly one task): its row out - first_slot of S is stored,
// not added to (f32 rows for bf16x3: the task's f64 sum rounded to f32).
// width: columns per task, COLS (wide) or 16 (narrow: the waves split the rows)
// xchunk: tasks dealt to the XCDs in runs of xchunk consecutive tasks (0: block b runs
// task b).
void contract_f64(hipStream_t s, int p, int width, uint32_t ntasks, const uint32_t *task_group_off,
                  const uint16_t *group_table, const uint32_t *group_src, const uint32_t *col_out,
                  const double *tables_swizzled, const double *X, double *Y, uint32_t first_slot, double *S,
                  uint32_t xchunk);
// M2M tasks of 16 parent columns whose groups are the child octants' tables: wave w
// takes groups w, w + 8, ..., the eight waves' sums added in a fixed tree
// (contract.hip k_m2m_split; near the root one workgroup per task and M-tile,
// k_m2m_rows).  Y may be X.  form: 0 the per-tile form when fewer (task, M-tile) pairs
// than CUs, else the split form; 1 the per-tile form; 2 the split form.
void m2m_split(hipStream_t s, int p, uint32_t ntasks, const uint32_t *task_group_off, const uint16_t *group_table,
               const uint32_t *group_src, const uint32_t *col_out, const double *tsw, const double *X, double *Y,
               int form = 0);
// stage: 0 or 1 one k-step per pipeline stage (two for the short tables of p <= 7); 2 two
// k-steps per stage in a ring of two buffers where the table allows it (an autotuned shape)
void contract_bf16(hipStream_t s, int p, uint32_t ntasks, const uint32_t *task_group_off,
                   const uint16_t *group_table, const uint32_t *group_src, const uint32_t *col_out,
                   const uint16_t *tables_swizzled, const uint16_t *X16, double *Y, uint32_t first_slot, float *S,
                   uint32_t xchunk, int stage = 0);
// X16: three bf16 planes (hi, mid, lo) of split_expansion_size(p) / 3 values per box
void split_bf16(hipStream_t s, int p, size_t nboxes, const double *X, uint16_t *X16);
// Y[row_box[r]] += sum of the scratch rows row_ent[row_off[r] .. row_off[r+1]) - first_slot
// in order, from S (f64) or, when S32 is not null, S32 (f32)
void reduce_slots(hipStream_t s, int p, uint32_t rows, const uint32_t *row_box, const uint32_t *row_off,
                  const uint32_t *row_ent, uint32_t first_slot, const double *S, const float *S32, double *Y);

// Swizzled table sizes (elements) per table at degree p, and the host-side swizzles.
size_t swizzled_table_size(int p);
size_t swizzled_table_size_bf16(int p);
size_t split_expansion_size(int p);
void swizzle_table(int p, const double *table_colmajor, double *out);
void swizzle_table_bf16(int p, const double *table_colmajor, uint16_t *out);

}  // namespace dev
}  // namespace fmmh
