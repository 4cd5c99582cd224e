// contract.hip -- grouped dense contractions on MFMA for the translations (M2M, L2L,
// M2L): for every task of COLS output columns and every group of the task (one table
// and one source expansion per column), Y[out[c]] += T_group X[src[group, c]].
//
// The tables of CONVENTIONS §3.12 are the A operand, staged through LDS in software-
// pipelined chunks (the next chunk's table slice and source coefficients are fetched into
// registers while the current one is multiplied); the source expansions are the B
// operand, read straight from HBM/L2; each of the 8 waves of a workgroup owns 16 columns
// and all M-tiles of them in registers.  The groups of a task are added in group order,
// so every output adds its translations in the order of its list (offset index for M2L,
// octant for M2M), the accumulation rule of fmm-plan/src/operator.rs:34-48.
//
// Two arithmetic variants:
//  - f64: v_mfma_f64_16x16x4f64, every product and sum in f64;
//  - bf16x3 (mixed precision): table and expansion are each split into three bf16 terms,
//    x = hi + mid + lo (24 significant bits), and the six products above 2^-24
//    (hi hi, hi mid, mid hi, hi lo, mid mid, lo hi) go through v_mfma_f32_16x16x32_bf16
//    with f32 accumulation; the f32 sum of each group (one offset) is added into f64
//    accumulators, so only one offset's worth of terms is ever summed in f32.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>

#include "kernels.hpp"

namespace fmmh {
namespace dev {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
// native vector types (the HIP struct vector types defeat scalar replacement of
// register arrays, which then live in scratch memory)
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// ---- layouts -------------------------------------------------------------------------

static int m_tiles(int p) { return ((p + 1) * (p + 1) + 15) / 16; }
// f64: chunks of KC k-steps of 4; k = chunk KC + kk + (lane / 16) KS, KS = chunks KC
static int k_chunks(int p) { return ((p + 1) * (p + 1) + 4 * KC - 1) / (4 * KC); }
// bf16: k-steps of 32; k = 32 ks + 8 (lane / 16) + e
static int k_steps16(int p) { return ((p + 1) * (p + 1) + 31) / 32; }

size_t swizzled_table_size(int p) { return size_t(k_chunks(p)) * m_tiles(p) * KC * 64; }
size_t swizzled_table_size_bf16(int p) { return size_t(k_steps16(p)) * 3 * m_tiles(p) * 64 * 8; }
size_t split_expansion_size(int p) { return size_t(3) * k_steps16(p) * 32; }

void swizzle_table(int p, const double *a, double *out) {
  const int nc = (p + 1) * (p + 1), MT = m_tiles(p), NCH = k_chunks(p), KS = NCH * KC;
  for (int c = 0; c < NCH; ++c)
    for (int mt = 0; mt < MT; ++mt)
      for (int kk = 0; kk < KC; ++kk)
        for (int lane = 0; lane < 64; ++lane) {
          const int row = mt * 16 + (lane & 15);
          const int k = c * KC + kk + (lane >> 4) * KS;
          const double v = (row < nc && k < nc) ? a[row + size_t(k) * nc] : 0.0;
          out[((size_t(c) * MT + mt) * KC + kk) * 64 + lane] = v;
        }
}

// Round to nearest even, f32 -> bf16 bits (finite inputs).
__host__ __device__ static inline uint16_t bf16_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

__host__ __device__ static inline double bf16_value(uint16_t b) {
  const uint32_t u = uint32_t(b) << 16;
  float f;
  memcpy(&f, &u, 4);
  return double(f);
}

// x = hi + mid + lo, each a bf16 (residuals formed exactly in f64).
__host__ __device__ static inline void split3(double x, uint16_t &hi, uint16_t &mid, uint16_t &lo) {
  hi = bf16_bits(float(x));
  const double r1 = x - bf16_value(hi);
  mid = bf16_bits(float(r1));
  const double r2 = r1 - bf16_value(mid);
  lo = bf16_bits(float(r2));
}

// [ks][plane][mt][lane][e]: row = 16 mt + lane % 16, k = 32 ks + 8 (lane / 16) + e
void swizzle_table_bf16(int p, const double *a, uint16_t *out) {
  const int nc = (p + 1) * (p + 1), MT = m_tiles(p), KS = k_steps16(p);
  for (int ks = 0; ks < KS; ++ks)
    for (int mt = 0; mt < MT; ++mt)
      for (int lane = 0; lane < 64; ++lane)
        for (int e = 0; e < 8; ++e) {
          const int row = mt * 16 + (lane & 15), k = 32 * ks + 8 * (lane >> 4) + e;
          const double v = (row < nc && k < nc) ? a[row + size_t(k) * nc] : 0.0;
          uint16_t s[3];
          split3(v, s[0], s[1], s[2]);
          for (int pl = 0; pl < 3; ++pl)
            out[((((size_t(ks) * 3 + pl) * MT + mt) * 64 + lane) * 8) + e] = s[pl];
        }
}

// ---- f64 contraction ---------------------------------------------------------------------

// 16 bytes per lane global -> LDS without registers (global_load_lds_dwordx4): lane i
// writes lds + 16 i; lds is the same for the whole wave.  The builtin exists only in the
// device compilation.
__device__ __forceinline__ void dma16(const u4 *g, u4 *lds) {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0);
#else
  (void)g;
  (void)lds;
#endif
}

// s_waitcnt with vmcnt = V (expcnt and lgkmcnt left free), gfx9 encoding
template <int V>
__device__ __forceinline__ void wait_vm() {
  static_assert(V >= 0 && V < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((V & 15) | ((V >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Wave w owns the columns of group w / WM (16 each, 8 / WM groups, so a task has
// 128 / WM columns) and the M-tiles mt = w % WM, w % WM + WM, ...: WM = 1 for wide tasks
// (every wave all M-tiles of its 16 columns), WM = 8 for narrow tasks of 16 columns on
// the coarse levels, where the waves split the rows instead.
//
// Pipeline as in the bf16x3 kernel below: table chunks global -> LDS by DMA into a ring
// of 3 buffers two chunks ahead, source coefficients to registers one chunk ahead, one
// barrier per chunk.  X and Y may be one array (L2L reads the parents' locals and adds
// into the children's), so neither is declared restrict.
template <int MT, int WM>
__global__ __launch_bounds__(512) void k_contract_f64(int nc, int nchunks, uint32_t ntasks, uint32_t xchunk,
                                                      const uint32_t *__restrict__ task_group_off,
                                                      const uint16_t *__restrict__ group_table,
                                                      const uint32_t *__restrict__ group_src,
                                                      const uint32_t *__restrict__ col_out,
                                                      const double *__restrict__ tsw, const double *X,
                                                      double *Y, uint32_t first_slot,
                                                      double *__restrict__ S) {
  constexpr int CHUNK2 = MT * KC * 32;  // 16-byte units per chunk
  constexpr int SEGS = CHUNK2 / 64;     // 1 KB segments (MT KC / 2 of them)
  constexpr int SEGW = (SEGS + 7) / 8;  // DMA instructions per wave per chunk
  constexpr int NST = 3;
  constexpr int W = 128 / WM;              // columns per task
  constexpr int MTW = (MT + WM - 1) / WM;  // M-tiles per wave
  static_assert(CHUNK2 % 64 == 0, "whole segments");
  __shared__ u4 As[NST][CHUNK2];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = wave % WM, col = (wave / WM) * 16 + (lane & 15), g = lane >> 4;
  const int KS = nchunks * KC;
  const uint32_t task = xcd_index(ntasks, xchunk);
  if (task >= ntasks) return;  // uniform over the workgroup
  const uint32_t out = col_out[size_t(task) * W + col];
  const uint32_t g0 = task_group_off[task], g1 = task_group_off[task + 1];
  const int nsteps = int(g1 - g0) * nchunks;
  if (nsteps == 0) return;  // uniform over the workgroup
  d4 acc[MTW];
#pragma unroll
  for (int i = 0; i < MTW; ++i) acc[i] = d4{0.0, 0.0, 0.0, 0.0};
  double pb[KC], b[KC];
  uint32_t src_n = 0;
  auto table_of = [&](uint32_t gi) -> uint32_t {
    const uint32_t w = reinterpret_cast<const uint32_t *>(group_table)[gi >> 1];
    return (gi & 1) ? (w >> 16) : (w & 0xFFFFu);
  };
  auto dma = [&](int step, int buf) {
    const uint32_t gi = g0 + uint32_t(step / nchunks);
    const int c = step % nchunks;
    const u4 *gsrc = reinterpret_cast<const u4 *>(tsw + (size_t(table_of(gi)) * nchunks + c) * (CHUNK2 * 2));
#pragma unroll
    for (int i = 0; i < SEGW; ++i) {
      const int seg = min(wave + 8 * i, SEGS - 1);
      dma16(gsrc + seg * 64 + lane, &As[buf][seg * 64]);
    }
  };
  auto fetch_src = [&](int step) { src_n = group_src[size_t(g0 + uint32_t(step / nchunks)) * W + col]; };
  // The loads are unconditional (a column without a source reads row 0, k is clamped;
  // k >= nc meets zero table entries) and the value is zeroed afterwards for a column
  // without a source: every path issues the same loads.
  auto fetch_b = [&](int step) {
    const int c = step % nchunks;
    const size_t s = src_n != 0xFFFFFFFFu ? src_n : 0;
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) {
      const int k = c * KC + kk + g * KS;
      const double v = X[s * nc + min(k, nc - 1)];
      pb[kk] = src_n != 0xFFFFFFFFu ? v : 0.0;
    }
  };
  auto take_b = [&]() {
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) b[kk] = pb[kk];
  };
  const int last = nsteps - 1;
  fetch_src(0);
  fetch_b(0);
  fetch_src(min(1, last));
  dma(0, 0);
  dma(min(1, last), 1);
  for (int step = 0; step < nsteps; ++step) {
    const int buf = step % NST;
    wait_vm<SEGW>();               // chunk s and its B loads landed (only the DMA of s + 1 may not)
    __builtin_amdgcn_s_barrier();  // every wave's part; ring slot s - 1 free
    __builtin_amdgcn_sched_barrier(0);  // nothing that reads this step's data moves above
    take_b();
    fetch_b(min(step + 1, last));
    fetch_src(min(step + 2, last));
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    dma(min(step + NST - 1, last), (step + NST - 1) % NST);
    const double *a = reinterpret_cast<const double *>(As[buf]);
#pragma unroll
    for (int kk = 0; kk < KC; ++kk)
#pragma unroll
      for (int i = 0; i < MTW; ++i) {
        const int mt = wm + i * WM;
        if (mt < MT)
          acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[(mt * KC + kk) * 64 + lane], b[kk], acc[i], 0, 0, 0);
      }
  }
  wait_vm<0>();  // no DMA may land in LDS after the workgroup ends
  if (out != 0xFFFFFFFFu) {
    // a box row adds; a scratch slot (one writer) is stored
    const bool slot = out >= first_slot;
    double *y = slot ? S + size_t(out - first_slot) * nc : Y + size_t(out) * nc;
    // accumulator element r of lane (g, j): row 4 r + g of the M-tile (f64 16x16x4)
#pragma unroll
    for (int i = 0; i < MTW; ++i) {
      const int mt = wm + i * WM;
      if (mt >= MT) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = mt * 16 + 4 * r + g;
        if (row < nc) y[row] = slot ? acc[i][r] : y[row] + acc[i][r];
      }
    }
  }
}

// ---- M2M with the groups split over the waves ------------------------------------------------
// Tasks of 16 parent columns whose groups are the child octants (8, or 16 when each
// octant has a real and an imaginary table): wave w forms the products of groups w,
// w + 8, ... (its tables read straight from L2 in the swizzled layout, the children's
// expansions as B) over all M-tiles, with no LDS staging and no barrier per chunk, so a
// task is 11 chunks long (p = 12) instead of 8 x 11.  The eight waves' sums then meet in
// LDS in a fixed tree, ((w0 + w4) + (w2 + w6)) + ((w1 + w5) + (w3 + w7)), and wave 0
// adds the result to the parents' rows.  X and Y are one array (children read, parents
// written), so neither is declared restrict.
template <int MT>
__global__ __launch_bounds__(512) void k_m2m_split(int nc, int nchunks, uint32_t ntasks,
                                                   const uint32_t *__restrict__ task_group_off,
                                                   const uint16_t *__restrict__ group_table,
                                                   const uint32_t *__restrict__ group_src,
                                                   const uint32_t *__restrict__ col_out,
                                                   const double *__restrict__ tsw, const double *X,
                                                   double *Y) {
  __shared__ d4 red[4][MT][64];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 15, g = lane >> 4;
  const uint32_t task = blockIdx.x;
  if (task >= ntasks) return;  // uniform over the workgroup
  const uint32_t g0 = task_group_off[task], g1 = task_group_off[task + 1];
  const int KS = nchunks * KC;
  const size_t tsize = size_t(nchunks) * MT * KC * 64;  // one swizzled table
  d4 acc[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) acc[i] = d4{0.0, 0.0, 0.0, 0.0};
  for (uint32_t gi = g0 + uint32_t(wave); gi < g1; gi += 8) {
    const uint32_t w2 = reinterpret_cast<const uint32_t *>(group_table)[gi >> 1];
    const uint32_t table = (gi & 1) ? (w2 >> 16) : (w2 & 0xFFFFu);
    const double *a = tsw + size_t(table) * tsize;
    const uint32_t src = group_src[size_t(gi) * 16 + col];
    const bool has = src != 0xFFFFFFFFu;
    const double *x = X + size_t(has ? src : 0) * nc;
    // half chunks (KH k-steps) in two register sets: the loads of half h + 1 are issued
    // before the MFMAs of half h, so their memory latency runs under those MFMAs (the same
    // MFMA order per M-tile as one loop over the chunks)
    constexpr int KH = KC / 2;
    double a0[KH][MT], b0[KH], a1[KH][MT], b1[KH];
    auto load = [&](int h, double (&av)[KH][MT], double (&bv)[KH]) {
      const int c = h / 2, k0 = (h % 2) * KH;
#pragma unroll
      for (int kk = 0; kk < KH; ++kk) {
        const int k = c * KC + k0 + kk + g * KS;
        const double v = x[min(k, nc - 1)];
        bv[kk] = (has && k < nc) ? v : 0.0;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) av[kk][mt] = a[((size_t(c) * MT + mt) * KC + k0 + kk) * 64 + lane];
      }
    };
    auto run = [&](const double (&av)[KH][MT], const double (&bv)[KH]) {
#pragma unroll
      for (int kk = 0; kk < KH; ++kk)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
          acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk][mt], bv[kk], acc[mt], 0, 0, 0);
    };
    const int nh = 2 * nchunks;
    if constexpr (MT <= 14) {
      load(0, a0, b0);
      for (int h = 0; h < nh; h += 2) {
        if (h + 1 < nh) load(h + 1, a1, b1);
        run(a0, b0);
        if (h + 1 < nh) {
          if (h + 2 < nh) load(h + 2, a0, b0);
          run(a1, b1);
        }
      }
    } else {
      // p >= 14: the two register sets would spill; one half at a time
      for (int h = 0; h < nh; ++h) {
        load(h, a0, b0);
        run(a0, b0);
      }
    }
  }
  // fixed-order tree over the waves' partial sums
#pragma unroll
  for (int half = 4; half >= 1; half >>= 1) {
    __syncthreads();
    if (wave >= half && wave < 2 * half)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) red[wave - half][mt][lane] = acc[mt];
    __syncthreads();
    if (wave < half)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) acc[mt] += red[wave][mt][lane];
  }
  if (wave == 0) {
    const uint32_t out = col_out[size_t(task) * 16 + col];
    if (out != 0xFFFFFFFFu) {
      double *y = Y + size_t(out) * nc;
      // accumulator element r of lane (g, j): row 4 r + g of the M-tile (f64 16x16x4)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = mt * 16 + 4 * r + g;
          if (row < nc) y[row] += acc[mt][r];
        }
    }
  }
}

// The same M2M with one workgroup per (task, M-tile): wave w takes group g0 + w (and
// + 8, ... if a task has more), issues every table and source load of that group and
// M-tile at once (nchunks <= MT chunks of KC k-steps, in registers) and then runs the
// MFMA chain, so a task pays the memory latency once per group instead of once per
// chunk.  The per-tile MFMA order and the wave tree are those of k_m2m_split, so the
// sums are the same bits.  Up to MT = 11 (p <= 12: 88 doubles of operands per lane) all
// chunks at once, above in two batches; for the levels near the root, where a few tasks
// would leave most CUs idle.
template <int MT>
__global__ __launch_bounds__(512) void k_m2m_rows(int nc, int nchunks, uint32_t ntasks,
                                                  const uint32_t *__restrict__ task_group_off,
                                                  const uint16_t *__restrict__ group_table,
                                                  const uint32_t *__restrict__ group_src,
                                                  const uint32_t *__restrict__ col_out,
                                                  const double *__restrict__ tsw, const double *X,
                                                  double *Y) {
  __shared__ d4 red[4][64];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 15, g = lane >> 4;
  const uint32_t task = blockIdx.x / MT;
  const int mt = int(blockIdx.x % MT);
  if (task >= ntasks) return;  // uniform over the workgroup
  const uint32_t g0 = task_group_off[task], g1 = task_group_off[task + 1];
  const int KS = nchunks * KC;
  const size_t tsize = size_t(nchunks) * MT * KC * 64;  // one swizzled table
  d4 acc = d4{0.0, 0.0, 0.0, 0.0};
  for (uint32_t gi = g0 + uint32_t(wave); gi < g1; gi += 8) {
    const uint32_t w2 = reinterpret_cast<const uint32_t *>(group_table)[gi >> 1];
    const uint32_t table = (gi & 1) ? (w2 >> 16) : (w2 & 0xFFFFu);
    const double *a = tsw + size_t(table) * tsize;
    const uint32_t src = group_src[size_t(gi) * 16 + col];
    const bool has = src != 0xFFFFFFFFu;
    // the run of k = g KS .. g KS + KS - 1; k past nc is clamped to the row's last
    // element (in bounds) and masked
    const double *xr = X + size_t(has ? src : 0) * nc;
    // the chunks in NB batches of CH (one batch up to MT = 11; two above, whose operands
    // would not fit the registers at once)
    constexpr int NB = MT <= 11 ? 1 : 2, CH = (MT + NB - 1) / NB;
#pragma unroll
    for (int bt = 0; bt < NB; ++bt) {
      double av[CH * KC], bv[CH * KC];
#pragma unroll
      for (int cc = 0; cc < CH; ++cc) {
        const int c = bt * CH + cc;
        if (c < nchunks)
#pragma unroll
          for (int kk = 0; kk < KC; ++kk) {
            const int k = c * KC + kk + g * KS;
            const double v = xr[k < nc ? k : nc - 1];
            bv[cc * KC + kk] = (has && k < nc) ? v : 0.0;
            av[cc * KC + kk] = a[((size_t(c) * MT + mt) * KC + kk) * 64 + lane];
          }
      }
#pragma unroll
      for (int cc = 0; cc < CH; ++cc)
        if (bt * CH + cc < nchunks)
#pragma unroll
          for (int kk = 0; kk < KC; ++kk)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[cc * KC + kk], bv[cc * KC + kk], acc, 0, 0, 0);
    }
  }
  // fixed-order tree over the waves' partial sums
#pragma unroll
  for (int half = 4; half >= 1; half >>= 1) {
    __syncthreads();
    if (wave >= half && wave < 2 * half) red[wave - half][lane] = acc;
    __syncthreads();
    if (wave < half) acc += red[wave][lane];
  }
  if (wave == 0) {
    const uint32_t out = col_out[size_t(task) * 16 + col];
    if (out != 0xFFFFFFFFu) {
      double *y = Y + size_t(out) * nc;
      // accumulator element r of lane (g, j): row 4 r + g of the M-tile (f64 16x16x4)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = mt * 16 + 4 * r + g;
        if (row < nc) y[row] += acc[r];
      }
    }
  }
}

// ---- bf16x3 contraction ---------------------------------------------------------------------

__device__ __forceinline__ bf8 as_bf8(u4 v) {
  bf8 r;
  memcpy(&r, &v, 16);
  return r;
}

// Waves tile the task two-dimensionally: wave w takes the M-tiles of half w % 2 (MH of
// them) for the 32 columns of pair w / 2, so each A fragment read from LDS feeds two
// N-tiles (half the LDS traffic of one column tile per wave).
//
// Pipeline: the table slices go global -> LDS by DMA (global_load_lds) into a ring of
// NST buffers, issued NST - 1 stages ahead; the source expansions (B operand) go to
// registers one stage ahead.  A stage is KG consecutive k-steps of one group.  NST = 3,
// KG = 1: one k-step per stage, DMA two stages ahead.  KG = 2 halves the wait-and-barrier
// rounds (the short tables of p <= 7 always; larger tables as an autotuned shape, with
// NST = 2 so two stages of up to 13 M-tiles fit the LDS: DMA one stage ahead, covered by
// a stage of twice the MFMAs).  One barrier per stage: after it every wave's DMA for this
// stage has landed and every wave is done with the buffer the next DMA refills.
template <int MT, int KG, int NST = 3>
__global__ __launch_bounds__(512) void k_contract_bf16(int nc, int ksteps, uint32_t ntasks, uint32_t xchunk,
                                                       const uint32_t *__restrict__ task_group_off,
                                                       const uint16_t *__restrict__ group_table,
                                                       const uint32_t *__restrict__ group_src,
                                                       const uint32_t *__restrict__ col_out, const u4 *__restrict__ tsw,
                                                       const u4 *__restrict__ X16, double *__restrict__ Y,
                                                       uint32_t first_slot, float *__restrict__ S) {
  constexpr int CHUNK = MT * 3 * 64;    // u4 per k-step: [plane][mt][lane]
  constexpr int STAGE = KG * CHUNK;     // u4 per stage (KG consecutive k-steps)
  constexpr int SEGS = STAGE / 64;      // 1 KB segments, one DMA instruction each
  constexpr int SEGW = (SEGS + 7) / 8;  // DMA instructions per wave per stage
  static_assert(NST == 2 || NST == 3, "ring of 2 or 3 stages");
  constexpr int MH = (MT + 1) / 2;
  __shared__ u4 As[NST][STAGE];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int half = wave & 1, pair = wave >> 1, g = lane >> 4;
  const int mt0 = half * MH;
  const int kq = ksteps * 4;  // u4 per plane of a split expansion
  const int spg = ksteps / KG;  // stages per group
  const uint32_t task = xcd_index(ntasks, xchunk);
  if (task >= ntasks) return;  // uniform over the workgroup
  int col[2];
  uint32_t out[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    col[nt] = pair * 32 + nt * 16 + (lane & 15);
    out[nt] = col_out[size_t(task) * COLS + col[nt]];
  }
  const uint32_t g0 = task_group_off[task], g1 = task_group_off[task + 1];
  const int nsteps = int(g1 - g0) * spg;
  if (nsteps == 0) return;
  f4 acc[MH][2];
  double sum[MH][2][4];
#pragma unroll
  for (int i = 0; i < MH; ++i)
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      acc[i][nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) sum[i][nt][r] = 0.0;
    }
  u4 pb[KG][2][3], b[KG][2][3];
  const u4 zero4 = {0u, 0u, 0u, 0u};
  // table slice of stage `step` into ring buffer `buf`: wave w copies the segments
  // w, w + 8, ...; a wave past the last segment copies the last one again (same bytes,
  // same place), so every wave issues SEGW DMA instructions
  // the table index of a group, through a 32-bit load of the pair holding it (uniform:
  // a scalar load, outside the vector-memory counter)
  auto table_of = [&](uint32_t gi) -> uint32_t {
    const uint32_t w = reinterpret_cast<const uint32_t *>(group_table)[gi >> 1];
    return (gi & 1) ? (w >> 16) : (w & 0xFFFFu);
  };
  auto dma = [&](int step, int buf) {
    const uint32_t gi = g0 + uint32_t(step / spg);
    const int ks = (step % spg) * KG;
    const u4 *gsrc = tsw + (size_t(table_of(gi)) * ksteps + ks) * CHUNK;
#pragma unroll
    for (int i = 0; i < SEGW; ++i) {
      const int seg = min(wave + 8 * i, SEGS - 1);
      dma16(gsrc + seg * 64 + lane, &As[buf][seg * 64]);
    }
  };
  // Every lane issues the same loads whether its column has a source or not (the value
  // is zeroed after the load), so the vector-memory counter advances the same way on
  // every path and the waits can count on it.  The source indices of a stage are loaded
  // a stage before its B loads, which are loaded a stage before use.
  uint32_t src_n[2];
  auto fetch_src = [&](int step) {
    const uint32_t gi = g0 + uint32_t(step / spg);
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) src_n[nt] = group_src[size_t(gi) * COLS + col[nt]];
  };
  auto fetch_b = [&](int step) {
    const int ks = (step % spg) * KG;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const uint32_t src = src_n[nt];
      const size_t s = src != 0xFFFFFFFFu ? src : 0;
#pragma unroll
      for (int kg = 0; kg < KG; ++kg)
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
          const u4 v = X16[(s * 3 + pl) * kq + (ks + kg) * 4 + g];
          pb[kg][nt][pl] = src != 0xFFFFFFFFu ? v : zero4;
        }
    }
  };
  // the MFMAs of k-step kg of the stage in ring slot buf, B operand bb
  auto mfma_kstep = [&](int buf, int kg, const u4 (&bb)[2][3]) {
    bf8 bh[2], bm[2], bl[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      bh[nt] = as_bf8(bb[nt][0]);
      bm[nt] = as_bf8(bb[nt][1]);
      bl[nt] = as_bf8(bb[nt][2]);
    }
    // the A fragments of M-tile i + 1 are read from LDS while M-tile i is multiplied
    u4 af[2][3];
    auto read_a = [&](int i, u4 *f) {
      const int mt = min(mt0 + i, MT - 1);
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) f[pl] = As[buf][kg * CHUNK + (pl * MT + mt) * 64 + lane];
    };
    read_a(0, af[0]);
#pragma unroll
    for (int i = 0; i < MH; ++i) {
      const int mt = mt0 + i;
      if (i + 1 < MH) read_a(i + 1, af[(i + 1) & 1]);
      if (mt < MT) {
        const bf8 ah = as_bf8(af[i & 1][0]);
        const bf8 am = as_bf8(af[i & 1][1]);
        const bf8 al = as_bf8(af[i & 1][2]);
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          // smallest terms first
          f4 a4 = acc[i][nt];
          a4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[nt], a4, 0, 0, 0);
          a4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm[nt], a4, 0, 0, 0);
          a4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[nt], a4, 0, 0, 0);
          a4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh[nt], a4, 0, 0, 0);
          a4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm[nt], a4, 0, 0, 0);
          a4 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[nt], a4, 0, 0, 0);
          acc[i][nt] = a4;
        }
      }
    }
  };
  // end of a group (one offset): its f32 sum goes into the f64 accumulators
  auto flush = [&]() {
#pragma unroll
    for (int i = 0; i < MH; ++i)
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
#pragma unroll
        for (int r = 0; r < 4; ++r) sum[i][nt][r] += double(acc[i][nt][r]);
        acc[i][nt] = f4{0.f, 0.f, 0.f, 0.f};
      }
  };
  const int last = nsteps - 1;
  if constexpr (NST == 3) {
    // Issue order in stage s: B loads of s + 1, source indices of s + 2, DMA of s + 2
    // (stages past the end repeat the last one, into a ring slot nobody reads again).  At
    // the top of stage s + 1 only the DMA of s + 2 is newer than the B loads of s + 1 and
    // the DMA of s + 1, so a wait for all but SEGW operations covers both.
    fetch_src(0);
    fetch_b(0);
    fetch_src(min(1, last));
    dma(0, 0);
    dma(min(1, last), 1);
    for (int step = 0; step < nsteps; ++step) {
      const int buf = step % NST;
      wait_vm<SEGW>();               // stage s and its B loads landed (only the DMA of s + 1 may not)
      __builtin_amdgcn_s_barrier();  // every wave's part; ring slot s - 1 free
      __builtin_amdgcn_sched_barrier(0);  // nothing that reads this stage's data moves above
#pragma unroll
      for (int kg = 0; kg < KG; ++kg)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
          for (int pl = 0; pl < 3; ++pl) b[kg][nt][pl] = pb[kg][nt][pl];
      fetch_b(min(step + 1, last));
      fetch_src(min(step + 2, last));
      // keep the issue order: the DMA after the loads (a wait for the B loads must not
      // retire the DMA with them)
      __builtin_amdgcn_sched_barrier(0);
      asm volatile("" ::: "memory");
      dma(min(step + 2, last), (step + 2) % NST);
#pragma unroll
      for (int kg = 0; kg < KG; ++kg) mfma_kstep(buf, kg, b[kg]);
      if ((step + 1) % spg == 0) flush();
    }
  } else {
    // Two k-steps per stage (KG = 2, an even k-step count per group, so a stage never
    // straddles groups), ring of two: the B operand goes one k-step (not one stage) ahead,
    // so one register set is in flight as in the one-step form.  Issue order in stage s
    // (k-steps t0 = 2 s, t1 = 2 s + 1): B of t1, DMA of s + 1, then during t1 the B of
    // t1 + 1 (the next group's source indices, loaded a group ahead, when t1 + 1 starts a
    // group).  Waits: at the top of a stage for everything (the DMA of s and the B of t0);
    // before t1 for all but the newest SEGW operations (the DMA of s + 1).
    static_assert(KG == 2, "the ring of two takes two k-steps per stage");
    uint32_t src_cur[2], src_next[2];
    auto load_src = [&](uint32_t gi, uint32_t (&dst)[2]) {
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) dst[nt] = group_src[size_t(gi) * COLS + col[nt]];
    };
    auto load_b = [&](int t, const uint32_t (&src)[2], u4 (&dst)[2][3]) {
      const int ks = t % ksteps;
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        const size_t s = src[nt] != 0xFFFFFFFFu ? src[nt] : 0;
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
          const u4 v = X16[(s * 3 + pl) * kq + ks * 4 + g];
          dst[nt][pl] = src[nt] != 0xFFFFFFFFu ? v : zero4;
        }
      }
    };
    const int lastk = 2 * nsteps - 1;
    load_src(g0, src_cur);
    load_src(min(g0 + 1, g1 - 1), src_next);
    load_b(0, src_cur, pb[0]);
    dma(0, 0);
    for (int step = 0; step < nsteps; ++step) {
      const int buf = step & 1, t0 = 2 * step, t1 = t0 + 1;
      wait_vm<0>();                  // the DMA of this stage and the B of t0 landed
      __builtin_amdgcn_s_barrier();  // every wave's part; the other ring slot free
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) b[0][nt][pl] = pb[0][nt][pl];
      load_b(t1, src_cur, pb[0]);  // t1 is in t0's group
      __builtin_amdgcn_sched_barrier(0);
      asm volatile("" ::: "memory");
      dma(min(step + 1, last), buf ^ 1);
      mfma_kstep(buf, 0, b[0]);
      wait_vm<SEGW>();  // the B of t1 landed (only the DMA of s + 1 may not)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) b[0][nt][pl] = pb[0][nt][pl];
      const int tn = min(t1 + 1, lastk);
      if (tn % ksteps == 0 && tn != t1) {
        // t1 + 1 opens the next group: its indices were loaded a group ago
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) src_cur[nt] = src_next[nt];
        load_src(min(g0 + uint32_t(tn / ksteps) + 1, g1 - 1), src_next);
      }
      load_b(tn, src_cur, pb[0]);
      __builtin_amdgcn_sched_barrier(0);
      mfma_kstep(buf, 1, b[0]);
      if ((t1 + 1) % ksteps == 0) flush();
    }
  }
  // no DMA may land in LDS after the workgroup ends
  wait_vm<0>();
  // accumulator element r of lane (g, j): row 4 g + r of the M-tile (bf16 16x16x32)
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    if (out[nt] == 0xFFFFFFFFu) continue;
    if (out[nt] >= first_slot) {
      // scratch slot (written by this task only): its f64 sum rounded to f32
      float *s = S + size_t(out[nt] - first_slot) * nc;
#pragma unroll
      for (int i = 0; i < MH; ++i) {
        const int mt = mt0 + i;
        if (mt >= MT) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = mt * 16 + 4 * g + r;
          if (row < nc) s[row] = float(sum[i][nt][r]);
        }
      }
      continue;
    }
    double *y = Y + size_t(out[nt]) * nc;
#pragma unroll
    for (int i = 0; i < MH; ++i) {
      const int mt = mt0 + i;
      if (mt >= MT) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = mt * 16 + 4 * g + r;
        if (row < nc) y[row] += sum[i][nt][r];
      }
    }
  }
}

// Split expansions (f64, nc per box) into three bf16 planes of kp = 32 ksteps values.
__global__ void k_split_bf16(size_t nboxes, int nc, int kp, const double *X, uint16_t *X16) {
  const size_t total = nboxes * size_t(kp);
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < total; i += size_t(gridDim.x) * blockDim.x) {
    const size_t box = i / kp;
    const int k = int(i % kp);
    uint16_t h = 0, m = 0, l = 0;
    if (k < nc) split3(X[box * nc + k], h, m, l);
    X16[(box * 3 + 0) * kp + k] = h;
    X16[(box * 3 + 1) * kp + k] = m;
    X16[(box * 3 + 2) * kp + k] = l;
  }
}

// Ordered reduction of per-pair results: Y[row_box[r]] += sum_e S[row_ent[e]] in row
// order, one workgroup per output box, one thread per coefficient.
template <typename T>
__global__ void k_reduce_slots(int nc, uint32_t rows, const uint32_t *row_box, const uint32_t *row_off,
                               const uint32_t *row_ent, uint32_t first_slot, const T *S, double *Y) {
  const uint32_t r = xcd_index(rows);
  if (r >= rows) return;
  const uint32_t box = row_box[r];
  const uint32_t e0 = row_off[r], e1 = row_off[r + 1];
  for (int i = threadIdx.x; i < nc; i += blockDim.x) {
    double acc = Y[size_t(box) * nc + i];
    // eight rows' loads in flight at a time, added in row order
    uint32_t e = e0;
    for (; e + 8 <= e1; e += 8) {
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = S[size_t(row_ent[e + k] - first_slot) * nc + i];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc += double(v[k]);
    }
    for (; e < e1; ++e) acc += double(S[size_t(row_ent[e] - first_slot) * nc + i]);
    Y[size_t(box) * nc + i] = acc;
  }
}

// ---- host wrappers ---------------------------------------------------------------------------

void contract_f64(hipStream_t s, int p, int width, uint32_t ntasks, const uint32_t *task_group_off,
                  const uint16_t *group_table, const uint32_t *group_src, const uint32_t *col_out, const double *tsw,
                  const double *X, double *Y, uint32_t first_slot, double *S, uint32_t xchunk) {
  if (ntasks == 0) return;
  const int nc = (p + 1) * (p + 1), nch = k_chunks(p);
#define FMMH_CONTRACT(MT)                                                                                        \
  case MT:                                                                                                       \
    if (width == 16)                                                                                             \
      hipLaunchKernelGGL((k_contract_f64<MT, 8>), dim3(xcd_grid(ntasks, xchunk)), dim3(512), 0, s, nc, nch,      \
                         ntasks, xchunk,                                                                         \
                         task_group_off,                                                                         \
                         group_table, group_src, col_out, tsw, X, Y, first_slot, S);                             \
    else if (width == 64)                                                                                        \
      hipLaunchKernelGGL((k_contract_f64<MT, 2>), dim3(xcd_grid(ntasks, xchunk)), dim3(512), 0, s, nc, nch,      \
                         ntasks, xchunk, task_group_off, group_table, group_src, col_out, tsw, X, Y, first_slot, \
                         S);                                                                                     \
    else                                                                                                         \
      hipLaunchKernelGGL((k_contract_f64<MT, 1>), dim3(xcd_grid(ntasks, xchunk)), dim3(512), 0, s, nc, nch,      \
                         ntasks, xchunk,                                                                         \
                         task_group_off,                                                                         \
                         group_table, group_src, col_out, tsw, X, Y, first_slot, S);                             \
    break;
  switch (m_tiles(p)) {
    FMMH_CONTRACT(1)
    FMMH_CONTRACT(2)
    FMMH_CONTRACT(3)
    FMMH_CONTRACT(4)
    FMMH_CONTRACT(5)
    FMMH_CONTRACT(6)
    FMMH_CONTRACT(7)
    FMMH_CONTRACT(8)
    FMMH_CONTRACT(9)
    FMMH_CONTRACT(10)
    FMMH_CONTRACT(11)
    FMMH_CONTRACT(12)
    FMMH_CONTRACT(13)
    FMMH_CONTRACT(14)
    FMMH_CONTRACT(15)
    FMMH_CONTRACT(16)
    default:
      break;
  }
#undef FMMH_CONTRACT
}

void m2m_split(hipStream_t s, int p, uint32_t ntasks, const uint32_t *task_group_off, const uint16_t *group_table,
               const uint32_t *group_src, const uint32_t *col_out, const double *tsw, const double *X, double *Y,
               int form) {
  if (ntasks == 0) return;
  const int nc = (p + 1) * (p + 1), nch = k_chunks(p);
#define FMMH_M2M_ROWS(MT)                                                                                        \
  case MT:                                                                                                       \
    hipLaunchKernelGGL((k_m2m_rows<MT>), dim3(ntasks * MT), dim3(512), 0, s, nc, nch, ntasks, task_group_off,   \
                       group_table, group_src, col_out, tsw, X, Y);                                              \
    return;
  // fewer (task, M-tile) pairs than CUs: one workgroup per pair (each
  // re-reads the source expansions, so a large level keeps the split form: 4M-point
  // tree, level 0: 57 -> 14 us, level 5: 226 us split, 787 us per pair)
  const bool per_tile = form == 1 || (form == 0 && ntasks * uint32_t(m_tiles(p)) < 256);
  if (per_tile) switch (m_tiles(p)) {
    FMMH_M2M_ROWS(1)
    FMMH_M2M_ROWS(2)
    FMMH_M2M_ROWS(3)
    FMMH_M2M_ROWS(4)
    FMMH_M2M_ROWS(5)
    FMMH_M2M_ROWS(6)
    FMMH_M2M_ROWS(7)
    FMMH_M2M_ROWS(8)
    FMMH_M2M_ROWS(9)
    FMMH_M2M_ROWS(10)
    FMMH_M2M_ROWS(11)
    FMMH_M2M_ROWS(12)
    FMMH_M2M_ROWS(13)
    FMMH_M2M_ROWS(14)
    FMMH_M2M_ROWS(15)
    FMMH_M2M_ROWS(16)
    default:
      break;
  }
#undef FMMH_M2M_ROWS
#define FMMH_M2M(MT)                                                                                             \
  case MT:                                                                                                       \
    hipLaunchKernelGGL((k_m2m_split<MT>), dim3(ntasks), dim3(512), 0, s, nc, nch, ntasks, task_group_off,       \
                       group_table, group_src, col_out, tsw, X, Y);                                              \
    break;
  switch (m_tiles(p)) {
    FMMH_M2M(1)
    FMMH_M2M(2)
    FMMH_M2M(3)
    FMMH_M2M(4)
    FMMH_M2M(5)
    FMMH_M2M(6)
    FMMH_M2M(7)
    FMMH_M2M(8)
    FMMH_M2M(9)
    FMMH_M2M(10)
    FMMH_M2M(11)
    FMMH_M2M(12)
    FMMH_M2M(13)
    FMMH_M2M(14)
    FMMH_M2M(15)
    FMMH_M2M(16)
    default:
      break;
  }
#undef FMMH_M2M
}

void contract_bf16(hipStream_t s, int p, uint32_t ntasks, const uint32_t *task_group_off, const uint16_t *group_table,
                   const uint32_t *group_src, const uint32_t *col_out, const uint16_t *tsw16, const uint16_t *X16,
                   double *Y, uint32_t first_slot, float *S, uint32_t xchunk, int stage) {
  if (ntasks == 0) return;
  const int nc = (p + 1) * (p + 1), ks = k_steps16(p);
  const u4 *t4 = reinterpret_cast<const u4 *>(tsw16);
  const u4 *x4 = reinterpret_cast<const u4 *>(X16);
  // two k-steps per stage for the short tables (at most 4 M-tiles and 2 k-steps)
  const bool pair_steps = m_tiles(p) <= 4 && ks == 2;
  // two k-steps per stage in a ring of two (stage = 2): an even k-step count and at most
  // 13 M-tiles (2 x 2 x 13 x 3 KB of LDS)
  if (stage == 2 && ks % 2 == 0 && m_tiles(p) >= 5 && m_tiles(p) <= 13) {
#define FMMH_CONTRACT22(MT)                                                                                        case MT:                                                                                                           hipLaunchKernelGGL((k_contract_bf16<MT, 2, 2>), dim3(xcd_grid(ntasks, xchunk)), dim3(512), 0, s, nc, ks,                           ntasks, xchunk, task_group_off, group_table, group_src, col_out, t4, x4, Y, first_slot, S);     return;
    switch (m_tiles(p)) {
      FMMH_CONTRACT22(5)
      FMMH_CONTRACT22(6)
      FMMH_CONTRACT22(7)
      FMMH_CONTRACT22(8)
      FMMH_CONTRACT22(9)
      FMMH_CONTRACT22(10)
      FMMH_CONTRACT22(11)
      FMMH_CONTRACT22(12)
      FMMH_CONTRACT22(13)
      default:
        break;
    }
#undef FMMH_CONTRACT22
  }
#define FMMH_CONTRACT(MT)                                                                                        \
  case MT:                                                                                                       \
    hipLaunchKernelGGL((k_contract_bf16<MT, 1>), dim3(xcd_grid(ntasks, xchunk)), dim3(512), 0, s, nc, ks, ntasks, \
                       xchunk,                                                                                   \
                       task_group_off, group_table,                                                              \
                       group_src, col_out, t4, x4, Y, first_slot, S);                                            \
    break;
#define FMMH_CONTRACT2(MT)                                                                                       \
  case MT:                                                                                                       \
    if (pair_steps)                                                                                              \
      hipLaunchKernelGGL((k_contract_bf16<MT, 2>), dim3(xcd_grid(ntasks, xchunk)), dim3(512), 0, s, nc, ks,      \
                         ntasks, xchunk, task_group_off, group_table, group_src, col_out, t4, x4, Y, first_slot, S); \
    else                                                                                                         \
      hipLaunchKernelGGL((k_contract_bf16<MT, 1>), dim3(xcd_grid(ntasks, xchunk)), dim3(512), 0, s, nc, ks,      \
                         ntasks, xchunk, task_group_off, group_table, group_src, col_out, t4, x4, Y, first_slot, S); \
    break;
  switch (m_tiles(p)) {
    FMMH_CONTRACT2(1)
    FMMH_CONTRACT2(2)
    FMMH_CONTRACT2(3)
    FMMH_CONTRACT2(4)
    FMMH_CONTRACT(5)
    FMMH_CONTRACT(6)
    FMMH_CONTRACT(7)
    FMMH_CONTRACT(8)
    FMMH_CONTRACT(9)
    FMMH_CONTRACT(10)
    FMMH_CONTRACT(11)
    FMMH_CONTRACT(12)
    FMMH_CONTRACT(13)
    FMMH_CONTRACT(14)
    FMMH_CONTRACT(15)
    FMMH_CONTRACT(16)
    default:
      break;
  }
#undef FMMH_CONTRACT
#undef FMMH_CONTRACT2
}

void split_bf16(hipStream_t s, int p, size_t nboxes, const double *X, uint16_t *X16) {
  if (nboxes == 0) return;
  const int nc = (p + 1) * (p + 1), kp = 32 * k_steps16(p);
  size_t total = nboxes * size_t(kp), grid = (total + 255) / 256;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(k_split_bf16, dim3(unsigned(grid)), dim3(256), 0, s, nboxes, nc, kp, X, X16);
}

void reduce_slots(hipStream_t s, int p, uint32_t rows, const uint32_t *row_box, const uint32_t *row_off,
                  const uint32_t *row_ent, uint32_t first_slot, const double *S, const float *S32, double *Y) {
  if (rows == 0) return;
  const int nc = (p + 1) * (p + 1);
  if (S32)
    hipLaunchKernelGGL(k_reduce_slots<float>, dim3(xcd_grid(rows)), dim3(256), 0, s, nc, rows, row_box, row_off,
                       row_ent, first_slot, S32, Y);
  else
    hipLaunchKernelGGL(k_reduce_slots<double>, dim3(xcd_grid(rows)), dim3(256), 0, s, nc, rows, row_box, row_off,
                       row_ent, first_slot, S, Y);
}

}  // namespace dev
}  // namespace fmmh
