// kernels.hip -- HIP kernels of the FMM evaluation for gfx950 (MI355X).
//
// Conventions (reference docs/CONVENTIONS.md): real storage of expansions, idx(n, m) =
// n^2 + n + m (§3.6); scaled coefficients (§3.7); leaf-scaled point coordinates and
// exact relative frames from integer cells (§3.13); the device P2P rules (§3.13
// "Device kernels").
//
// Layout in HBM:
//  - points in leaf order as structure of arrays x, y, z, q (sources) and x, y, z
//    (targets); the points of a leaf are one contiguous range (src_start / tgt_start);
//  - cells: 4 u32 (level, i, j, k) per leaf and per box;
//  - expansions of every level in one array, box-major: box gid (level base + box index)
//    at gid * nc, with nc = (p+1)^2 coefficients contiguous;
//  - outputs phi^ and g^ in leaf order (f64), turned into phi and grad in caller order by
//    the output pass.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>

#include "kernels.hpp"

namespace fmmh {
namespace dev {

__constant__ double c_inv_reg[kMaxCoefGrad];  // 1 / ((n + m)(n - m)) at idx(n, m), n >= m + 2
__constant__ double c_inv_2m[kMaxDegree + 3];  // 1 / (2m)

hipError_t upload_constants() {
  double inv_reg[kMaxCoefGrad] = {0};
  for (int n = 0; n <= kMaxDegree + 1; ++n)
    for (int m = 0; m <= n; ++m)
      if (n >= m + 2) inv_reg[n * n + n + m] = 1.0 / double((n + m) * (n - m));
  double inv2m[kMaxDegree + 3] = {0};
  for (int m = 1; m <= kMaxDegree + 2; ++m) inv2m[m] = 1.0 / (2.0 * m);
  const hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_inv_reg), inv_reg, sizeof(inv_reg));
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_inv_2m), inv2m, sizeof(inv2m));
}

__device__ __forceinline__ int idx(int n, int m) { return n * n + n + m; }

// ---- frames (CONVENTIONS §3.13 "Relative frames") -----------------------------------
// c^(s|t) = (C_L(s) - C_L(t)) 2^(l_t - L) and r^(s|t) = 2^(l_t - l_s), L = max(l_s, l_t),
// C_L = (2 i + 1) 2^(L - l): integers below 2^17 times powers of two, exact in f32/f64.
template <typename T>
__device__ __forceinline__ void relative_frame(const uint32_t *cs, const uint32_t *ct, T c[3], T &r) {
  const int ls = int(cs[0]), lt = int(ct[0]);
  const int L = ls > lt ? ls : lt;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int64_t a = int64_t(2 * cs[k + 1] + 1) << (L - ls);
    const int64_t b = int64_t(2 * ct[k + 1] + 1) << (L - lt);
    c[k] = T(ldexp(double(a - b), lt - L));
  }
  r = T(ldexp(1.0, lt - ls));
}

template <typename T>
__device__ __forceinline__ T rsqrt_t(T x);

template <>
__device__ __forceinline__ float rsqrt_t<float>(float x) {
  return __builtin_amdgcn_rsqf(x);
}

template <>
__device__ __forceinline__ double rsqrt_t<double>(double x) {
  // v_rsq_f64 estimate refined by two Newton steps y += y (1 - x y^2) / 2
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  return y;
}

// ---- fills and moves ----------------------------------------------------------------

__global__ void k_fill(double *p, size_t n, double v) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void k_gather_charges(const double *q, const uint32_t *order, size_t n, double *q64, float *rec) {
  for (size_t j = blockIdx.x * size_t(blockDim.x) + threadIdx.x; j < n; j += size_t(gridDim.x) * blockDim.x) {
    const double v = q[order[j]];
    q64[j] = v;
    if (rec) rec[4 * j + 3] = float(v);
  }
}

// Output pass (fmm-exec/src/fmm.rs:117-145; CONVENTIONS §3.13 "Output"): per leaf,
// phi = phi^ / (4 pi r_t) and grad = g^ / (4 pi r_t^2) in f64, into caller order; r_t is
// the leaf's radius (its own cloud's in a forest).
__global__ void k_output(uint32_t nleaves, const uint32_t *tgt_start, const double *leaf_r,
                         const uint32_t *order, const double *phi_hat, const double *grad_hat,
                         double *phi, double *grad, int gradients) {
  const uint32_t lf = xcd_index(nleaves);
  if (lf >= nleaves) return;
  const uint32_t t0 = tgt_start[lf], t1 = tgt_start[lf + 1];
  const double rt = leaf_r[lf];
  const double s1 = 4.0 * 3.141592653589793 * rt, s2 = s1 * rt;
  for (uint32_t i = t0 + threadIdx.x; i < t1; i += blockDim.x) {
    const uint32_t c = order[i];
    phi[c] = phi_hat[i] / s1;
    if (gradients) {
      grad[3 * size_t(c) + 0] = grad_hat[3 * size_t(i) + 0] / s2;
      grad[3 * size_t(c) + 1] = grad_hat[3 * size_t(i) + 1] / s2;
      grad[3 * size_t(c) + 2] = grad_hat[3 * size_t(i) + 2] / s2;
    }
  }
}

// ---- direct sum (fmm-ref/src/p2p.rs:147-205, term :370-403, Neumaier :407-431) --------
// Compensated f64 sums per target over all sources, unscaled (sum q / r and its
// gradient -sum q d / r^3, d = x - y); a source equal to the target is skipped.  Thread
// = target; blockIdx.y = a chunk of sources whose compensated partials (sum, comp for phi
// and the three gradient components) go to part[chunk][target]; a second kernel adds the
// chunks in order.

struct Neumaier {
  double s = 0.0, c = 0.0;
  __device__ void add(double t) {
    const double u = s + t;
    c += fabs(s) >= fabs(t) ? (s - u) + t : (t - u) + s;
    s = u;
  }
};

__global__ __launch_bounds__(256) void k_direct_partial(const double *src, const double *q, int64_t ns,
                                                        const double *tgt, int64_t nt, int64_t chunk,
                                                        int gradients, double *part) {
  __shared__ double ys[256][4];
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int64_t s0 = int64_t(blockIdx.y) * chunk, s1 = s0 + chunk < ns ? s0 + chunk : ns;
  double x[3] = {0.0, 0.0, 0.0};
  if (t < nt)
    for (int k = 0; k < 3; ++k) x[k] = tgt[3 * t + k];
  Neumaier p, g[3];
  for (int64_t sb = s0; sb < s1; sb += 256) {
    const int cnt = int((s1 - sb) < 256 ? (s1 - sb) : 256);
    __syncthreads();
    if (int(threadIdx.x) < cnt) {
      const int64_t s = sb + threadIdx.x;
      ys[threadIdx.x][0] = src[3 * s];
      ys[threadIdx.x][1] = src[3 * s + 1];
      ys[threadIdx.x][2] = src[3 * s + 2];
      ys[threadIdx.x][3] = q[s];
    }
    __syncthreads();
    if (t < nt)
      for (int j = 0; j < cnt; ++j) {
        const double y0 = ys[j][0], y1 = ys[j][1], y2 = ys[j][2], qj = ys[j][3];
        if (x[0] == y0 && x[1] == y1 && x[2] == y2) continue;
        const double d0 = x[0] - y0, d1 = x[1] - y1, d2 = x[2] - y2;
        const double r2 = (d0 * d0 + d1 * d1) + d2 * d2;
        const double r = sqrt(r2);
        p.add(qj / r);
        if (gradients) {
          const double r3 = r2 * r;
          g[0].add(-(qj * d0 / r3));
          g[1].add(-(qj * d1 / r3));
          g[2].add(-(qj * d2 / r3));
        }
      }
  }
  if (t < nt) {
    double *o = part + (size_t(blockIdx.y) * nt + t) * 8;
    o[0] = p.s;
    o[1] = p.c;
    for (int k = 0; k < 3; ++k) {
      o[2 + 2 * k] = g[k].s;
      o[3 + 2 * k] = g[k].c;
    }
  }
}

__global__ void k_direct_combine(int64_t nt, int nchunks, int gradients, const double *part, double *phi,
                                 double *grad) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  Neumaier acc[4];
  for (int c = 0; c < nchunks; ++c) {
    const double *o = part + (size_t(c) * nt + t) * 8;
    for (int k = 0; k < 4; ++k) {
      acc[k].add(o[2 * k]);
      acc[k].add(o[2 * k + 1]);
    }
  }
  phi[t] = acc[0].s + acc[0].c;
  if (gradients)
    for (int k = 0; k < 3; ++k) grad[3 * t + k] = acc[1 + k].s + acc[1 + k].c;
}

// Adjoint of a potential-based loss (sources = targets, caller order): with psi and
// grad psi the field of the charges g = dL/dphi, dL/dq_j = psi_j and
// dL/dz_j = q_j grad psi_j + g_j grad phi_j (the second term from moving z_j as a target).
__global__ void k_vjp_combine(size_t n, const double *q, const double *g, const double *psi,
                              const double *dpsi, const double *field, double *grad_q, double *grad_pos) {
  for (size_t j = blockIdx.x * size_t(blockDim.x) + threadIdx.x; j < n; j += size_t(gridDim.x) * blockDim.x) {
    grad_q[j] = psi[j];
    if (grad_pos)
      for (int k = 0; k < 3; ++k) grad_pos[3 * j + k] = q[j] * dpsi[3 * j + k] + g[j] * field[3 * j + k];
  }
}

// ---- P2P (CONVENTIONS §3.13 "Device kernels") ---------------------------------------
// One wave per target leaf, targets in registers; every source leaf of the near row is
// staged through LDS in tiles of 64 points, with
// y^ = fma(r^, u_s, c^) formed once per source in f64 (u_s itself for s = t).  A pair adds
// nothing when r^2 == 0: the select comes before the charge is touched.
//
// T = double: every operation in f64.  T = float (mixed precision): coordinates stay
// f64 in HBM; each is split into hi + lo floats (hi = fl32(u), lo = fl32(u - hi)) and the
// separation is formed as (t_hi - y_hi) + (t_lo - y_lo), which keeps it accurate to about
// 2^-24 of |d| even for close pairs (a plain f32 difference loses 2^-24 of |u| >> |d|);
// r^2, the inverse square root and the sums of a tile (at most 64 sources) are f32; the
// tile sums are added in f64.  Coincident points give d = 0 exactly in both forms.

// Lane mapping (one wave per target leaf, targets processed in chunks of 128): a chunk of
// n <= 64 targets uses np = n lanes per source phase and floor(64 / n) phases, phase f
// taking the sources j = f, f + phases, ... of every tile (tiles of phases x floor(64 /
// phases) sources); a chunk of 65..128 targets puts two targets in each lane.  A tile is
// padded with zero charges far away to a multiple of the phase count, so every lane runs
// the same number of steps; the phases' partial sums are added in phase order through
// LDS, so results are the same on every run.  The next tile's coordinates and charges
// are loaded into registers while the current tile is computed.
//
// Only a target's own leaf can hold a source at r^2 == 0 (equal coordinates give equal
// leaves): that tile selects 0 for r^2 == 0; other tiles clamp r^2 below at the smallest
// normal number, which changes nothing for distinct points (r^2 >= 1e-22 in leaf units
// at the deepest level) and keeps a separation rounded to zero finite.
// Chunks of at least `umin` targets in the mixed path (33 by the static rule): each lane
// holds one target (TWO: two, the second 64 further on) and at every step all lanes take
// the same source, read from its 16-byte record by uniform loads, so no LDS staging and
// no lanes split over phases.  A record holds the source's coordinates in its own leaf's
// frame as 32-bit fixed point (u 2^28, |u| <= 1) and its charge as f32.  Per near entry
// the targets are moved into the source leaf's frame once, x' = c'' + rho u_t in f64
// (c'' = (C_t - C_s) / r_s, rho = r_t / r_s), and rounded to the same fixed point: in a
// 2:1-balanced tree a near leaf (U list, or W/X pairs moved to P2P by the near rule) has
// |x'| <= 1 + 2 rho + 2 <= 7 per axis, so |x' 2^28| < 2^31 and the separation
// d' = x' - u_s is one exact integer subtraction per axis, then one conversion to f32
// (relative rounding 2^-24).  The coordinates' quantisation, 2^-29 source-leaf units,
// bounds the error of a pair at 2^-29 / |d'| relative: below 1e-7 from |d'| = 2^-4.5
// half-widths on; a closer pair (r^2 under 2^47 integer units, rare: one in ~10^6 pairs
// of a uniform leaf) forms d' again from the f64 coordinates, x' - u_s, scaled and
// rounded once, so close pairs keep 2^-24 of |d'| as in the double-float form.
// In integer units r^2 = |d'|^2 2^56: the f32 sums of an entry (at most 128 sources) are
// scaled by rho 2^28 (potential) and rho^2 2^56 (gradient) as they go into the f64 sums,
// since q / |d| = rho q / |d'| and q d / |d|^3 = rho^2 q d' / |d'|^3 in target units.
// Off the own leaf r^2 starts from 1/4 (integer units) in the quantised form and from the
// smallest normal float in the f64 form: distinct points keep a finite r^2.
template <bool GRAD, bool TWO>
__device__ __forceinline__ void p2p_uniform(uint32_t lf, uint32_t tb, uint32_t n, uint32_t e0, uint32_t e1,
                                            const uint32_t *__restrict__ near_ent,
                                            const uint32_t *__restrict__ leaf_cell,
                                            const uint32_t *__restrict__ src_start, const int4 *__restrict__ rec,
                                            const uint8_t *__restrict__ near_close,
                                            const double *__restrict__ sx, const double *__restrict__ sy,
                                            const double *__restrict__ sz, const double *__restrict__ tx,
                                            const double *__restrict__ ty, const double *__restrict__ tz,
                                            double *__restrict__ phi, double *__restrict__ grad) {
  constexpr int A = TWO ? 2 : 1;
  constexpr float CLOSE = kP2pClose;
  constexpr float TINY = 1.17549435e-38f;
  const int lane = threadIdx.x;
  double u[A][3];
  bool ok[A];
#pragma unroll
  for (int a = 0; a < A; ++a) {
    const uint32_t i = tb + uint32_t(lane + 64 * a);
    ok[a] = uint32_t(lane + 64 * a) < n;
    u[a][0] = ok[a] ? tx[i] : 0.0;
    u[a][1] = ok[a] ? ty[i] : 0.0;
    u[a][2] = ok[a] ? tz[i] : 0.0;
  }
  double cp[A], cx[A], cy[A], cz[A];
#pragma unroll
  for (int a = 0; a < A; ++a) cp[a] = cx[a] = cy[a] = cz[a] = 0.0;
  const uint32_t *ct = leaf_cell + 4 * size_t(lf);
  for (uint32_t e = e0; e < e1; ++e) {
    const uint32_t sl = near_ent[e];
    const uint32_t s0 = src_start[sl], s1 = src_start[sl + 1];
    if (s0 == s1) continue;
    double c[3], rho;
    relative_frame<double>(ct, leaf_cell + 4 * size_t(sl), c, rho);
    int xi[A][3];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) xi[a][k] = __double2int_rn(fma(rho, u[a][k], c[k]) * kP2pFixed);
    const bool own = sl == lf;
    // entries flagged at build (a pair of the entry closer than the threshold) and the
    // own leaf test their pairs; the others run the separations unchecked
    const bool close_entry = own || near_close[e] != 0;
    for (uint32_t jb = s0; jb < s1; jb += 128) {
      const uint32_t je = (s1 - jb) < 128u ? s1 : jb + 128;
      float ap[A], ax[A], ay[A], az[A];
#pragma unroll
      for (int a = 0; a < A; ++a) ap[a] = ax[a] = ay[a] = az[a] = 0.0f;
      // the records of the next 4 sources are loaded while the current 4 are computed
      auto run = [&](auto check, auto test_close) {
        // K sources: separations and r^2 of every pair first, then (flagged entries) one
        // wave-uniform test for a close pair among them (the f64 recomputation runs only
        // then, per lane), then the terms, so the pairs' chains interleave
        auto body = [&](const int4 *h, uint32_t js, auto kk) {
          constexpr int K = decltype(kk)::value;
          float dx[K][A], dy[K][A], dz[K][A], r2[K][A];
          float rmin = CLOSE;
#pragma unroll
          for (int k = 0; k < K; ++k)
#pragma unroll
            for (int a = 0; a < A; ++a) {
              dx[k][a] = float(xi[a][0] - h[k].x);
              dy[k][a] = float(xi[a][1] - h[k].y);
              dz[k][a] = float(xi[a][2] - h[k].z);
              r2[k][a] = decltype(check)::value
                             ? dx[k][a] * dx[k][a] + dy[k][a] * dy[k][a] + dz[k][a] * dz[k][a]
                             : fmaf(dz[k][a], dz[k][a], fmaf(dy[k][a], dy[k][a], fmaf(dx[k][a], dx[k][a], 0.25f)));
              rmin = fminf(rmin, r2[k][a]);
            }
          if (decltype(test_close)::value && __any(rmin < CLOSE)) {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
              for (int a = 0; a < A; ++a)
                if (r2[k][a] < CLOSE) {
                  // a close pair: d' = x' - u_s in f64, then scaled and rounded once
                  const uint32_t jj = js + uint32_t(k);
                  dx[k][a] = float((fma(rho, u[a][0], c[0]) - sx[jj]) * kP2pFixed);
                  dy[k][a] = float((fma(rho, u[a][1], c[1]) - sy[jj]) * kP2pFixed);
                  dz[k][a] = float((fma(rho, u[a][2], c[2]) - sz[jj]) * kP2pFixed);
                  r2[k][a] = decltype(check)::value
                                 ? dx[k][a] * dx[k][a] + dy[k][a] * dy[k][a] + dz[k][a] * dz[k][a]
                                 : fmaf(dz[k][a], dz[k][a], fmaf(dy[k][a], dy[k][a], fmaf(dx[k][a], dx[k][a], TINY)));
                }
          }
#pragma unroll
          for (int k = 0; k < K; ++k) {
            const float q = __int_as_float(h[k].w);
#pragma unroll
            for (int a = 0; a < A; ++a) {
              const float inv = decltype(check)::value ? (r2[k][a] == 0.0f ? 0.0f : __builtin_amdgcn_rsqf(r2[k][a]))
                                                       : __builtin_amdgcn_rsqf(r2[k][a]);
              const float qi = q * inv;
              ap[a] += qi;
              if (GRAD) {
                const float qi3 = qi * (inv * inv);
                ax[a] -= qi3 * dx[k][a];
                ay[a] -= qi3 * dy[k][a];
                az[a] -= qi3 * dz[k][a];
              }
            }
          }
        };
        uint32_t j = jb;
        if (je - j >= 4) {
          int4 c[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) c[k] = rec[size_t(j) + k];
          for (; j + 8 <= je; j += 4) {
            int4 nx[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) nx[k] = rec[size_t(j + 4) + k];
            body(c, j, std::integral_constant<int, 4>{});
#pragma unroll
            for (int k = 0; k < 4; ++k) c[k] = nx[k];
          }
          body(c, j, std::integral_constant<int, 4>{});
          j += 4;
        }
        for (; j < je; ++j) {
          const int4 h1 = rec[j];
          body(&h1, j, std::integral_constant<int, 1>{});
        }
      };
      if (own)
        run(std::true_type{}, std::true_type{});
      else if (close_entry)
        run(std::false_type{}, std::true_type{});
      else
        run(std::false_type{}, std::false_type{});
      const double s1f = rho * kP2pFixed, s2f = (rho * rho) * (kP2pFixed * kP2pFixed);
#pragma unroll
      for (int a = 0; a < A; ++a) {
        cp[a] += double(ap[a]) * s1f;
        if (GRAD) {
          cx[a] += double(ax[a]) * s2f;
          cy[a] += double(ay[a]) * s2f;
          cz[a] += double(az[a]) * s2f;
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < A; ++a) {
    const uint32_t i = tb + uint32_t(lane + 64 * a);
    if (ok[a]) {
      phi[i] += cp[a];
      if (GRAD) {
        grad[3 * size_t(i) + 0] += cx[a];
        grad[3 * size_t(i) + 1] += cy[a];
        grad[3 * size_t(i) + 2] += cz[a];
      }
    }
  }
}

template <typename T, bool GRAD>
__global__ __launch_bounds__(64) void k_p2p(uint32_t leaf_first, uint32_t count, uint32_t umin,
                                            const uint32_t *__restrict__ near_off,
                                            const uint32_t *__restrict__ near_ent,
                                            const uint32_t *__restrict__ leaf_cell,
                                            const uint32_t *__restrict__ src_start,
                                            const uint32_t *__restrict__ tgt_start, const double *__restrict__ sx,
                                            const double *__restrict__ sy, const double *__restrict__ sz,
                                            const double *__restrict__ sq, const int4 *__restrict__ rec,
                                            const uint8_t *__restrict__ near_close,
                                            const double *__restrict__ tx, const double *__restrict__ ty,
                                            const double *__restrict__ tz, double *__restrict__ phi,
                                            double *__restrict__ grad) {
  constexpr bool SPLIT = sizeof(T) == 4;
  constexpr int TILE = 64;
  constexpr T TINY = sizeof(T) == 4 ? T(1.17549435e-38f) : T(2.2250738585072014e-308);
  constexpr double FAR = 1024.0;  // padding position, in target-leaf units
  // a source in LDS: (x, y, z, q) and, for the split form, (x_lo, y_lo, z_lo, 0), so a
  // lane reads it with one or two 16-byte loads
  using V4 = typename std::conditional<SPLIT, float4, double4>::type;
  __shared__ V4 ys[TILE];
  __shared__ float4 yl[SPLIT ? TILE : 1];
  __shared__ double fr[TILE][4];       // frame (c, r) of each entry of a chunk of the near row
  __shared__ uint32_t fstart[TILE];    // first source of each entry
  __shared__ uint32_t fcum[TILE + 1];  // exclusive prefix sums of the entries' source counts
  const int lane = threadIdx.x;
  const uint32_t li = xcd_index(count);
  if (li >= count) return;
  const uint32_t lf = leaf_first + li;
  const uint32_t t0 = tgt_start[lf], t1 = tgt_start[lf + 1];
  if (t0 == t1) return;
  const uint32_t *ct = leaf_cell + 4 * size_t(lf);
  const uint32_t e0 = near_off[lf], e1 = near_off[lf + 1];
  for (uint32_t tb = t0; tb < t1; tb += 128) {
    const uint32_t n = (t1 - tb) < 128u ? (t1 - tb) : 128u;
    if (SPLIT && n >= umin) {
      if (n > 64)
        p2p_uniform<GRAD, true>(lf, tb, n, e0, e1, near_ent, leaf_cell, src_start, rec, near_close, sx, sy, sz, tx,
                                ty, tz, phi, grad);
      else
        p2p_uniform<GRAD, false>(lf, tb, n, e0, e1, near_ent, leaf_cell, src_start, rec, near_close, sx, sy, sz, tx,
                                 ty, tz, phi, grad);
      continue;
    }
    // np = n lanes per phase and floor(64 / n) phases (a chunk of 17..21 targets runs 3
    // phases, where the next power of two would allow 2); lanes past np x phases idle
    int np = 64, phases = 1;
    if (n <= 64) {
      np = int(n);
      phases = 64 / np;
    }
    const int tsz = phases * (64 / phases);  // sources per tile: a multiple of the phases
    const bool lane_ok = lane < np * phases;
    const int slot = lane % np, phase = lane_ok ? lane / np : 0;
    const bool two = n > 64;
    T ux[2], uy[2], uz[2], lx[2], ly[2], lz[2], ap[2], ax[2], ay[2], az[2];
    bool ok[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const uint32_t i = tb + slot + a * 64;
      ok[a] = lane_ok && ((a == 0) ? (uint32_t(slot) < n) : (two && uint32_t(slot + 64) < n));
      const double x = ok[a] ? tx[i] : 0.0, y = ok[a] ? ty[i] : 0.0, z = ok[a] ? tz[i] : 0.0;
      ux[a] = T(x);
      uy[a] = T(y);
      uz[a] = T(z);
      lx[a] = T(x - double(ux[a]));
      ly[a] = T(y - double(uy[a]));
      lz[a] = T(z - double(uz[a]));
      ap[a] = ax[a] = ay[a] = az[a] = T(0);
    }
    // f32 sums are per tile (at most 64 terms) and go into f64 after each tile, so long
    // near rows do not lose digits to one growing f32 sum
    double cp[2] = {0.0, 0.0}, cx[2] = {0.0, 0.0}, cy[2] = {0.0, 0.0}, cz[2] = {0.0, 0.0};
    auto flush = [&]() {
      if (SPLIT) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          cp[a] += double(ap[a]);
          cx[a] += double(ax[a]);
          cy[a] += double(ay[a]);
          cz[a] += double(az[a]);
          ap[a] = ax[a] = ay[a] = az[a] = T(0);
        }
      }
    };
    // one source j of the tile against the lane's targets
    auto step = [&](int j, bool check) {
      const V4 v = ys[j];
      const T yx = v.x, yy = v.y, yz = v.z, q = v.w;
      T yxl = T(0), yyl = T(0), yzl = T(0);
      if (SPLIT) {
        const float4 w = yl[j];
        yxl = w.x;
        yyl = w.y;
        yzl = w.z;
      }
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        if (a == 1 && !two) break;
        T dx = ux[a] - yx, dy = uy[a] - yy, dz = uz[a] - yz;
        if (SPLIT) {
          dx += lx[a] - yxl;
          dy += ly[a] - yyl;
          dz += lz[a] - yzl;
        }
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T inv = check ? (r2 == T(0) ? T(0) : rsqrt_t<T>(r2)) : rsqrt_t<T>(r2 > TINY ? r2 : TINY);
        const T qi = q * inv;
        ap[a] += qi;
        if (GRAD) {
          const T qi3 = qi * (inv * inv);
          ax[a] -= qi3 * dx;
          ay[a] -= qi3 * dy;
          az[a] -= qi3 * dz;
        }
      }
    };
    // stage one tile: lane < cnt holds the source (x, y, z, q) in the target frame;
    // lanes cnt .. steps * phases - 1 hold zero charges far away
    auto stage = [&](int cnt, int steps, bool valid, double yx, double yy, double yz, double q) {
      __syncthreads();
      if (lane < steps * phases) {
        if (!(valid && lane < cnt)) {
          yx = yy = yz = FAR;
          q = 0.0;
        }
        V4 v;
        v.x = T(yx);
        v.y = T(yy);
        v.z = T(yz);
        v.w = T(q);
        ys[lane] = v;
        if (SPLIT)
          yl[lane] = make_float4(float(yx - double(float(yx))), float(yy - double(float(yy))),
                                 float(yz - double(float(yz))), 0.0f);
      }
      __syncthreads();
    };
    // The near row is walked in chunks of up to 64 entries.  The own leaf's sources are
    // computed apart (r^2 == 0 select); the other entries' sources form one stream, cut
    // into full tiles of 64 whatever the leaf sizes, each lane staging a source of the
    // entry its stream position falls in (frames and prefix sums in LDS).
    for (uint32_t cb = e0; cb < e1; cb += TILE) {
      const int nce = int((e1 - cb) < uint32_t(TILE) ? (e1 - cb) : uint32_t(TILE));
      uint32_t st = 0, cnt = 0;
      bool own = false;
      double c[3] = {0.0, 0.0, 0.0}, r = 1.0;
      if (lane < nce) {
        const uint32_t sl = near_ent[cb + lane];
        st = src_start[sl];
        cnt = src_start[sl + 1] - st;
        own = sl == lf;
        if (own) cnt = 0;
        relative_frame<double>(leaf_cell + 4 * size_t(sl), ct, c, r);
      }
      uint32_t inc = cnt;  // inclusive prefix sum over the lanes
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(inc, off);
        if (lane >= off) inc += v;
      }
      const uint32_t total = __shfl(inc, 63);
      const bool has_own = __ballot(own) != 0;
      __syncthreads();  // the previous chunk's tiles are done with the entry table
      fr[lane][0] = c[0];
      fr[lane][1] = c[1];
      fr[lane][2] = c[2];
      fr[lane][3] = r;
      fstart[lane] = st;
      fcum[lane + 1] = inc;
      if (lane == 0) fcum[0] = 0;
      if (has_own) {
        // the own leaf: y = u_s; coincident points add nothing
        const uint32_t s0 = src_start[lf], s1 = src_start[lf + 1];
        for (uint32_t sb = s0; sb < s1; sb += uint32_t(tsz)) {
          const int cn = int((s1 - sb) < uint32_t(tsz) ? (s1 - sb) : uint32_t(tsz));
          const int steps = (cn + phases - 1) / phases;
          const uint32_t s = sb + uint32_t(lane);
          const bool valid = s < s1;
          stage(cn, steps, valid, valid ? sx[s] : 0.0, valid ? sy[s] : 0.0, valid ? sz[s] : 0.0,
                valid ? sq[s] : 0.0);
          for (int k2 = 0; k2 < steps; ++k2) step(phase + k2 * phases, true);
          flush();
        }
      } else {
        __syncthreads();  // the entry table is visible to every lane
      }
      // stream position f of lane: entry e with fcum[e] <= f < fcum[e + 1]
      uint32_t ea = 0;  // first entry of the current tile (uniform)
      uint32_t pe = 0;
      bool pvalid = false;
      double px = 0.0, py = 0.0, pz = 0.0, pq = 0.0;
      auto prefetch = [&](uint32_t base) {
        while (fcum[ea + 1] <= base) ++ea;
        const uint32_t f = base + uint32_t(lane);
        pvalid = f < total;
        if (pvalid) {
          uint32_t e = ea;
          while (fcum[e + 1] <= f) ++e;
          pe = e;
          const uint32_t s = fstart[e] + (f - fcum[e]);
          px = sx[s];
          py = sy[s];
          pz = sz[s];
          pq = sq[s];
        }
      };
      if (total > 0) prefetch(0);
      for (uint32_t base = 0; base < total; base += uint32_t(tsz)) {
        const int cn = int((total - base) < uint32_t(tsz) ? (total - base) : uint32_t(tsz));
        const int steps = (cn + phases - 1) / phases;
        double yx = 0.0, yy = 0.0, yz = 0.0;
        if (pvalid) {
          yx = fma(fr[pe][3], px, fr[pe][0]);
          yy = fma(fr[pe][3], py, fr[pe][1]);
          yz = fma(fr[pe][3], pz, fr[pe][2]);
        }
        stage(cn, steps, pvalid, yx, yy, yz, pq);
        // start the next tile's loads before computing this one
        if (base + uint32_t(tsz) < total) prefetch(base + uint32_t(tsz));
        int k2 = 0;
        for (; k2 + 2 <= steps; k2 += 2) {
          step(phase + k2 * phases, false);
          step(phase + (k2 + 1) * phases, false);
        }
        if (k2 < steps) step(phase + k2 * phases, false);
        flush();
      }
    }
    if (!SPLIT) {
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        cp[a] = double(ap[a]);
        cx[a] = double(ax[a]);
        cy[a] = double(ay[a]);
        cz[a] = double(az[a]);
      }
    }
    // add the phases' partial sums (lanes slot + np f share a target) in phase order,
    // through the frame table's LDS (free once the row's tiles are done)
    if (phases > 1) {
      __syncthreads();
      fr[lane][0] = cp[0];
      fr[lane][1] = cx[0];
      fr[lane][2] = cy[0];
      fr[lane][3] = cz[0];
      __syncthreads();
      if (phase == 0)
        for (int f = 1; f < phases; ++f) {
          const int o = f * np + slot;
          cp[0] += fr[o][0];
          if (GRAD) {
            cx[0] += fr[o][1];
            cy[0] += fr[o][2];
            cz[0] += fr[o][3];
          }
        }
      __syncthreads();  // the next chunk writes the frame table again
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const uint32_t i = tb + slot + a * 64;
      if (ok[a] && phase == 0) {
        phi[i] += cp[a];
        if (GRAD) {
          grad[3 * size_t(i) + 0] += cx[a];
          grad[3 * size_t(i) + 1] += cy[a];
          grad[3 * size_t(i) + 2] += cz[a];
        }
      }
    }
  }
}

// ---- expansions from points: P2M (regular, unit frame) and P2L (irregular) ----------
// One wave per output box; the row's source leaves in order, their points in tiles of
// 32.  Per tile, lane j < 32 forms the diagonal X_m^m (m = 0..p) of point j into LDS; then
// lane (column pair g, phase f) walks columns g and p - g (degrees g..p, then p-g..p: p + 2
// entries, one column of p/2 + 1 entries for g = p/2) of the points j = f, f + F, ... of
// the tile (F = 64 / ceil((p+1)/2) phases), adding q conj X_n^m to its entries in
// registers, so every lane takes the same number of steps per point.  The phases' columns
// meet in LDS (over the diagonals' space, as many phases at a time as fit) and are added
// in phase order.  About 10 KB of LDS per wave, so 16 waves share a CU.

template <bool IRREG>
__global__ __launch_bounds__(64) void k_expand(int p, uint32_t rows, const uint32_t *row_box, const uint32_t *row_cell,
                                               const uint32_t *row_off, const uint32_t *row_ent,
                                               const uint32_t *leaf_cell, const uint32_t *src_start,
                                               const double *sx, const double *sy, const double *sz,
                                               const double *sq, double *coeffs) {
  constexpr int TP = 32;                       // points per tile
  constexpr int LD = TP + 1;                   // row stride of a diagonal plane (no bank conflicts)
  constexpr int DW = (kMaxDegree + 1) * LD;    // doubles per diagonal plane
  constexpr int NE = kMaxDegree + 2;           // register entries per lane (two columns)
  __shared__ double pt[3][TP];                 // z, q, r2 (regular) or 1/r2 (irregular)
  __shared__ double diag[2 * DW];              // Re, Im of X_m^m at [m][j]; then the phase sums
  __shared__ double inv_tab[kMaxDegree + 1][kMaxDegree + 1];  // 1 / ((n + m)(n - m)) at [n - m][m]
  double *dre = diag, *dim = diag + DW, *red = diag;
  const int lane = threadIdx.x;
  const int P1 = p + 1, nc = P1 * P1;
  const int groups = (P1 + 1) / 2;
  const int phases = 64 / groups;
  const int g = lane % groups, ph = lane / groups;
  const bool active = ph < phases;
  // this lane's columns: m1 = g (entries 0 .. len1 - 1), then m2 = p - g (entries
  // len1 .. tot - 1) unless it is the same column
  const int m1 = g, m2 = p - g;
  const int len1 = P1 - m1, tot = m2 != m1 ? p + 2 : len1;
  const uint32_t row = xcd_index(rows);
  if (row >= rows) return;
  const uint32_t box = row_box[row];
  if (!IRREG)
    for (int t = lane; t < (kMaxDegree + 1) * (kMaxDegree + 1); t += 64) {
      const int kk = t / (kMaxDegree + 1), mm = t % (kMaxDegree + 1);
      inv_tab[kk][mm] = kk >= 1 ? 1.0 / double((2 * mm + kk) * kk) : 0.0;
    }
  double are[NE], aim[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) are[k] = aim[k] = 0.0;
  for (uint32_t e = row_off[row]; e < row_off[row + 1]; ++e) {
    const uint32_t sl = row_ent[e];
    const uint32_t s0 = src_start[sl], s1 = src_start[sl + 1];
    double c[3] = {0.0, 0.0, 0.0}, r = 1.0;
    // P2L from leaf s into box t: frame (c^(t|s), r^(t|s)), v = (u_s - c) / r
    if (IRREG) relative_frame<double>(row_cell + 4 * size_t(row), leaf_cell + 4 * size_t(sl), c, r);
    const double inv_r = 1.0 / r;
    for (uint32_t sb = s0; sb < s1; sb += TP) {
      const int cnt = int((s1 - sb) < uint32_t(TP) ? (s1 - sb) : uint32_t(TP));
      __syncthreads();
      if (lane < cnt) {
        const uint32_t s = sb + lane;
        double x = sx[s], y = sy[s], z = sz[s];
        if (IRREG) {
          x = (x - c[0]) * inv_r;
          y = (y - c[1]) * inv_r;
          z = (z - c[2]) * inv_r;
        }
        const double r2 = x * x + y * y + z * z;
        double dr, di = 0.0, f2 = 0.0;
        if (IRREG) {
          f2 = 1.0 / r2;
          dr = rsqrt_t<double>(r2);
        } else {
          dr = 1.0;
        }
        dre[lane] = dr;
        dim[lane] = 0.0;
        for (int k = 1; k <= p; ++k) {
          const double f = IRREG ? (2.0 * k - 1.0) * f2 : c_inv_2m[k];
          const double nr = f * (x * dr - y * di), ni = f * (x * di + y * dr);
          dr = nr;
          di = ni;
          dre[k * LD + lane] = dr;
          dim[k * LD + lane] = di;
        }
        pt[0][lane] = z;
        pt[1][lane] = sq[s];
        pt[2][lane] = IRREG ? f2 : r2;
      }
      __syncthreads();
      if (active)
        for (int j = ph; j < cnt; j += phases) {
          const double z = pt[0][j], q = pt[1][j], w = pt[2][j];
          const double d2r = dre[m2 * LD + j], d2i = dim[m2 * LD + j];
          double pr = 0.0, pi = 0.0, cr = dre[m1 * LD + j], ci = dim[m1 * LD + j];
          int mm = m1, k = 0;  // current column and degree above its diagonal
#pragma unroll
          for (int kk = 0; kk < NE; ++kk) {
            if (kk >= tot) continue;
            if (kk > 0) {
              // X_n^m from X_{n-1}^m and X_{n-2}^m (n = mm + k + 1); at the column change
              // the second column's diagonal instead
              ++k;
              const int n = mm + k;
              const double a = (2.0 * n - 1.0) * z;
              double nr, ni;
              if (IRREG) {
                const double b = double((n - 1) * (n - 1) - mm * mm);
                nr = (a * cr - b * pr) * w;
                ni = (a * ci - b * pi) * w;
              } else {
                const double inv = inv_tab[k][mm];
                nr = (a * cr - w * pr) * inv;
                ni = (a * ci - w * pi) * inv;
              }
              const bool restart = kk == len1;
              pr = restart ? 0.0 : cr;
              pi = restart ? 0.0 : ci;
              cr = restart ? d2r : nr;
              ci = restart ? d2i : ni;
              if (restart) {
                mm = m2;
                k = 0;
              }
            }
            are[kk] += q * cr;
            aim[kk] -= q * ci;
          }
        }
    }
  }
  // phase sums, as many phases at a time as the diagonals' space holds, in phase order
  const int chunk = (2 * DW) / nc;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};  // slots lane, lane + 64, ... (nc <= 256)
  for (int f0 = 0; f0 < phases; f0 += chunk) {
    __syncthreads();
    if (active && ph >= f0 && ph < f0 + chunk)
#pragma unroll
      for (int kk = 0; kk < NE; ++kk) {
        const int mm = kk < len1 ? m1 : m2, n = kk < len1 ? m1 + kk : m2 + (kk - len1);
        if (kk < tot) red[(ph - f0) * nc + idx(n, mm)] = are[kk];
        if (kk < tot && mm > 0) red[(ph - f0) * nc + idx(n, -mm)] = aim[kk];
      }
    __syncthreads();
    const int f1 = f0 + chunk < phases ? f0 + chunk : phases;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int slot = lane + 64 * u;
      if (slot < nc)
        for (int f = f0; f < f1; ++f) acc[u] += red[(f - f0) * nc + slot];
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int slot = lane + 64 * u;
    if (slot < nc) coeffs[size_t(box) * nc + slot] += acc[u];
  }
}

// ---- evaluation at targets: L2P (regular) and M2P (irregular) ------------------------
// One wave per target leaf; every entry's coefficients go to LDS with the derived
// coefficients of the gradient (the ladders of CONVENTIONS §3.11 "L2P and M2P" moved
// onto the coefficients), then the lanes of a target (one, or up to 8 for leaves of at
// most 32 targets) walk the recursion of its columns and contract on the fly.  Regular: grad_k = sum_{n<p} G_k R_n, with
// Gz_n^m = L_{n+1}^m, Gx_n^m = (L_{n+1}^{m+1} - L_{n+1}^{m-1})/2,
// Gy_n^m = i (L_{n+1}^{m+1} + L_{n+1}^{m-1})/2.  Irregular: grad_k = sum_{1<=n<=p+1}
// G_k I_n, with Gz_n^m = -M_{n-1}^m, Gx_n^m = (M_{n-1}^{m+1} - M_{n-1}^{m-1})/2,
// Gy_n^m = i (M_{n-1}^{m+1} + M_{n-1}^{m-1})/2.

__device__ __forceinline__ void read_order(const double *c, int n, int q, int pmax, double &a, double &b) {
  if (n < 0 || n > pmax || q > n || -q > n) {
    a = b = 0.0;
    return;
  }
  if (q >= 0) {
    a = c[idx(n, q)];
    b = q > 0 ? c[idx(n, -q)] : 0.0;
  } else {
    const double s = ((-q) & 1) ? -1.0 : 1.0;
    a = s * c[idx(n, -q)];
    b = -s * c[idx(n, q)];
  }
}

template <bool IRREG, bool GRAD>
__global__ __launch_bounds__(IRREG ? 128 : 64) void k_evaluate(int p, uint32_t rows, const uint32_t *row_leaf,
                                                           const uint32_t *row_off,
                                                           const uint32_t *row_box, const uint32_t *box_cell,
                                                           const uint32_t *leaf_cell, const uint32_t *tgt_start,
                                                           const double *tx, const double *ty, const double *tz,
                                                           const double *coeffs, double *phi, double *grad) {
  // LDS sized to the degree (eval_lds_bytes): the coefficients, then the three gradient
  // coefficient sets, so that more waves share a CU at low p
  extern __shared__ __align__(16) double eval_smem[];
  const int nc = (p + 1) * (p + 1);
  const uint32_t row = xcd_index(rows);
  if (row >= rows) return;
  const uint32_t lf = row_leaf[row];
  const uint32_t t0 = tgt_start[lf], t1 = tgt_start[lf + 1];
  if (t0 == t1) return;
  const int pg = IRREG ? p + 1 : p - 1;  // degree of the gradient coefficients
  const int ng = pg >= 0 ? (pg + 1) * (pg + 1) : 0;
  double *C = eval_smem, *G0 = C + nc, *G1 = G0 + ng, *G2 = G1 + ng;
  // L2P: one wave per leaf; leaves of more than 32 targets one lane per target, in
  // passes of 64, smaller ones L = min(8, 64 / n) sub-lanes per target, which share its
  // columns.  M2P: two waves, one thread per target (a row's W-list boxes are taken in
  // turn, so the per-box set-up and passes stay short on long rows).
  constexpr int NT = IRREG ? 128 : 64;
  const int lane = threadIdx.x;
  const uint32_t nt = t1 - t0;
  const int L = (IRREG || nt > 32) ? 1 : (64 / int(nt) < 8 ? 64 / int(nt) : 8);
  const int per = NT / L;  // targets per pass
  const int u = lane % L, tl = lane / L;
  for (uint32_t e = row_off[row]; e < row_off[row + 1]; ++e) {
    const uint32_t box = row_box[e];
    double c[3] = {0.0, 0.0, 0.0}, r = 1.0;
    if (IRREG) relative_frame<double>(box_cell + 4 * size_t(box), leaf_cell + 4 * size_t(lf), c, r);
    __syncthreads();
    for (int k = lane; k < nc; k += NT) C[k] = coeffs[size_t(box) * nc + k];
    __syncthreads();
    if (GRAD)
      for (int k = lane; k < ng; k += NT) {
        int n = 0;
        while ((n + 1) * (n + 1) <= k) ++n;
        const int m = k - n * n - n;
        if (m < 0) continue;  // slot -m is written with slot +m
        const int src = IRREG ? n - 1 : n + 1;
        double ar, ai, br, bi, zr, zi;
        read_order(C, src, m + 1, p, ar, ai);
        read_order(C, src, m - 1, p, br, bi);
        read_order(C, src, m, p, zr, zi);
        const double xr = 0.5 * (ar - br), xi = 0.5 * (ai - bi);
        const double yr = -0.5 * (ai + bi), yi = 0.5 * (ar + br);
        if (IRREG) {
          zr = -zr;
          zi = -zi;
        }
        G0[idx(n, m)] = xr;
        G1[idx(n, m)] = yr;
        G2[idx(n, m)] = zr;
        if (m > 0) {
          G0[idx(n, -m)] = xi;
          G1[idx(n, -m)] = yi;
          G2[idx(n, -m)] = zi;
        }
      }
    __syncthreads();
    const double inv_r = 1.0 / r;
    for (uint32_t b = 0; b < nt; b += uint32_t(per)) {
      const uint32_t ti = b + uint32_t(tl);
      const bool valid = tl < per && ti < nt;
      const uint32_t i = t0 + (valid ? ti : 0u);
      double sp = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
      if (valid) {
        double x = tx[i], y = ty[i], z = tz[i];
        if (IRREG) {
          x = (x - c[0]) * inv_r;
          y = (y - c[1]) * inv_r;
          z = (z - c[2]) * inv_r;
        }
        const double r2 = x * x + y * y + z * z;
        const double inv_r2 = IRREG ? 1.0 / r2 : 0.0;
        double dr = IRREG ? rsqrt_t<double>(r2) : 1.0, di = 0.0;
        int md = 0;  // order of the diagonal entry (dr, di)
        const int ptop = IRREG ? (GRAD ? p + 1 : p) : p;
        // columns m of this sub-lane: block jb holds m = jb L .. jb L + L - 1, dealt to
        // the sub-lanes forwards in even blocks and backwards in odd ones
        for (int jb = 0; jb * L <= ptop; ++jb) {
          const int m = jb * L + ((jb & 1) ? L - 1 - u : u);
          if (m > ptop) continue;
          while (md < m) {
            ++md;
            const double f = IRREG ? (2.0 * md - 1.0) * inv_r2 : c_inv_2m[md];
            const double nr = f * (x * dr - y * di), ni = f * (x * di + y * dr);
            dr = nr;
            di = ni;
          }
          double pr = 0.0, pi = 0.0, cr = dr, ci = di;
          const double w = m == 0 ? 1.0 : 2.0;
          for (int n = m; n <= ptop; ++n) {
            if (n > m) {
              double nr, ni;
              if (IRREG) {
                if (n == m + 1) {
                  const double f = (2.0 * m + 1.0) * z * inv_r2;
                  nr = f * cr;
                  ni = f * ci;
                } else {
                  const double a = (2.0 * n - 1.0) * z, bb = double((n - 1) * (n - 1) - m * m);
                  nr = (a * cr - bb * pr) * inv_r2;
                  ni = (a * ci - bb * pi) * inv_r2;
                }
              } else {
                if (n == m + 1) {
                  nr = z * cr;
                  ni = z * ci;
                } else {
                  const double a = (2.0 * n - 1.0) * z, inv = c_inv_reg[idx(n, m)];
                  nr = (a * cr - r2 * pr) * inv;
                  ni = (a * ci - r2 * pi) * inv;
                }
              }
              pr = cr;
              pi = ci;
              cr = nr;
              ci = ni;
            }
            if (n <= p) {
              const double lr = C[idx(n, m)], li = m > 0 ? C[idx(n, -m)] : 0.0;
              sp += w * (lr * cr - li * ci);
            }
            if (GRAD && n <= pg && (!IRREG || n >= 1)) {
              const int ia = idx(n, m), ib = idx(n, -m);
              const double xr = G0[ia], yr = G1[ia], zr = G2[ia];
              const double xi = m > 0 ? G0[ib] : 0.0, yi = m > 0 ? G1[ib] : 0.0, zi = m > 0 ? G2[ib] : 0.0;
              sx += w * (xr * cr - xi * ci);
              sy += w * (yr * cr - yi * ci);
              sz += w * (zr * cr - zi * ci);
            }
          }
        }
      }
      if (L > 1) {
        // the target's sub-lanes add their sums in sub-lane order (every lane takes part)
        const int base = lane - u;
        double tp = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
        for (int k = 0; k < L; ++k) {
          const int src = base + k < 64 ? base + k : 63;
          tp += __shfl(sp, src);
          if (GRAD) {
            ux += __shfl(sx, src);
            uy += __shfl(sy, src);
            uz += __shfl(sz, src);
          }
        }
        sp = tp;
        sx = ux;
        sy = uy;
        sz = uz;
      }
      if (valid && u == 0) {
        phi[i] += sp * inv_r;
        if (GRAD) {
          const double s2 = inv_r * inv_r;
          grad[3 * size_t(i) + 0] += sx * s2;
          grad[3 * size_t(i) + 1] += sy * s2;
          grad[3 * size_t(i) + 2] += sz * s2;
        }
      }
    }
  }
}

// ---- host wrappers ---------------------------------------------------------------------

static inline unsigned grid_for(size_t n, int threads) {
  size_t g = (n + threads - 1) / threads;
  if (g > 65535u * 16u) g = 65535u * 16u;
  return unsigned(g == 0 ? 1 : g);
}

void fill(hipStream_t s, double *p, size_t n, double v) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}

int direct_chunks(int64_t ns, int64_t nt) {
  const int64_t tblocks = (nt + 255) / 256;
  int64_t nch = (1024 + tblocks - 1) / tblocks;
  const int64_t most = (ns + 1023) / 1024;  // at least 1024 sources per chunk
  if (nch > most) nch = most;
  if (nch < 1) nch = 1;
  if (nch > 65535) nch = 65535;
  return int(nch);
}

void direct_sum(hipStream_t s, const double *src, const double *q, int64_t ns, const double *tgt, int64_t nt,
                int gradients, double *part, int nchunks, double *phi, double *grad) {
  if (nt == 0) return;
  const int64_t chunk = (ns + nchunks - 1) / nchunks;
  hipLaunchKernelGGL(k_direct_partial, dim3(unsigned((nt + 255) / 256), unsigned(nchunks)), dim3(256), 0, s, src, q,
                     ns, tgt, nt, chunk, gradients, part);
  hipLaunchKernelGGL(k_direct_combine, dim3(unsigned((nt + 255) / 256)), dim3(256), 0, s, nt, nchunks, gradients,
                     part, phi, grad);
}

void vjp_combine(hipStream_t s, size_t n, const double *q, const double *g, const double *psi, const double *dpsi,
                 const double *field, double *grad_q, double *grad_pos) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_vjp_combine, dim3(grid_for(n, 256)), dim3(256), 0, s, n, q, g, psi, dpsi, field, grad_q,
                     grad_pos);
}

void build_source_records(const double *sx, const double *sy, const double *sz, size_t n, int32_t *host_rec) {
  for (size_t j = 0; j < n; ++j) {
    const double u[3] = {sx[j], sy[j], sz[j]};
    int32_t *r = host_rec + 4 * j;
    for (int k = 0; k < 3; ++k) r[k] = int32_t(std::lrint(u[k] * kP2pFixed));
    r[3] = 0;  // the charge's bits, written by gather_charges
  }
}

void gather_charges(hipStream_t s, const double *q, const uint32_t *order, size_t n, double *q64, float *rec) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_charges, dim3(grid_for(n, 256)), dim3(256), 0, s, q, order, n, q64, rec);
}

void output_pass(hipStream_t s, uint32_t nleaves, const uint32_t *tgt_start, const double *leaf_r,
                 const uint32_t *order, const double *phi_hat, const double *grad_hat, double *phi, double *grad,
                 int gradients) {
  if (nleaves == 0) return;
  hipLaunchKernelGGL(k_output, dim3(xcd_grid(nleaves)), dim3(128), 0, s, nleaves, tgt_start, leaf_r, order, phi_hat,
                     grad_hat, phi, grad, gradients);
}

template <typename T>
void p2p(hipStream_t s, uint32_t leaf_first, uint32_t count, int gradients, uint32_t uniform_min,
         const uint32_t *near_off, const uint32_t *near_ent, const uint32_t *leaf_cell, const uint32_t *src_start,
         const uint32_t *tgt_start, const double *sx, const double *sy, const double *sz, const double *sq,
         const float *rec, const uint8_t *near_close, const double *tx, const double *ty, const double *tz,
         double *phi, double *grad) {
  if (count == 0) return;
  const int4 *r4 = reinterpret_cast<const int4 *>(rec);
  const uint32_t umin = uniform_min < 1 ? 1 : uniform_min;
  if (gradients)
    hipLaunchKernelGGL((k_p2p<T, true>), dim3(xcd_grid(count)), dim3(64), 0, s, leaf_first, count, umin, near_off,
                       near_ent, leaf_cell, src_start, tgt_start, sx, sy, sz, sq, r4, near_close, tx, ty, tz, phi,
                       grad);
  else
    hipLaunchKernelGGL((k_p2p<T, false>), dim3(xcd_grid(count)), dim3(64), 0, s, leaf_first, count, umin, near_off,
                       near_ent, leaf_cell, src_start, tgt_start, sx, sy, sz, sq, r4, near_close, tx, ty, tz, phi,
                       grad);
}

template void p2p<float>(hipStream_t, uint32_t, uint32_t, int, uint32_t, const uint32_t *, const uint32_t *,
                         const uint32_t *, const uint32_t *, const uint32_t *, const double *, const double *,
                         const double *, const double *, const float *, const uint8_t *, const double *,
                         const double *, const double *, double *, double *);
template void p2p<double>(hipStream_t, uint32_t, uint32_t, int, uint32_t, const uint32_t *, const uint32_t *,
                          const uint32_t *, const uint32_t *, const uint32_t *, const double *, const double *,
                          const double *, const double *, const float *, const uint8_t *, const double *,
                          const double *, const double *, double *, double *);

// Near entries with a close pair for the mixed P2P's fixed-point form: one wave per
// target leaf; per entry the smallest r^2 over the leaf's targets and the entry's
// sources in the integer units the uniform-source form computes (the same rounding of
// x' and the same quarter-unit seed), flag = r^2 < kP2pClose.  The own leaf is always
// flagged (its coincident points and quantised-equal pairs need the test).
__global__ __launch_bounds__(64) void k_p2p_close_flags(uint32_t count, const uint32_t *__restrict__ near_off,
                                                        const uint32_t *__restrict__ near_ent,
                                                        const uint32_t *__restrict__ leaf_cell,
                                                        const uint32_t *__restrict__ src_start,
                                                        const uint32_t *__restrict__ tgt_start,
                                                        const int4 *__restrict__ rec, const double *__restrict__ tx,
                                                        const double *__restrict__ ty, const double *__restrict__ tz,
                                                        uint8_t *__restrict__ flags) {
  const uint32_t lf = blockIdx.x;
  if (lf >= count) return;
  const int lane = threadIdx.x;
  const uint32_t t0 = tgt_start[lf], t1 = tgt_start[lf + 1];
  const uint32_t *ct = leaf_cell + 4 * size_t(lf);
  for (uint32_t e = near_off[lf]; e < near_off[lf + 1]; ++e) {
    const uint32_t sl = near_ent[e];
    if (sl == lf) {
      if (lane == 0) flags[e] = 1;
      continue;
    }
    const uint32_t s0 = src_start[sl], s1 = src_start[sl + 1];
    double c[3], rho;
    relative_frame<double>(ct, leaf_cell + 4 * size_t(sl), c, rho);
    float rmin = kP2pClose;
    for (uint32_t tb = t0; tb < t1; tb += 64) {
      const uint32_t i = tb + uint32_t(lane);
      if (i >= t1) break;
      const int xi0 = __double2int_rn(fma(rho, tx[i], c[0]) * kP2pFixed);
      const int xi1 = __double2int_rn(fma(rho, ty[i], c[1]) * kP2pFixed);
      const int xi2 = __double2int_rn(fma(rho, tz[i], c[2]) * kP2pFixed);
      for (uint32_t j = s0; j < s1; ++j) {
        const int4 h = rec[j];
        const float dx = float(xi0 - h.x), dy = float(xi1 - h.y), dz = float(xi2 - h.z);
        rmin = fminf(rmin, fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 0.25f))));
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) rmin = fminf(rmin, __shfl_xor(rmin, off));
    if (lane == 0) flags[e] = rmin < kP2pClose ? 1 : 0;
  }
}

void p2p_close_flags(hipStream_t s, uint32_t nleaves, const uint32_t *near_off, const uint32_t *near_ent,
                     const uint32_t *leaf_cell, const uint32_t *src_start, const uint32_t *tgt_start,
                     const float *rec, const double *tx, const double *ty, const double *tz, uint8_t *flags) {
  if (nleaves == 0) return;
  hipLaunchKernelGGL(k_p2p_close_flags, dim3(nleaves), dim3(64), 0, s, nleaves, near_off, near_ent, leaf_cell,
                     src_start, tgt_start, reinterpret_cast<const int4 *>(rec), tx, ty, tz, flags);
}

void expand(hipStream_t s, int p, int irregular, uint32_t rows, const uint32_t *row_box, const uint32_t *row_cell,
            const uint32_t *row_off, const uint32_t *row_ent, const uint32_t *leaf_cell,
            const uint32_t *src_start, const double *sx, const double *sy, const double *sz, const double *sq,
            double *coeffs) {
  if (rows == 0) return;
  if (irregular)
    hipLaunchKernelGGL(k_expand<true>, dim3(xcd_grid(rows)), dim3(64), 0, s, p, rows, row_box, row_cell, row_off,
                       row_ent, leaf_cell, src_start, sx, sy, sz, sq, coeffs);
  else
    hipLaunchKernelGGL(k_expand<false>, dim3(xcd_grid(rows)), dim3(64), 0, s, p, rows, row_box, row_cell, row_off,
                       row_ent, leaf_cell, src_start, sx, sy, sz, sq, coeffs);
}

void evaluate(hipStream_t s, int p, int irregular, int gradients, uint32_t rows, const uint32_t *row_leaf,
              const uint32_t *row_off, const uint32_t *row_box, const uint32_t *box_cell,
              const uint32_t *leaf_cell, const uint32_t *tgt_start, const double *tx, const double *ty,
              const double *tz, const double *coeffs, double *phi, double *grad) {
  if (rows == 0) return;
  // LDS: (p+1)^2 coefficients and, with gradients, three sets of (pg+1)^2 (pg = p + 1 for
  // M2P, p - 1 for L2P); at most 1,123 doubles at p = 15
  const int pg = irregular ? p + 1 : p - 1;
  const size_t lds = sizeof(double) * (size_t((p + 1) * (p + 1)) + (gradients && pg >= 0 ? 3 * size_t((pg + 1) * (pg + 1)) : 0));
#define FMMH_EVAL(I, G)                                                                                          \
  hipLaunchKernelGGL((k_evaluate<I, G>), dim3(xcd_grid(rows)), dim3(I ? 128 : 64), lds, s, p, rows, row_leaf, \
                     row_off, row_box, box_cell, leaf_cell, tgt_start, tx, ty, tz, coeffs, phi, grad)
  if (irregular) {
    if (gradients)
      FMMH_EVAL(true, true);
    else
      FMMH_EVAL(true, false);
  } else {
    if (gradients)
      FMMH_EVAL(false, true);
    else
      FMMH_EVAL(false, false);
  }
#undef FMMH_EVAL
}

}  // namespace dev
}  // namespace fmmh
