// plan_device.hip -- the plan's per-point stages on the GPU (octree/src/octree/
// implementation.rs:682-716,1014 builds the tree from sorted Morton keys): the bounding box
// of the points, their level-16 Morton keys and their sort (device radix sort; the
// distinct set taken on the host), and, once the host has refined, balanced and listed the tree from those
// keys, each point's leaf (the last leaf at or before its key in Morton order), the points
// in leaf order (a stable radix sort by leaf: caller order within a leaf) and their
// leaf-scaled coordinates.  Every value is the host build's: the keys and coordinates are
// the same IEEE operations (no contraction), the sorts and the unique are exact.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <hipcub/hipcub.hpp>
#include <vector>

#include "common.hpp"
#include "operator.hpp"
#include "plan.hpp"

namespace fmmh {
namespace {

constexpr int kThreads = 256;

struct Box6 {
  double v[6];
};

// per block: min x, y, z, max x, y, z over its points and the count of non-finite coordinates
__global__ __launch_bounds__(kThreads) void k_bounds(const double *__restrict__ x, int64_t n,
                                                     double *__restrict__ part, uint32_t *__restrict__ bad) {
  __shared__ double s[6][kThreads];
  __shared__ uint32_t sb[kThreads];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t nb = 0;
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    for (int k = 0; k < 3; ++k) {
      const double v = x[3 * i + k];
      if (!isfinite(v)) {
        ++nb;
        continue;
      }
      lo[k] = fmin(lo[k], v);
      hi[k] = fmax(hi[k], v);
    }
  for (int k = 0; k < 3; ++k) {
    s[k][threadIdx.x] = lo[k];
    s[3 + k][threadIdx.x] = hi[k];
  }
  sb[threadIdx.x] = nb;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w /= 2) {
    if (int(threadIdx.x) < w) {
      for (int k = 0; k < 3; ++k) {
        s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        s[3 + k][threadIdx.x] = fmax(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
      }
      sb[threadIdx.x] += sb[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
  if (threadIdx.x == 0) bad[blockIdx.x] = sb[0];
}

__device__ __forceinline__ uint64_t spread3_dev(uint64_t v) {
  uint64_t r = 0;
  for (int b = 0; b < 16; ++b) r |= ((v >> b) & 1ull) << (3 * b);
  return r;
}

// morton::from_physical_point at level 16 (plan.cpp fine_key): the same operations
__device__ __forceinline__ uint64_t fine_key_dev(const double *x, const Box6 &box) {
  uint64_t idx[3];
  for (int k = 0; k < 3; ++k) {
    const double ref = (x[k] - box.v[k]) / (box.v[k + 3] - box.v[k]);
    const double v = ref * 65536.0;
    const uint64_t i = (v > 0.0) ? static_cast<uint64_t>(v) : 0;
    idx[k] = i < 65535 ? i : 65535;
  }
  const uint64_t pos = (spread3_dev(idx[0]) << 2) | (spread3_dev(idx[1]) << 1) | spread3_dev(idx[2]);
  return (pos << kLevelBits) | static_cast<uint64_t>(kDeepest);
}

__global__ __launch_bounds__(kThreads) void k_fine_keys(const double *__restrict__ x, int64_t n, Box6 box,
                                                        uint64_t *__restrict__ keys) {
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads)
    keys[i] = fine_key_dev(x + 3 * i, box);
}

// the leaf of each point: the last of the Morton-ordered leaves at or before its key
__global__ __launch_bounds__(kThreads) void k_point_leaf(const double *__restrict__ x, int64_t n, Box6 box,
                                                         const uint64_t *__restrict__ morton, uint32_t nl,
                                                         const uint32_t *__restrict__ morton_to_leaf,
                                                         uint32_t *__restrict__ leaf_of) {
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kThreads) {
    const uint64_t k = fine_key_dev(x + 3 * i, box);
    uint32_t lo = 0, hi = nl;  // first leaf with key > k
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (morton[mid] <= k)
        lo = mid + 1;
      else
        hi = mid;
    }
    leaf_of[i] = morton_to_leaf[lo - 1];
  }
}

// start[l] = the first position of leaf l among the sorted leaves of the points (l <= nl)
__global__ __launch_bounds__(kThreads) void k_leaf_starts(const uint32_t *__restrict__ sorted, int64_t n, uint32_t nl,
                                                          uint32_t *__restrict__ start) {
  const int64_t l = blockIdx.x * int64_t(kThreads) + threadIdx.x;
  if (l > nl) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (sorted[mid] < uint32_t(l))
      lo = mid + 1;
    else
      hi = mid;
  }
  start[l] = uint32_t(lo);
}

// leaf-scaled coordinates (plan.cpp leaf_coordinates): fl(fl(fl(x - a) 2^(l+1)) / w - (2i + 1))
__global__ __launch_bounds__(kThreads) void k_leaf_coordinates(const double *__restrict__ x,
                                                               const uint32_t *__restrict__ order,
                                                               const uint32_t *__restrict__ leaf_sorted, int64_t n,
                                                               const uint32_t *__restrict__ cell, double lx, double ly,
                                                               double lz, double side, double *__restrict__ u) {
  for (int64_t j = blockIdx.x * int64_t(kThreads) + threadIdx.x; j < n; j += int64_t(gridDim.x) * kThreads) {
    const double *p = x + 3 * size_t(order[j]);
    const uint32_t *c = cell + 4 * size_t(leaf_sorted[j]);
    const double scale = ldexp(1.0, int(c[0]) + 1), lower[3] = {lx, ly, lz};
    for (int k = 0; k < 3; ++k) {
      const double d = p[k] - lower[k];
      const double s = d * scale;
      u[3 * j + k] = s / side - static_cast<double>(2 * int64_t(c[1 + k]) + 1);
    }
  }
}

unsigned grid_of(int64_t n) {
  const int64_t b = (n + kThreads - 1) / kThreads;
  return unsigned(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

// the points of one set (sources or targets) placed in the leaves of the plan
void place_points(hipStream_t st, const double *dx, int64_t n, const Box6 &box, const DevBuf<uint64_t> &morton,
                  const DevBuf<uint32_t> &m2l, const DevBuf<uint32_t> &cells, const Plan &plan,
                  std::vector<uint32_t> &start, std::vector<uint32_t> &order, std::vector<double> &u) {
  const uint32_t nl = plan.nleaves();
  start.assign(nl + 1, 0);
  order.clear();
  u.clear();
  if (n == 0) return;
  DevBuf<uint32_t> leaf_of, leaf_sorted, iota, ord, dstart;
  DevBuf<double> du;
  leaf_of.alloc(n);
  leaf_sorted.alloc(n);
  iota.alloc(n);
  ord.alloc(n);
  dstart.alloc(nl + 1);
  du.alloc(3 * size_t(n));
  hipLaunchKernelGGL(k_point_leaf, dim3(grid_of(n)), dim3(kThreads), 0, st, dx, n, box, morton.ptr, nl, m2l.ptr,
                     leaf_of.ptr);
  std::vector<uint32_t> seq(n);
  for (int64_t i = 0; i < n; ++i) seq[i] = uint32_t(i);
  hip_check(hipMemcpyAsync(iota.ptr, seq.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st), "plan upload");
  int bits = 1;
  while ((uint64_t(1) << bits) <= nl) ++bits;
  size_t tmp_bytes = 0;
  hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, leaf_of.ptr, leaf_sorted.ptr, iota.ptr, ord.ptr,
                                               n, 0, bits, st),
            "plan sort (size)");
  DevBuf<uint8_t> tmp;
  tmp.alloc(tmp_bytes);
  hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, leaf_of.ptr, leaf_sorted.ptr, iota.ptr, ord.ptr,
                                               n, 0, bits, st),
            "plan sort by leaf");
  hipLaunchKernelGGL(k_leaf_starts, dim3(grid_of(int64_t(nl) + 1)), dim3(kThreads), 0, st, leaf_sorted.ptr, n, nl,
                     dstart.ptr);
  hipLaunchKernelGGL(k_leaf_coordinates, dim3(grid_of(n)), dim3(kThreads), 0, st, dx, ord.ptr, leaf_sorted.ptr, n,
                     cells.ptr, plan.lower[0], plan.lower[1], plan.lower[2], plan.side, du.ptr);
  order.resize(n);
  u.resize(3 * size_t(n));
  hip_check(hipMemcpyAsync(start.data(), dstart.ptr, (nl + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, st),
            "plan download");
  hip_check(hipMemcpyAsync(order.data(), ord.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "plan download");
  hip_check(hipMemcpyAsync(u.data(), du.ptr, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st), "plan download");
  hip_check(hipStreamSynchronize(st), "plan points");
}

}  // namespace

Plan build_plan_device(int device, const double *sources, int64_t ns, const double *targets, int64_t nt,
                       const PlanOptions &options) {
  Plan plan = plan_begin(ns, nt, sources == targets && ns == nt, options);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(kNoDevice, "no HIP device");
  require(device >= 0 && device < ndev, "device ordinal out of range", kNoDevice);
  hip_check(hipSetDevice(device), "hipSetDevice");
  hipStream_t st = nullptr;
  hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  const int64_t n = ns + nt;
  DevBuf<double> pts;
  pts.alloc(3 * size_t(n));
  if (ns) hip_check(hipMemcpyAsync(pts.ptr, sources, 3 * ns * sizeof(double), hipMemcpyHostToDevice, st), "upload");
  if (nt)
    hip_check(hipMemcpyAsync(pts.ptr + 3 * ns, targets, 3 * nt * sizeof(double), hipMemcpyHostToDevice, st), "upload");
  // the bounding box: per-block partials, then the host
  const unsigned nb = grid_of(n) < 1024 ? grid_of(n) : 1024;
  DevBuf<double> part;
  DevBuf<uint32_t> bad;
  part.alloc(6 * size_t(nb));
  bad.alloc(nb);
  hipLaunchKernelGGL(k_bounds, dim3(nb), dim3(kThreads), 0, st, pts.ptr, n, part.ptr, bad.ptr);
  std::vector<double> hp(6 * size_t(nb));
  std::vector<uint32_t> hb(nb);
  hip_check(hipMemcpyAsync(hp.data(), part.ptr, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st), "download");
  hip_check(hipMemcpyAsync(hb.data(), bad.ptr, hb.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "download");
  hip_check(hipStreamSynchronize(st), "plan bounds");
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint64_t nbad = 0;
  for (unsigned b = 0; b < nb; ++b) {
    nbad += hb[b];
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], hp[6 * b + k]);
      hi[k] = std::max(hi[k], hp[6 * b + 3 + k]);
    }
  }
  require(nbad == 0, "build_plan: a coordinate is not finite");
  plan_domain(plan, lo, hi);
  Box6 box;
  for (int k = 0; k < 6; ++k) box.v[k] = plan.box[k];
  // the fine keys of all points, sorted on the device; their distinct set on the host
  DevBuf<uint64_t> keys, sorted;
  keys.alloc(n);
  sorted.alloc(n);
  hipLaunchKernelGGL(k_fine_keys, dim3(grid_of(n)), dim3(kThreads), 0, st, pts.ptr, n, box, keys.ptr);
  size_t sort_bytes = 0;
  hip_check(hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, keys.ptr, sorted.ptr, n, 0, 64, st),
            "plan sort (size)");
  DevBuf<uint8_t> tmp;
  tmp.alloc(sort_bytes);
  hip_check(hipcub::DeviceRadixSort::SortKeys(tmp.ptr, sort_bytes, keys.ptr, sorted.ptr, n, 0, 64, st),
            "plan sort of the keys");
  std::vector<uint64_t> fine(n);
  hip_check(hipMemcpyAsync(fine.data(), sorted.ptr, n * sizeof(uint64_t), hipMemcpyDeviceToHost, st), "download");
  hip_check(hipStreamSynchronize(st), "plan keys");
  keys.release();
  sorted.release();
  tmp.release();
  require(std::is_sorted(fine.begin(), fine.end()), "build_plan_device: the device sort returned unsorted keys",
          kDeviceError);
  fine.erase(std::unique(fine.begin(), fine.end()), fine.end());
  // the tree and lists on the host, then the points on the device
  plan_tree(plan, std::move(fine));
  DevBuf<uint64_t> morton;
  DevBuf<uint32_t> m2l, cells;
  morton.upload(plan.morton_leaves);
  m2l.upload(plan.morton_to_leaf);
  std::vector<uint32_t> cell(4 * size_t(plan.nleaves()));
  for (uint32_t i = 0; i < plan.nleaves(); ++i) {
    const Cell c = decode(plan.leaf_key[i]);
    cell[4 * i] = uint32_t(c.level);
    cell[4 * i + 1] = uint32_t(c.i);
    cell[4 * i + 2] = uint32_t(c.j);
    cell[4 * i + 3] = uint32_t(c.k);
  }
  cells.upload(cell);
  place_points(st, pts.ptr, ns, box, morton, m2l, cells, plan, plan.src_start, plan.src_order, plan.us);
  place_points(st, pts.ptr + 3 * ns, nt, box, morton, m2l, cells, plan, plan.tgt_start, plan.tgt_order, plan.ut);
  plan_finish(plan);
  return plan;
}

}  // namespace fmmh
