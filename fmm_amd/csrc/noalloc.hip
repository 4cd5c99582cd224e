// noalloc.hip -- the compensated direct sum without any allocation, for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
// that forbid hipMalloc).
//
// fmmh_direct_sum allocates its per-chunk partial sums (8 doubles per target and source
// chunk) and a stream on every call.  This variant takes device pointers and the caller's
// stream, and keeps the partial sums in `scratch`, one fixed array per GPU that lives as
// long as the process; more than kNoAllocTargets targets is refused.  Calls that overlap
// on one GPU share the array and race; the caller orders them (one stream, or events).
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"
#include "operator.hpp"

// 32 MiB: 8 doubles per (target, source chunk).  The allocating path's chunk count is
// ceil(1024 / ceil(nt / 256)) for nt <= 262,144 targets, so chunks x targets stays below
// 262,144 + 256 ceil(nt / 256) <= 524,288: every accepted call uses the same chunks (and
// gives the same bits) as fmmh_direct_sum.
constexpr size_t kNoAllocDoubles = size_t(4) << 20;
constexpr int64_t kNoAllocTargets = 262144;

// Host code never uses this name as an address: the device address of each GPU's copy
// comes from hipGetSymbolAddress with that GPU current (scratch_on).
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace fmmh {

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
std::array<double *, kMaxDevices> g_scratch_ptr{};

// The device address of `scratch` on `device` (the current device), looked up once.
double *scratch_on(int device) {
  require(device >= 0 && device < kMaxDevices, "direct_sum_noalloc: device ordinal out of range", kNoDevice);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    double *ptr = nullptr;
    hip_check(hipGetSymbolAddress(reinterpret_cast<void **>(&ptr), HIP_SYMBOL(scratch)),
              "direct_sum_noalloc: hipGetSymbolAddress");
    require(ptr != nullptr, "direct_sum_noalloc: no device address for the scratch array", kDeviceError);
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

}  // namespace

int64_t direct_sum_noalloc_max_targets() { return kNoAllocTargets; }

void direct_sum_noalloc(int device, const double *sources, const double *charges, int64_t ns, const double *targets,
                        int64_t nt, double *phi, double *grad, hipStream_t stream) {
  require(ns >= 0 && nt >= 0, "direct_sum_noalloc: negative size");
  require(nt <= direct_sum_noalloc_max_targets(),
          "direct_sum_noalloc: more targets than the scratch array holds (262,144)", kTooLarge);
  if (nt == 0) return;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(kNoDevice, "no HIP device");
  require(device < ndev, "direct_sum_noalloc: device ordinal out of range", kNoDevice);
  hip_check(hipSetDevice(device), "hipSetDevice");
  double *part = scratch_on(device);
  // source chunks: as the allocating path chooses them (the cut below never applies to
  // an accepted size; it keeps the array's bound explicit)
  int nch = dev::direct_chunks(ns, nt);
  const int64_t fit = int64_t(kNoAllocDoubles / (8 * size_t(nt)));
  if (nch > fit) nch = int(fit);
  // every (chunk, target) slot the combine reads is written by the partial kernel first,
  // so the array needs no clearing between calls
  dev::direct_sum(stream, sources, charges, ns, targets, nt, grad != nullptr, part, nch, phi, grad);
  hip_check(hipGetLastError(), "direct_sum_noalloc: launch");
}

}  // namespace fmmh
