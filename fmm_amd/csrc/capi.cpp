// capi.cpp -- the extern "C" boundary declared in include/fmm_amd.h.
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fmm_amd.h"
#include "hoperator.hpp"
#include "hpwtables.hpp"
#include "htables.hpp"
#include "kernels.hpp"
#include "operator.hpp"
#include "plan.hpp"
#include "pwtables.hpp"
#include "table_cache.hpp"
#include "tune.hpp"

struct fmmh_plan {
  fmmh::Plan plan;
};

struct fmmh_operator {
  std::unique_ptr<fmmh::DeviceOperator> op;
  int64_t nlevels = 0;
  int64_t nleaves = 0;
  // The first device failure of an evaluation, kept: every later call that touches the
  // device returns it (fmm-exec/src/device.rs:414-425).  inject: the next such call
  // fails with a device error of its own (the test hook, device.rs:4621).
  std::string failed;
  bool inject = false;
};

struct fmmh_helmholtz {
  std::unique_ptr<fmmh::HelmholtzOperator> op;
  int64_t nlevels = 0;
  int64_t nleaves = 0;
  int64_t nboxes = 0;
};

namespace {

thread_local std::string g_last_error;

template <class F>
int32_t guard(F &&f) {
  try {
    f();
    g_last_error.clear();
    return FMMH_OK;
  } catch (const fmmh::Error &e) {
    g_last_error = e.what();
    return static_cast<int32_t>(e.status);
  } catch (const std::bad_alloc &) {
    g_last_error = "host out of memory";
    return FMMH_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return FMMH_DEVICE;
  }
}

void need(bool ok, const char *what) {
  if (!ok) throw fmmh::Error(fmmh::kInvalidArgument, what);
}

void need(bool ok, const std::string &what) {
  if (!ok) throw fmmh::Error(fmmh::kInvalidArgument, what);
}

// A call on an operator that touches the device: the kept failure, if any, is returned
// without device work; a device error raised by the call is kept.  Argument errors are
// not kept (they are the caller's, found before any device work).
template <class F>
int32_t device_call(fmmh_operator *op, const char *name, F &&f) {
  return guard([&] {
    need(op != nullptr, (std::string(name) + ": NULL operator").c_str());
    if (op->inject) {
      op->inject = false;
      op->failed = std::string(name) + ": injected device error (test hook)";
    }
    if (!op->failed.empty()) throw fmmh::Error(fmmh::kDeviceError, op->failed);
    try {
      f();
    } catch (const fmmh::Error &e) {
      if (e.status == fmmh::kDeviceError) op->failed = e.what();
      throw;
    }
  });
}

}  // namespace

extern "C" {

const char *fmmh_last_error(void) { return g_last_error.c_str(); }

int32_t fmmh_abi_version(void) { return 6; }

int32_t fmmh_plan_build(const double *sources, int64_t nsources, const double *targets, int64_t ntargets,
                        const fmmh_plan_options *options, fmmh_plan **out) {
  return guard([&] {
    need(out != nullptr, "fmmh_plan_build: out is NULL");
    need(nsources == 0 || sources != nullptr, "fmmh_plan_build: sources is NULL");
    need(ntargets == 0 || targets != nullptr, "fmmh_plan_build: targets is NULL");
    fmmh::PlanOptions o;
    if (options) {
      o.leaf_points = options->leaf_points;
      o.max_level = options->max_level;
      o.near_w = options->near_w;
      o.near_x = options->near_x;
    }
    need(o.max_level >= 0 && o.max_level <= fmmh::kDeepest, "fmmh_plan_build: max_level must lie in 0..16");
    auto p = std::make_unique<fmmh_plan>();
    p->plan = fmmh::build_plan(sources, nsources, targets, ntargets, o);
    *out = p.release();
  });
}

int32_t fmmh_plan_build_device(int32_t device, const double *sources, int64_t nsources, const double *targets,
                               int64_t ntargets, const fmmh_plan_options *options, fmmh_plan **out) {
  return guard([&] {
    need(out != nullptr, "fmmh_plan_build_device: out is NULL");
    need(nsources == 0 || sources != nullptr, "fmmh_plan_build_device: sources is NULL");
    need(ntargets == 0 || targets != nullptr, "fmmh_plan_build_device: targets is NULL");
    fmmh::PlanOptions o;
    if (options) {
      o.leaf_points = options->leaf_points;
      o.max_level = options->max_level;
      o.near_w = options->near_w;
      o.near_x = options->near_x;
    }
    need(o.max_level >= 0 && o.max_level <= fmmh::kDeepest, "fmmh_plan_build_device: max_level must lie in 0..16");
    auto p = std::make_unique<fmmh_plan>();
    p->plan = fmmh::build_plan_device(device, sources, nsources, targets, ntargets, o);
    *out = p.release();
  });
}

int32_t fmmh_forest_build(const double *points, const int64_t *counts, int32_t nclouds,
                          const fmmh_plan_options *options, fmmh_plan **out) {
  return guard([&] {
    need(out != nullptr, "fmmh_forest_build: out is NULL");
    need(points != nullptr && counts != nullptr && nclouds >= 1, "fmmh_forest_build: no clouds");
    fmmh::PlanOptions o;
    if (options) {
      o.leaf_points = options->leaf_points;
      o.max_level = options->max_level;
      o.near_w = options->near_w;
      o.near_x = options->near_x;
    }
    need(o.max_level >= 0 && o.max_level <= fmmh::kDeepest, "fmmh_forest_build: max_level must lie in 0..16");
    auto p = std::make_unique<fmmh_plan>();
    p->plan = fmmh::build_forest(points, counts, nclouds, o);
    *out = p.release();
  });
}

void fmmh_plan_free(fmmh_plan *plan) { delete plan; }

int32_t fmmh_plan_nlevels(const fmmh_plan *plan) { return plan ? plan->plan.nlevels : -1; }

int64_t fmmh_plan_nboxes(const fmmh_plan *plan, int32_t level) {
  if (!plan || level < 0 || level >= plan->plan.nlevels) return 0;
  return plan->plan.nboxes(level);
}

int64_t fmmh_plan_nleaves(const fmmh_plan *plan) { return plan ? plan->plan.nleaves() : -1; }

int32_t fmmh_plan_domain(const fmmh_plan *plan, double *box, double *side) {
  return guard([&] {
    need(plan && box && side, "fmmh_plan_domain: NULL argument");
    std::memcpy(box, plan->plan.box.data(), 6 * sizeof(double));
    *side = plan->plan.side;
  });
}

int32_t fmmh_plan_box_keys(const fmmh_plan *plan, int32_t level, uint64_t *out) {
  return guard([&] {
    need(plan && out, "fmmh_plan_box_keys: NULL argument");
    need(level >= 0 && level < plan->plan.nlevels, "fmmh_plan_box_keys: level out of range");
    const auto &k = plan->plan.keys[level];
    std::memcpy(out, k.data(), k.size() * sizeof(uint64_t));
  });
}

int32_t fmmh_plan_leaf_keys(const fmmh_plan *plan, uint64_t *out) {
  return guard([&] {
    need(plan && out, "fmmh_plan_leaf_keys: NULL argument");
    std::memcpy(out, plan->plan.leaf_key.data(), plan->plan.leaf_key.size() * sizeof(uint64_t));
  });
}

int64_t fmmh_plan_list(const fmmh_plan *plan, int32_t kind, int32_t level, uint32_t *offsets, uint32_t *entries,
                       uint16_t *aux) {
  if (!plan || level < 0 || level >= plan->plan.nlevels || kind < 0 || kind > 3) {
    g_last_error = "fmmh_plan_list: bad plan, kind or level";
    return -1;
  }
  const fmmh::Plan &p = plan->plan;
  const fmmh::Csr &c = kind == 0 ? p.near[level] : kind == 1 ? p.v[level] : kind == 2 ? p.w[level] : p.x[level];
  if (offsets && !c.off.empty()) std::memcpy(offsets, c.off.data(), c.off.size() * sizeof(uint32_t));
  if (entries && !c.ent.empty()) std::memcpy(entries, c.ent.data(), c.ent.size() * sizeof(uint32_t));
  if (aux && kind == 1 && !c.aux.empty()) std::memcpy(aux, c.aux.data(), c.aux.size() * sizeof(uint16_t));
  return static_cast<int64_t>(c.size());
}

int32_t fmmh_plan_point_leaves(const fmmh_plan *plan, int32_t targets, uint32_t *out) {
  return guard([&] {
    need(plan && out, "fmmh_plan_point_leaves: NULL argument");
    const fmmh::Plan &p = plan->plan;
    const auto &start = targets ? p.tgt_start : p.src_start;
    const auto &order = targets ? p.tgt_order : p.src_order;
    for (uint32_t l = 0; l < p.nleaves(); ++l)
      for (uint32_t j = start[l]; j < start[l + 1]; ++j) out[order[j]] = l;
  });
}

int32_t fmmh_plan_leaf_coordinates(const fmmh_plan *plan, int32_t targets, double *coords, uint32_t *order) {
  return guard([&] {
    need(plan != nullptr, "fmmh_plan_leaf_coordinates: NULL plan");
    const fmmh::Plan &p = plan->plan;
    const auto &u = targets ? p.ut : p.us;
    const auto &o = targets ? p.tgt_order : p.src_order;
    if (coords) std::memcpy(coords, u.data(), u.size() * sizeof(double));
    if (order) std::memcpy(order, o.data(), o.size() * sizeof(uint32_t));
  });
}

int32_t fmmh_plan_leaf_cells(const fmmh_plan *plan, uint32_t *out) {
  return guard([&] {
    need(plan && out, "fmmh_plan_leaf_cells: NULL argument");
    const fmmh::Plan &p = plan->plan;
    for (size_t i = 0; i < p.leaf_key.size(); ++i) {
      const fmmh::Cell c = fmmh::decode(p.leaf_key[i]);
      out[4 * i] = uint32_t(c.level);
      out[4 * i + 1] = uint32_t(c.i);
      out[4 * i + 2] = uint32_t(c.j);
      out[4 * i + 3] = uint32_t(c.k);
    }
  });
}

int32_t fmmh_plan_box_cells(const fmmh_plan *plan, int32_t level, uint32_t *out) {
  return guard([&] {
    need(plan && out, "fmmh_plan_box_cells: NULL argument");
    const fmmh::Plan &p = plan->plan;
    need(level >= 0 && level < p.nlevels, "fmmh_plan_box_cells: level out of range");
    for (size_t i = 0; i < p.keys[level].size(); ++i) {
      const fmmh::Cell c = fmmh::decode(p.keys[level][i]);
      out[4 * i] = uint32_t(c.level);
      out[4 * i + 1] = uint32_t(c.i);
      out[4 * i + 2] = uint32_t(c.j);
      out[4 * i + 3] = uint32_t(c.k);
    }
  });
}

int32_t fmmh_plan_point_starts(const fmmh_plan *plan, int32_t targets, uint32_t *out) {
  return guard([&] {
    need(plan && out, "fmmh_plan_point_starts: NULL argument");
    const auto &v = targets ? plan->plan.tgt_start : plan->plan.src_start;
    std::memcpy(out, v.data(), v.size() * sizeof(uint32_t));
  });
}

int32_t fmmh_m2l_rotation_table(int32_t p, int32_t offset, double *table) {
  return guard([&] {
    need(table != nullptr && offset >= 0 && offset < 316, "fmmh_m2l_rotation_table: bad argument");
    std::vector<double> fwd, bwd, coax, dist2;
    std::vector<uint16_t> dist_of;
    fmmh::build_rotation_tables(p, fwd, bwd, coax, dist_of, dist2);
    const int nc = (p + 1) * (p + 1), RB = fmmh::rotation_block_size(p), CB = fmmh::coaxial_block_size(p);
    // column c of G A F: F e_c, then A, then G
    std::vector<double> x(nc), y(nc);
    for (int c = 0; c < nc; ++c) {
      std::fill(x.begin(), x.end(), 0.0);
      x[c] = 1.0;
      const int n0 = int(std::sqrt(double(c)));
      // F: degree n0 block
      std::fill(y.begin(), y.end(), 0.0);
      size_t bo = 0;
      for (int n = 0; n < n0; ++n) bo += size_t(2 * n + 1) * (2 * n + 1);
      const int k = 2 * n0 + 1, cc = c - n0 * n0;
      for (int r = 0; r < k; ++r) y[n0 * n0 + r] = fwd[size_t(offset) * RB + bo + r + size_t(cc) * k];
      // A: per order
      std::fill(x.begin(), x.end(), 0.0);
      size_t ao = size_t(dist_of[offset]) * CB;
      for (int i = 0; i <= p; ++i) {
        const int w = p + 1 - i;
        for (int j = i; j <= p; ++j)
          for (int n = i; n <= p; ++n) {
            const double a = coax[ao + (j - i) + size_t(n - i) * w];
            x[j * j + j + i] += a * y[n * n + n + i];
            if (i > 0) x[j * j + j - i] += a * y[n * n + n - i];
          }
        ao += size_t(w) * w;
      }
      // G: per degree
      bo = 0;
      for (int n = 0; n <= p; ++n) {
        const int kk = 2 * n + 1;
        for (int r = 0; r < kk; ++r) {
          double s = 0.0;
          for (int q = 0; q < kk; ++q) s += bwd[size_t(offset) * RB + bo + r + size_t(q) * kk] * x[n * n + q];
          table[(n * n + r) + size_t(c) * nc] = s;
        }
        bo += size_t(kk) * kk;
      }
    }
  });
}

int32_t fmmh_set_table_cache(const char *directory) {
  return guard([&] { fmmh::set_table_cache(directory); });
}

int32_t fmmh_table_cache_stats(int64_t *hits, int64_t *misses) {
  return guard([&] { fmmh::table_cache_stats(hits, misses); });
}

int32_t fmmh_tables(int32_t p, double *m2m, double *l2l, double *m2l) {
  return guard([&] {
    need(p >= 0 && p <= 30, "fmmh_tables: p must lie in 0..30");
    std::vector<double> a, b, c;
    fmmh::cached_tables(p, a, b, c);
    if (m2m) std::memcpy(m2m, a.data(), a.size() * sizeof(double));
    if (l2l) std::memcpy(l2l, b.data(), b.size() * sizeof(double));
    if (m2l) std::memcpy(m2l, c.data(), c.size() * sizeof(double));
  });
}

int32_t fmmh_device_count(int32_t *count) {
  return guard([&] {
    need(count != nullptr, "fmmh_device_count: NULL count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

int32_t fmmh_device_describe(int32_t device, char *name, int32_t capacity) {
  return guard([&] {
    need(name != nullptr && capacity > 0, "fmmh_device_describe: NULL or empty name buffer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    need(device >= 0 && device < n, "fmmh_device_describe: no such device");
    hipDeviceProp_t prop;
    fmmh::hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    const std::string s = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount) + "cu";
    need(int64_t(s.size()) < capacity, "fmmh_device_describe: name buffer too small");
    std::memcpy(name, s.c_str(), s.size() + 1);
  });
}

int32_t fmmh_helmholtz_plane_wave_quadrature(double kappa_re, double kappa_im, int32_t capacity, int32_t *domain,
                                             int32_t *nodes, double *stated_error, double *lambda, double *t,
                                             double *weight, int32_t *alpha, int32_t *fourier) {
  return guard([&] {
    need(domain && nodes, "fmmh_helmholtz_plane_wave_quadrature: NULL domain or nodes");
    const fmmh::HpwQuadrature q = fmmh::hpw_quadrature(fmmh::cplx(kappa_re, kappa_im));
    *domain = q.domain;
    *nodes = static_cast<int32_t>(q.lambda.size());
    if (stated_error) *stated_error = q.stated_error;
    need(int(q.lambda.size()) <= capacity, "fmmh_helmholtz_plane_wave_quadrature: capacity below the node count");
    for (size_t k = 0; k < q.lambda.size(); ++k) {
      if (lambda) lambda[2 * k] = q.lambda[k].real(), lambda[2 * k + 1] = q.lambda[k].imag();
      if (t) t[2 * k] = q.t[k].real(), t[2 * k + 1] = q.t[k].imag();
      if (weight) weight[2 * k] = q.weight[k].real(), weight[2 * k + 1] = q.weight[k].imag();
      if (alpha) alpha[k] = q.alpha[k];
      if (fourier) fourier[k] = q.fourier[k];
    }
  });
}

int32_t fmmh_bessel_j_norms(double z_re, double z_im, int32_t top, double *out) {
  return guard([&] {
    need(out != nullptr && top >= 0, "fmmh_bessel_j_norms: NULL out or negative top");
    const std::vector<double> v = fmmh::hpw_bessel_j_norms(fmmh::cplx(z_re, z_im), top);
    std::copy(v.begin(), v.end(), out);
  });
}

// The last Helmholtz plane-wave level built by the host restatements, kept for the next
// call of the same (p, kh).
static const fmmh::HpwLevel &cached_hpw_level(int32_t p, fmmh::cplx kh, std::unique_lock<std::mutex> &lock) {
  static std::mutex mu;
  static std::unique_ptr<fmmh::HpwLevel> last;
  lock = std::unique_lock<std::mutex>(mu);
  if (!last || last->p != p || last->kh != kh) {
    need(p >= 0 && p <= 30, "helmholtz plane wave: p must lie in 0..30");
    last = std::make_unique<fmmh::HpwLevel>(fmmh::hpw_level(p, kh, fmmh::hpw_quadrature(2.0 * kh)));
  }
  return *last;
}

int32_t fmmh_helmholtz_plane_wave_m2l(int32_t p, double kh_re, double kh_im, int32_t offset, int32_t direction,
                                      const double *in, double *out) {
  return guard([&] {
    need(in && out, "fmmh_helmholtz_plane_wave_m2l: NULL argument");
    need(offset >= 0 && offset < 316, "fmmh_helmholtz_plane_wave_m2l: offset index 0..315");
    need(direction >= -1 && direction < 6, "fmmh_helmholtz_plane_wave_m2l: direction -1..5");
    std::unique_lock<std::mutex> lock;
    const fmmh::HpwLevel &L = cached_hpw_level(p, fmmh::cplx(kh_re, kh_im), lock);
    int d[3];
    fmmh::offset_of_index(offset, d);
    fmmh::hpw_m2l_host(L, d, direction, in, out);
  });
}

int32_t fmmh_helmholtz_plane_wave_merge(int32_t p, double kh_re, double kh_im, const double *slot_multipoles,
                                        const uint8_t *present, double *out) {
  return guard([&] {
    need(slot_multipoles && present && out, "fmmh_helmholtz_plane_wave_merge: NULL argument");
    std::unique_lock<std::mutex> lock;
    const fmmh::HpwLevel &L = cached_hpw_level(p, fmmh::cplx(kh_re, kh_im), lock);
    const size_t nc = size_t(p + 1) * (p + 1);
    std::vector<const double *> slots(fmmh::kPwSlots, nullptr);
    for (int a = 0; a < fmmh::kPwSlots; ++a)
      if (present[a]) slots[a] = slot_multipoles + a * 2 * nc;
    fmmh::hpw_merge_host(L, slots.data(), out);
  });
}

// ---- kernels: thin launches of the operator's own kernels on caller memory ------------

static hipStream_t kernel_stream(int32_t device, void *stream) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw fmmh::Error(fmmh::kNoDevice, "no HIP device");
  need(device >= 0 && device < n, "device ordinal out of range");
  fmmh::hip_check(hipSetDevice(device), "hipSetDevice");
  return static_cast<hipStream_t>(stream);
}

int32_t fmmh_kernel_p2p(int32_t device, const fmmh_p2p_inputs *in, int32_t gradients, double *phi_hat,
                        double *grad_hat, void *stream) {
  return guard([&] {
    need(in && phi_hat && (!gradients || grad_hat), "fmmh_kernel_p2p: NULL argument");
    need(in->near_off && in->near_ent && in->leaf_cell && in->src_start && in->tgt_start && in->sx && in->sy &&
             in->sz && in->q && in->tx && in->ty && in->tz,
         "fmmh_kernel_p2p: NULL input array");
    const hipStream_t s = kernel_stream(device, stream);
    fmmh::dev::p2p<double>(s, in->leaf_first, in->count, gradients ? 1 : 0, 0, in->near_off, in->near_ent, in->leaf_cell,
                   in->src_start, in->tgt_start, in->sx, in->sy, in->sz, in->q, nullptr, nullptr, in->tx, in->ty,
                   in->tz, phi_hat, gradients ? grad_hat : nullptr);
    fmmh::hip_check(hipGetLastError(), "fmmh_kernel_p2p launch");
  });
}

static int32_t kernel_expand(int32_t device, int32_t p, int irregular, const fmmh_source_inputs *in, double *coeffs,
                             void *stream, const char *name) {
  return guard([&] {
    need(in && coeffs, std::string(name) + ": NULL argument");
    need(p >= 0 && p <= fmmh::dev::kMaxDegree, std::string(name) + ": 0 <= p <= 15");
    need(in->rows == 0 || (in->row_box && in->row_cell && in->row_off && in->row_ent && in->leaf_cell &&
                           in->src_start && in->sx && in->sy && in->sz && in->q),
         std::string(name) + ": NULL input array");
    const hipStream_t s = kernel_stream(device, stream);
    fmmh::dev::expand(s, p, irregular, in->rows, in->row_box, in->row_cell, in->row_off, in->row_ent, in->leaf_cell,
                      in->src_start, in->sx, in->sy, in->sz, in->q, coeffs);
    fmmh::hip_check(hipGetLastError(), (std::string(name) + " launch").c_str());
  });
}

int32_t fmmh_kernel_p2m(int32_t device, int32_t p, const fmmh_source_inputs *in, double *coeffs, void *stream) {
  return kernel_expand(device, p, 0, in, coeffs, stream, "fmmh_kernel_p2m");
}

int32_t fmmh_kernel_p2l(int32_t device, int32_t p, const fmmh_source_inputs *in, double *coeffs, void *stream) {
  return kernel_expand(device, p, 1, in, coeffs, stream, "fmmh_kernel_p2l");
}

static int32_t kernel_evaluate(int32_t device, int32_t p, int irregular, int32_t gradients,
                               const fmmh_target_inputs *in, const double *coeffs, double *phi_hat, double *grad_hat,
                               void *stream, const char *name) {
  return guard([&] {
    need(in && coeffs && phi_hat && (!gradients || grad_hat), std::string(name) + ": NULL argument");
    need(p >= 0 && p <= fmmh::dev::kMaxDegree, std::string(name) + ": 0 <= p <= 15");
    need(in->rows == 0 || (in->row_leaf && in->row_off && in->row_box && in->box_cell && in->leaf_cell &&
                           in->tgt_start && in->tx && in->ty && in->tz),
         std::string(name) + ": NULL input array");
    const hipStream_t s = kernel_stream(device, stream);
    fmmh::dev::evaluate(s, p, irregular, gradients ? 1 : 0, in->rows, in->row_leaf, in->row_off, in->row_box,
                        in->box_cell, in->leaf_cell, in->tgt_start, in->tx, in->ty, in->tz, coeffs, phi_hat,
                        gradients ? grad_hat : nullptr);
    fmmh::hip_check(hipGetLastError(), (std::string(name) + " launch").c_str());
  });
}

int32_t fmmh_kernel_l2p(int32_t device, int32_t p, int32_t gradients, const fmmh_target_inputs *in,
                        const double *coeffs, double *phi_hat, double *grad_hat, void *stream) {
  return kernel_evaluate(device, p, 0, gradients, in, coeffs, phi_hat, grad_hat, stream, "fmmh_kernel_l2p");
}

int32_t fmmh_kernel_m2p(int32_t device, int32_t p, int32_t gradients, const fmmh_target_inputs *in,
                        const double *coeffs, double *phi_hat, double *grad_hat, void *stream) {
  return kernel_evaluate(device, p, 1, gradients, in, coeffs, phi_hat, grad_hat, stream, "fmmh_kernel_m2p");
}

int32_t fmmh_kernel_zero(int32_t device, double *values, int64_t count, void *stream) {
  return guard([&] {
    need(count >= 0 && (count == 0 || values), "fmmh_kernel_zero: NULL values or negative count");
    const hipStream_t s = kernel_stream(device, stream);
    fmmh::dev::fill(s, values, size_t(count), 0.0);
    fmmh::hip_check(hipGetLastError(), "fmmh_kernel_zero launch");
  });
}

int32_t fmmh_kernel_gather_output(int32_t device, uint32_t nleaves, const uint32_t *tgt_start, const double *leaf_r,
                                  const uint32_t *order, const double *phi_hat, const double *grad_hat,
                                  int32_t gradients, double *phi, double *grad, void *stream) {
  return guard([&] {
    need(nleaves == 0 || (tgt_start && leaf_r && order && phi_hat && phi && (!gradients || (grad_hat && grad))),
         "fmmh_kernel_gather_output: NULL argument");
    const hipStream_t s = kernel_stream(device, stream);
    fmmh::dev::output_pass(s, nleaves, tgt_start, leaf_r, order, phi_hat, grad_hat, phi, grad, gradients ? 1 : 0);
    fmmh::hip_check(hipGetLastError(), "fmmh_kernel_gather_output launch");
  });
}

}  // extern "C"

namespace {

fmmh::OperatorOptions options_of(const fmmh_operator_options *options, const fmmh_tuning *tuning) {
  fmmh::OperatorOptions o;
  if (options) {
    o.p = options->p;
    o.precision = options->precision;
    o.gradients = options->gradients;
    o.device = options->device;
    o.m2l = options->m2l;
  }
  if (tuning) {
    o.tuning.p2p_uniform_min = tuning->p2p_uniform_min;
    o.tuning.m2l_stage = tuning->m2l_stage;
    for (int l = 0; l < fmmh::kTuneLevels; ++l) {
      o.tuning.m2l_mode[l] = tuning->m2l_mode[l];
      o.tuning.l2l_width[l] = tuning->l2l_width[l];
      o.tuning.m2m_form[l] = tuning->m2m_form[l];
    }
  }
  return o;
}

}  // namespace

extern "C" {

int32_t fmmh_operator_create(const fmmh_plan *plan, const fmmh_operator_options *options, fmmh_operator **out) {
  return fmmh_operator_create_tuned(plan, options, nullptr, out);
}

int32_t fmmh_operator_create_tuned(const fmmh_plan *plan, const fmmh_operator_options *options,
                                   const fmmh_tuning *tuning, fmmh_operator **out) {
  return guard([&] {
    need(plan && out, "fmmh_operator_create: NULL argument");
    if (tuning) {
      need(tuning->p2p_uniform_min >= 0, "fmmh_operator_create_tuned: p2p_uniform_min must be >= 0");
      need(tuning->m2l_stage >= 0 && tuning->m2l_stage <= 2, "fmmh_operator_create_tuned: M2L stage 0..2");
      for (int l = 0; l < fmmh::kTuneLevels; ++l) {
        need(tuning->m2l_mode[l] >= 0 && tuning->m2l_mode[l] <= 2, "fmmh_operator_create_tuned: M2L mode 0..2");
        const int w = tuning->l2l_width[l];
        need(w == 0 || w == 16 || w == 64 || w == 128, "fmmh_operator_create_tuned: L2L width 0, 16, 64, 128");
        need(tuning->m2m_form[l] >= 0 && tuning->m2m_form[l] <= 2, "fmmh_operator_create_tuned: M2M form 0..2");
      }
    }
    const fmmh::OperatorOptions o = options_of(options, tuning);
    auto h = std::make_unique<fmmh_operator>();
    h->op = std::make_unique<fmmh::DeviceOperator>(plan->plan, o);
    h->nlevels = plan->plan.nlevels;
    h->nleaves = plan->plan.nleaves();
    *out = h.release();
  });
}

void fmmh_operator_free(fmmh_operator *op) { delete op; }

int32_t fmmh_autotune(const fmmh_plan *plan, const fmmh_operator_options *options, double budget_seconds,
                      fmmh_tuning *out, double *report) {
  return guard([&] {
    need(plan && options && out, "fmmh_autotune: NULL argument");
    need(budget_seconds >= 0.0, "fmmh_autotune: negative budget");
    const fmmh::OperatorOptions o = options_of(options, nullptr);
    fmmh::Tuning t;
    fmmh::TuneReport r;
    fmmh::autotune(plan->plan, o, budget_seconds, t, r);
    out->p2p_uniform_min = t.p2p_uniform_min;
    out->m2l_stage = t.m2l_stage;
    for (int l = 0; l < fmmh::kTuneLevels; ++l) {
      out->m2l_mode[l] = t.m2l_mode[l];
      out->l2l_width[l] = t.l2l_width[l];
      out->m2m_form[l] = t.m2m_form[l];
    }
    if (report) std::memcpy(report, &r.ms[0][0][0], sizeof(r.ms));
  });
}

int32_t fmmh_operator_stats(const fmmh_operator *op, fmmh_stats *out) {
  return guard([&] {
    need(op && out, "fmmh_operator_stats: NULL argument");
    out->nlevels = op->op->nlevels();
    out->nleaves = op->nleaves;
    out->nboxes = static_cast<int64_t>(op->op->nboxes_total());
    out->p2p_pairs = op->op->p2p_pairs();
    out->m2l_pairs = op->op->m2l_pairs();
    out->m2l_slots = op->op->m2l_slots();
    out->device_bytes = op->op->device_bytes();
  });
}

int32_t fmmh_load_charges(fmmh_operator *op, const double *charges, int64_t count, int32_t on_device) {
  return device_call(op, "fmmh_load_charges", [&] {
    need(charges != nullptr, "fmmh_load_charges: NULL argument");
    need(count == op->op->nsources(), "fmmh_load_charges: count differs from the plan's source count");
    op->op->load_charges(charges, on_device != 0);
  });
}

int32_t fmmh_reset(fmmh_operator *op) {
  return device_call(op, "fmmh_reset", [&] {
    op->op->reset();
  });
}

#define FMMH_LEVEL_CALL(NAME)                                                      \
  int32_t fmmh_##NAME(fmmh_operator *op, int32_t level) {                          \
    if (op != nullptr && (level < 0 || level >= op->nlevels))                      \
      return guard([&] { need(false, "fmmh_" #NAME ": level out of range"); });    \
    return device_call(op, "fmmh_" #NAME, [&] { op->op->NAME(level); });          \
  }
FMMH_LEVEL_CALL(p2m)
FMMH_LEVEL_CALL(m2m)
FMMH_LEVEL_CALL(m2l)
FMMH_LEVEL_CALL(p2l)
FMMH_LEVEL_CALL(l2l)
FMMH_LEVEL_CALL(l2p)
FMMH_LEVEL_CALL(m2p)
FMMH_LEVEL_CALL(p2p)
#undef FMMH_LEVEL_CALL

int32_t fmmh_evaluate(fmmh_operator *op) {
  return device_call(op, "fmmh_evaluate", [&] {
    op->op->evaluate();
  });
}

int32_t fmmh_evaluate_graph(fmmh_operator *op) {
  return device_call(op, "fmmh_evaluate_graph", [&] {
    op->op->evaluate_graph();
  });
}

int32_t fmmh_read_output(fmmh_operator *op, double *phi, double *grad, int64_t count, int32_t on_device) {
  return device_call(op, "fmmh_read_output", [&] {
    need(phi != nullptr, "fmmh_read_output: NULL argument");
    need(count == op->op->ntargets(), "fmmh_read_output: count differs from the plan's target count");
    op->op->read_output(phi, grad, on_device != 0);
  });
}

int32_t fmmh_vjp(fmmh_operator *op, const double *charges, const double *cotangent, const double *field,
                 int64_t count, int32_t on_device, double *grad_charges, double *grad_points) {
  return device_call(op, "fmmh_vjp", [&] {
    need(charges && cotangent && grad_charges, "fmmh_vjp: NULL argument");
    need(count == op->op->nsources() && count == op->op->ntargets(),
         "fmmh_vjp: count differs from the plan's point count");
    need(grad_points == nullptr || field != nullptr, "fmmh_vjp: grad_points needs the forward field");
    op->op->vjp(charges, cotangent, field, on_device != 0, grad_charges, grad_points);
  });
}

int32_t fmmh_direct_sum(int32_t device, const double *sources, const double *charges, int64_t nsources,
                        const double *targets, int64_t ntargets, double *phi, double *grad, int32_t on_device) {
  return guard([&] {
    need(nsources == 0 || (sources && charges), "fmmh_direct_sum: NULL sources");
    need(ntargets == 0 || (targets && phi), "fmmh_direct_sum: NULL targets or output");
    fmmh::direct_sum(device, sources, charges, nsources, targets, ntargets, phi, grad, on_device != 0);
  });
}

int32_t fmmh_direct_sum_noalloc(int32_t device, const double *sources, const double *charges, int64_t nsources,
                                const double *targets, int64_t ntargets, double *phi, double *grad, void *stream) {
  return guard([&] {
    need(nsources == 0 || (sources && charges), "fmmh_direct_sum_noalloc: NULL sources");
    need(ntargets == 0 || (targets && phi), "fmmh_direct_sum_noalloc: NULL targets or output");
    fmmh::direct_sum_noalloc(device, sources, charges, nsources, targets, ntargets, phi, grad,
                             static_cast<hipStream_t>(stream));
  });
}

int32_t fmmh_stream_wait(fmmh_operator *op, void *stream) {
  return device_call(op, "fmmh_stream_wait", [&] {
    op->op->wait_for(static_cast<hipStream_t>(stream));
  });
}

int32_t fmmh_sync(fmmh_operator *op) {
  return device_call(op, "fmmh_sync", [&] {
    op->op->sync();
  });
}

int32_t fmmh_read_expansions(fmmh_operator *op, int32_t locals, int32_t level, double *out) {
  return device_call(op, "fmmh_read_expansions", [&] {
    need(out != nullptr, "fmmh_read_expansions: NULL argument");
    need(level >= 0 && level < op->nlevels, "fmmh_read_expansions: level out of range");
    op->op->read_expansions(locals != 0, level, out);
  });
}

int32_t fmmh_inject_error(fmmh_operator *op) {
  return guard([&] {
    need(op != nullptr, "fmmh_inject_error: NULL operator");
    op->inject = true;
  });
}

int32_t fmmh_kind_timing(fmmh_operator *op, int32_t enable) {
  return guard([&] {
    need(op != nullptr, "fmmh_kind_timing: NULL operator");
    op->op->set_timing(enable != 0);
  });
}

int32_t fmmh_kind_times(fmmh_operator *op, double *ms, int64_t *launches) {
  return device_call(op, "fmmh_kind_times", [&] {
    op->op->kind_times(ms, launches);
  });
}

int32_t fmmh_helmholtz_degree(double k_re, double radius, double eps) {
  if (!(eps > 0.0 && eps < 1.0) || !(radius > 0.0) || !(k_re >= 0.0)) return -1;
  const double L = std::log10(1.0 / eps), x = k_re * std::sqrt(3.0) * radius;
  const double low = std::ceil(0.30 * L * L + 0.25 * L + 2.60);
  const double high = std::ceil(x + 0.80 * std::pow(std::log(1.0 / eps), 2.0 / 3.0) * std::cbrt(x));
  return static_cast<int32_t>(low > high ? low : high);
}

int32_t fmmh_helmholtz_create(const fmmh_plan *plan, const fmmh_helmholtz_options *options, fmmh_helmholtz **out) {
  return guard([&] {
    need(plan && options && out, "fmmh_helmholtz_create: NULL argument");
    fmmh::HelmholtzOptions o;
    o.p = options->p;
    o.kre = options->k_re;
    o.kim = options->k_im;
    o.precision = options->precision;
    o.gradients = options->gradients;
    o.device = options->device;
    o.m2l = options->m2l;
    o.m2l_dense_budget = options->m2l_dense_budget;
    auto h = std::make_unique<fmmh_helmholtz>();
    h->op = std::make_unique<fmmh::HelmholtzOperator>(plan->plan, o);
    h->nlevels = plan->plan.nlevels;
    h->nleaves = plan->plan.nleaves();
    for (int l = 0; l < plan->plan.nlevels; ++l) h->nboxes += plan->plan.nboxes(l);
    *out = h.release();
  });
}

void fmmh_helmholtz_free(fmmh_helmholtz *op) { delete op; }

int32_t fmmh_helmholtz_stats(const fmmh_helmholtz *op, fmmh_stats *out) {
  return guard([&] {
    need(op && out, "fmmh_helmholtz_stats: NULL argument");
    out->nlevels = op->op->nlevels();
    out->nleaves = op->nleaves;
    out->nboxes = op->nboxes;
    out->p2p_pairs = op->op->p2p_pairs();
    out->m2l_pairs = op->op->m2l_pairs();
    out->m2l_slots = 0;
    out->device_bytes = op->op->device_bytes();
  });
}

int32_t fmmh_helmholtz_load_charges(fmmh_helmholtz *op, const double *charges, int64_t count, int32_t on_device) {
  return guard([&] {
    need(op && charges, "fmmh_helmholtz_load_charges: NULL argument");
    need(count == op->op->nsources(), "fmmh_helmholtz_load_charges: count differs from the plan's source count");
    op->op->load_charges(charges, on_device != 0);
  });
}

int32_t fmmh_helmholtz_stream_wait(fmmh_helmholtz *op, void *stream) {
  return guard([&] {
    need(op != nullptr, "fmmh_helmholtz_stream_wait: NULL operator");
    op->op->wait_for(static_cast<hipStream_t>(stream));
  });
}

int32_t fmmh_helmholtz_reset(fmmh_helmholtz *op) {
  return guard([&] {
    need(op != nullptr, "fmmh_helmholtz_reset: NULL operator");
    op->op->reset();
  });
}

int32_t fmmh_helmholtz_evaluate(fmmh_helmholtz *op) {
  return guard([&] {
    need(op != nullptr, "fmmh_helmholtz_evaluate: NULL operator");
    op->op->evaluate();
  });
}

int32_t fmmh_helmholtz_read_output(fmmh_helmholtz *op, double *phi, double *grad, int64_t count,
                                   int32_t on_device) {
  return guard([&] {
    need(op && phi, "fmmh_helmholtz_read_output: NULL argument");
    need(count == op->op->ntargets(), "fmmh_helmholtz_read_output: count differs from the plan's target count");
    op->op->read_output(phi, grad, on_device != 0);
  });
}

int32_t fmmh_helmholtz_read_expansions(fmmh_helmholtz *op, int32_t locals, int32_t level, double *out) {
  return guard([&] {
    need(op && out, "fmmh_helmholtz_read_expansions: NULL argument");
    need(level >= 0 && level < op->nlevels, "fmmh_helmholtz_read_expansions: level out of range");
    op->op->read_expansions(locals != 0, level, out);
  });
}

int32_t fmmh_helmholtz_kind_timing(fmmh_helmholtz *op, int32_t enable) {
  return guard([&] {
    need(op != nullptr, "fmmh_helmholtz_kind_timing: NULL operator");
    op->op->set_timing(enable != 0);
  });
}

int32_t fmmh_helmholtz_kind_times(fmmh_helmholtz *op, double *ms, int64_t *launches) {
  return guard([&] {
    need(op != nullptr, "fmmh_helmholtz_kind_times: NULL operator");
    op->op->kind_times(ms, launches);
  });
}

int32_t fmmh_helmholtz_table(int32_t p, double kh_re, double kh_im, int32_t kind, int32_t index, double *out) {
  return guard([&] {
    need(out != nullptr, "fmmh_helmholtz_table: NULL out");
    need(p >= 0 && p <= 30, "fmmh_helmholtz_table: p must lie in 0..30");
    need(kh_re != 0.0 || kh_im != 0.0, "fmmh_helmholtz_table: kh must be nonzero");
    const fmmh::cplx kh(kh_re, kh_im);
    std::vector<double> t;
    const size_t sz = 2 * size_t(fmmh::ncoef(p)) * fmmh::ncoef(p);
    if (kind == 0 || kind == 1) {
      need(index >= 0 && index < 8, "fmmh_helmholtz_table: octant must lie in 0..7");
      if (kind == 0)
        fmmh::h_m2m_tables(p, kh, t);
      else
        fmmh::h_l2l_tables(p, kh, t);
      std::memcpy(out, t.data() + size_t(index) * sz, sz * sizeof(double));
    } else if (kind == 2) {
      need(index >= 0 && index < 316, "fmmh_helmholtz_table: offset index 0..315");
      fmmh::h_m2l_tables(p, kh, {uint16_t(index)}, t);
      std::memcpy(out, t.data(), sz * sizeof(double));
    } else {
      // the M2L composed from the per-pair pieces the device applies: O^T C O
      need(kind == 3 && index >= 0 && index < 316, "fmmh_helmholtz_table: kind 0..3, offset index 0..315");
      need(p <= 15, "fmmh_helmholtz_table: the per-pair pieces support p <= 15");
      int d[3];
      fmmh::offset_of_index(index, d);
      std::vector<double> fwd, coax;
      fmmh::h_rcr_rotations(p, fwd);
      fmmh::h_rcr_coaxial(p, kh, {2.0 * std::sqrt(double(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))}, coax);
      const int N = fmmh::ncoef(p), rb = fmmh::h_rcr_rotation_size(p), cb = fmmh::h_rcr_coaxial_size(p);
      const double *F = fwd.data() + size_t(index) * rb;
      auto bo = [](int n) { return n * (2 * n - 1) * (2 * n + 1) / 3; };
      auto co = [p](int m) {
        int c = 0;
        for (int q = 0; q < m; ++q) c += (p + 1 - q) * (p + 1 - q);
        return c;
      };
      std::vector<double> zr(N), zi(N);
      for (int k = 0; k < N; ++k) {
        const int nk = int(std::sqrt(double(k))), ck = k - nk * nk, wk = 2 * nk + 1;
        std::fill(zr.begin(), zr.end(), 0.0);
        std::fill(zi.begin(), zi.end(), 0.0);
        for (int np = 0; np <= p; ++np)
          for (int m = -std::min(np, nk); m <= std::min(np, nk); ++m) {
            const int am = m < 0 ? -m : m, w = p + 1 - am;
            const double x = F[bo(nk) + (m + nk) + ck * wk];
            const size_t a = size_t(co(am)) + (np - am) + size_t(nk - am) * w;
            zr[np * np + np + m] = coax[a] * x;
            zi[np * np + np + m] = coax[cb + a] * x;
          }
        for (int np = 0; np <= p; ++np) {
          const int w = 2 * np + 1;
          for (int c = 0; c < w; ++c) {
            double vr = 0.0, vi = 0.0;
            for (int r = 0; r < w; ++r) {
              const double o = F[bo(np) + r + c * w];
              vr += o * zr[np * np + r];
              vi += o * zi[np * np + r];
            }
            out[size_t(k) * N + np * np + c] = vr;
            out[size_t(N) * N + size_t(k) * N + np * np + c] = vi;
          }
        }
      }
    }
  });
}

int32_t fmmh_helmholtz_direct_sum(int32_t device, double k_re, double k_im, const double *sources,
                                  const double *charges, int64_t nsources, const double *targets, int64_t ntargets,
                                  double *phi, double *grad, int32_t on_device) {
  return guard([&] {
    need(nsources == 0 || (sources && charges), "fmmh_helmholtz_direct_sum: NULL sources");
    need(ntargets == 0 || (targets && phi), "fmmh_helmholtz_direct_sum: NULL targets or output");
    fmmh::helmholtz_direct_sum(device, k_re, k_im, sources, charges, nsources, targets, ntargets, phi, grad,
                               on_device != 0);
  });
}

int32_t fmmh_plane_wave_quadrature(int32_t p, int32_t nodes, int32_t *out_nodes, int32_t *half_nodes, double *lambda,
                                   double *weight, int32_t *nalpha, int32_t *fourier, double *max_error) {
  return guard([&] {
    need(out_nodes != nullptr, "fmmh_plane_wave_quadrature: NULL nodes");
    const fmmh::PlaneWave pw = fmmh::build_plane_wave(p, nodes);
    *out_nodes = pw.nodes;
    if (half_nodes) *half_nodes = pw.H;
    if (max_error) *max_error = pw.max_error;
    for (int k = 0; k < pw.nodes; ++k) {
      if (lambda) lambda[k] = pw.lambda[k];
      if (weight) weight[k] = pw.weight[k];
      if (nalpha) nalpha[k] = pw.fmm3d_alpha[k];
      if (fourier) fourier[k] = pw.fourier[k];
    }
  });
}

// The last quadrature built by the host restatements, kept for the next call of the same
// (p, nodes).
static const fmmh::PlaneWave &cached_plane_wave(int32_t p, int32_t nodes, std::unique_lock<std::mutex> &lock) {
  static std::mutex mu;
  static std::unique_ptr<fmmh::PlaneWave> last;
  lock = std::unique_lock<std::mutex>(mu);
  if (!last || last->p != p || last->nodes != (nodes == 0 ? fmmh::plane_wave_nodes_for_degree(p) : nodes))
    last = std::make_unique<fmmh::PlaneWave>(fmmh::build_plane_wave(p, nodes));
  return *last;
}

int32_t fmmh_plane_wave_m2l(int32_t p, int32_t nodes, int32_t offset, int32_t direction, const double *in,
                            double *out) {
  return guard([&] {
    need(in && out, "fmmh_plane_wave_m2l: NULL argument");
    need(offset >= 0 && offset < 316, "fmmh_plane_wave_m2l: offset index 0..315");
    need(direction >= -1 && direction < 6, "fmmh_plane_wave_m2l: direction -1..5");
    std::unique_lock<std::mutex> lock;
    const fmmh::PlaneWave &pw = cached_plane_wave(p, nodes, lock);
    int d[3];
    fmmh::offset_of_index(offset, d);
    fmmh::plane_wave_m2l_host(pw, d, direction, in, out);
  });
}

int32_t fmmh_plane_wave_merge(int32_t p, int32_t nodes, const double *slot_multipoles, const uint8_t *present,
                              double *out) {
  return guard([&] {
    need(slot_multipoles && present && out, "fmmh_plane_wave_merge: NULL argument");
    std::unique_lock<std::mutex> lock;
    const fmmh::PlaneWave &pw = cached_plane_wave(p, nodes, lock);
    const size_t nc = size_t(p + 1) * (p + 1);
    std::vector<const double *> slots(fmmh::kPwSlots, nullptr);
    for (int a = 0; a < fmmh::kPwSlots; ++a)
      if (present[a]) slots[a] = slot_multipoles + a * nc;
    fmmh::plane_wave_merge_host(pw, slots.data(), out);
  });
}

int32_t fmmh_plane_wave_slot_direction(int32_t slot, int32_t child) {
  if (slot < 0 || slot >= fmmh::kPwSlots || child < 0 || child >= fmmh::kPwChildren) return -2;
  return fmmh::plane_wave_slot_direction(slot, child);
}

}  // extern "C"
