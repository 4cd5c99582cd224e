/*
 * rng.c -- oracle (test infrastructure): SplitMix64 and the seeded point
 * distributions of fmm-validate/src/rng.rs and fmm-validate/src/points.rs.
 */
#include <math.h>
#include "oracle.h"

/* SplitMix64 step, fmm-validate/src/rng.rs:29-35. */
uint64_t orc_next_u64(uint64_t *state) {
  *state += 0x9e3779b97f4a7c15ULL;
  uint64_t z = *state;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* Top 53 bits scaled to [0, 1), fmm-validate/src/rng.rs:38-40. */
double orc_uniform(uint64_t *state) {
  return (double)(orc_next_u64(state) >> 11) * (1.0 / 9007199254740992.0);
}

/* lo + (hi - lo) u, fmm-validate/src/rng.rs:43-45. */
double orc_range(uint64_t *state, double lo, double hi) {
  return lo + (hi - lo) * orc_uniform(state);
}

/* Uniform in 0..n from the top 32 bits, fmm-validate/src/fmm_accuracy.rs:384-386. */
int64_t orc_below(uint64_t *state, int64_t n) {
  return (int64_t)(((orc_next_u64(state) >> 32) * (uint64_t)n) >> 32);
}

/* Box-Muller, cosine branch, fmm-validate/src/points.rs:142-147. */
double orc_normal(uint64_t *state) {
  double radius = sqrt(-2.0 * log(1.0 - orc_uniform(state)));
  return radius * cos(2.0 * 3.141592653589793 * orc_uniform(state));
}

/* points::cube, fmm-validate/src/points.rs:27-31: x, y, z drawn in turn. */
void orc_cube(uint64_t *state, int64_t n, const double centre[3], double half, double *out) {
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) out[3 * i + k] = centre[k] + half * orc_range(state, -1.0, 1.0);
}

/* in_unit_ball, fmm-validate/src/points.rs:150-158. */
static void in_unit_ball(uint64_t *state, double min_norm, double v[3]) {
  for (;;) {
    for (int k = 0; k < 3; ++k) v[k] = orc_range(state, -1.0, 1.0);
    double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (s <= 1.0 && s >= min_norm * min_norm) return;
  }
}

/* points::ball, fmm-validate/src/points.rs:35-43. */
void orc_ball(uint64_t *state, int64_t n, const double centre[3], double radius, double *out) {
  double v[3];
  for (int64_t i = 0; i < n; ++i) {
    in_unit_ball(state, 0.0, v);
    for (int k = 0; k < 3; ++k) out[3 * i + k] = centre[k] + radius * v[k];
  }
}

/* points::sphere, fmm-validate/src/points.rs:48-57. */
void orc_sphere(uint64_t *state, int64_t n, const double centre[3], double radius, double *out) {
  double v[3];
  for (int64_t i = 0; i < n; ++i) {
    in_unit_ball(state, 1e-3, v);
    double norm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) out[3 * i + k] = centre[k] + radius * (v[k] / norm);
  }
}

/* points::plummer, fmm-validate/src/points.rs:81-98, truncated at 10 a. */
void orc_plummer(uint64_t *state, int64_t n, const double centre[3], double scale, double *out) {
  const double x = 10.0;
  const double mass = x * x * x / pow(1.0 + x * x, 1.5);
  double v[3];
  for (int64_t i = 0; i < n; ++i) {
    double u = mass * orc_uniform(state);
    double r = u > 0.0 ? scale / sqrt(pow(u, -2.0 / 3.0) - 1.0) : 0.0;
    in_unit_ball(state, 1e-3, v);
    double norm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) out[3 * i + k] = centre[k] + r * (v[k] / norm);
  }
}

/* points::gaussian_clusters, fmm-validate/src/points.rs:114-135 (truncated at 4 sigma). */
void orc_gaussian_clusters(uint64_t *state, int64_t n, const double *centres, int64_t k,
                           double width, double *out) {
  int64_t kk = k > 0 ? k : 1, at = 0;
  for (int64_t j = 0; j < k; ++j) {
    int64_t count = n / kk + (j < n % kk ? 1 : 0);
    for (int64_t c = 0; c < count; ++c) {
      double z[3];
      for (;;) {
        for (int a = 0; a < 3; ++a) z[a] = orc_normal(state);
        if (z[0] * z[0] + z[1] * z[1] + z[2] * z[2] <= 16.0) break;
      }
      for (int a = 0; a < 3; ++a) out[3 * at + a] = centres[3 * j + a] + width * z[a];
      ++at;
    }
  }
}

/* points::charges, fmm-validate/src/points.rs:137-139: uniform in [-1, 1). */
void orc_charges(uint64_t *state, int64_t n, double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc_range(state, -1.0, 1.0);
}

void orc_normals(uint64_t *state, int64_t n, double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc_normal(state);
}

/* Partial Fisher-Yates shuffle of 0..n, fmm-bench/src/measure.rs:52-57. */
void orc_sample(uint64_t *state, int64_t n, int64_t sampled, int64_t *out) {
  /* out has room for n entries; the first `sampled` are the sample. */
  for (int64_t i = 0; i < n; ++i) out[i] = i;
  if (sampled > n) sampled = n;
  for (int64_t i = 0; i < sampled; ++i) {
    int64_t j = i + orc_below(state, n - i);
    int64_t t = out[i];
    out[i] = out[j];
    out[j] = t;
  }
}
