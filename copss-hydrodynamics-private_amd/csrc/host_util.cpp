// host_util.cpp -- host-side input utilities of the C-ABI (no GPU needed):
// the reference's Brownian noise generator, a seeded random-sequential-
// addition bead generator, N(0,1) forces, and the HEX20 IF table.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "copss_ggem.h"

extern "C" {

// RandomGenerator::random_vector_normal (random_generator.C:43-67) with the
// engine std::default_random_engine seeded by set_seed (:96-99); the
// fixman_integrate call draws normal(0, sqrt(dt)) over 3N values and scales
// the vector by sqrt(2) (copss.C:1565-1572).  Each call of
// random_vector_normal builds a fresh distribution on the same engine, so a
// restart discards whole vectors (copss.C:1544-1555).
int ggem_util_reference_noise(int64_t n, uint64_t seed, double dt, int64_t skip_vectors,
                              double* out) {
  if (n < 0 || !out || !(dt > 0) || skip_vectors < 0) return GGEM_ERR_INVALID_ARG;
  std::default_random_engine gen;
  gen.seed(static_cast<std::default_random_engine::result_type>(seed));
  const double std_dev = std::sqrt(dt);
  for (int64_t v = 0; v < skip_vectors; ++v) {
    std::normal_distribution<double> d(0.0, 1.0);
    d.reset();
    for (int64_t i = 0; i < n; ++i) (void)d(gen);
  }
  std::normal_distribution<double> dist(0.0, std_dev);
  dist.reset();
  for (int64_t i = 0; i < n; ++i) out[i] = dist(gen);
  const double s2 = std::sqrt(2.0);
  for (int64_t i = 0; i < n; ++i) out[i] *= s2;
  return GGEM_OK;
}

int ggem_util_normal(int64_t n, uint64_t seed, double* out) {
  if (n < 0 || !out) return GGEM_ERR_INVALID_ARG;
  std::mt19937_64 gen(seed);
  std::normal_distribution<double> dist(0.0, 1.0);
  for (int64_t i = 0; i < n; ++i) out[i] = dist(gen);
  return GGEM_OK;
}

// Random sequential addition with a hash grid of cell size >= min_dist.
int ggem_util_rsa_positions(const ggem_params* p, int64_t n, double min_dist, uint64_t seed,
                            double* xyz) {
  if (!p || n < 0 || !xyz || !(min_dist >= 0)) return GGEM_ERR_INVALID_ARG;
  double L[3];
  int64_t g[3];
  for (int d = 0; d < 3; ++d) {
    L[d] = p->box_max[d] - p->box_min[d];
    if (!(L[d] > 0)) return GGEM_ERR_INVALID_ARG;
    g[d] = min_dist > 0 ? int64_t(std::floor(L[d] / min_dist)) : 1;
    if (g[d] < 1) g[d] = 1;
    if (g[d] > 4096) g[d] = 4096;
  }
  const int64_t ncell = g[0] * g[1] * g[2];
  std::vector<int64_t> head(ncell, -1), next(n, -1);
  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<double> ux(p->box_min[0], p->box_max[0]);
  std::uniform_real_distribution<double> uy(p->box_min[1], p->box_max[1]);
  std::uniform_real_distribution<double> uz(p->box_min[2], p->box_max[2]);
  const double md2 = min_dist * min_dist;
  const int64_t max_tries = 1000 * (n + 1);
  int64_t tries = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (;;) {
      if (++tries > max_tries) return GGEM_ERR_CAPACITY;
      double x[3];
      x[0] = ux(gen);
      x[1] = uy(gen);
      x[2] = uz(gen);
      int64_t c[3];
      for (int d = 0; d < 3; ++d) {
        c[d] = int64_t((x[d] - p->box_min[d]) / L[d] * double(g[d]));
        if (c[d] >= g[d]) c[d] = g[d] - 1;
        if (c[d] < 0) c[d] = 0;
      }
      bool ok = true;
      for (int dk = -1; dk <= 1 && ok; ++dk)
        for (int dj = -1; dj <= 1 && ok; ++dj)
          for (int di = -1; di <= 1 && ok; ++di) {
            int64_t q[3] = {c[0] + di, c[1] + dj, c[2] + dk};
            bool skip = false;
            for (int d = 0; d < 3; ++d) {
              if (q[d] < 0 || q[d] >= g[d]) {
                if (p->periodic[d]) q[d] = (q[d] + g[d]) % g[d];
                else skip = true;
              }
            }
            if (skip) continue;
            for (int64_t j = head[q[0] + g[0] * (q[1] + g[1] * q[2])]; j >= 0; j = next[j]) {
              double r2 = 0;
              for (int d = 0; d < 3; ++d) {
                double dd = xyz[3 * j + d] - x[d];
                if (p->periodic[d]) {
                  if (dd <= -L[d] / 2) dd += L[d];
                  if (dd > L[d] / 2) dd -= L[d];
                }
                r2 += dd * dd;
              }
              if (r2 < md2) {
                ok = false;
                break;
              }
            }
          }
      if (!ok) continue;
      for (int d = 0; d < 3; ++d) xyz[3 * i + d] = x[d];
      const int64_t cc = c[0] + g[0] * (c[1] + g[1] * c[2]);
      next[i] = head[cc];
      head[cc] = i;
      break;
    }
  }
  return GGEM_OK;
}

// HEX20 serendipity shape functions on [-1,1]^3 in libMesh's node order
// (8 vertices, then the 12 edge mid-nodes 0-1, 1-2, 2-3, 3-0, 0-4, 1-5, 2-6,
// 3-7, 4-5, 5-6, 6-7, 7-4); Gauss-Legendre points x fastest; an
// axis-aligned element of size h has JxW = w_a w_b w_c h_x h_y h_z / 8.
int ggem_util_hex20_if_table(const double* h, int32_t nqp_1d, double* IF) {
  if (!h || !IF || (nqp_1d != 2 && nqp_1d != 3)) return GGEM_ERR_INVALID_ARG;
  static const int V[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  static const int E[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                               {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
  double node[20][3];
  for (int v = 0; v < 8; ++v)
    for (int d = 0; d < 3; ++d) node[v][d] = V[v][d];
  for (int e = 0; e < 12; ++e)
    for (int d = 0; d < 3; ++d) node[8 + e][d] = 0.5 * (V[E[e][0]][d] + V[E[e][1]][d]);
  double gp[3], gw[3];
  if (nqp_1d == 2) {
    gp[0] = -std::sqrt(1. / 3.);
    gp[1] = std::sqrt(1. / 3.);
    gw[0] = gw[1] = 1.0;
  } else {
    gp[0] = -std::sqrt(3. / 5.);
    gp[1] = 0.0;
    gp[2] = std::sqrt(3. / 5.);
    gw[0] = gw[2] = 5. / 9.;
    gw[1] = 8. / 9.;
  }
  const int nqp = nqp_1d * nqp_1d * nqp_1d;
  const double detJ = (0.5 * h[0]) * (0.5 * h[1]) * (0.5 * h[2]);
  int qp = 0;
  for (int c = 0; c < nqp_1d; ++c)
    for (int b = 0; b < nqp_1d; ++b)
      for (int a = 0; a < nqp_1d; ++a) {
        const double xi = gp[a], eta = gp[b], zeta = gp[c];
        const double jxw = gw[a] * gw[b] * gw[c] * detJ;
        for (int l = 0; l < 20; ++l) {
          const double xl = node[l][0], yl = node[l][1], zl = node[l][2];
          double phi;
          if (l < 8) {
            phi = 0.125 * (1 + xi * xl) * (1 + eta * yl) * (1 + zeta * zl) *
                  (xi * xl + eta * yl + zeta * zl - 2);
          } else if (xl == 0) {
            phi = 0.25 * (1 - xi * xi) * (1 + eta * yl) * (1 + zeta * zl);
          } else if (yl == 0) {
            phi = 0.25 * (1 + xi * xl) * (1 - eta * eta) * (1 + zeta * zl);
          } else {
            phi = 0.25 * (1 + xi * xl) * (1 + eta * yl) * (1 - zeta * zeta);
          }
          IF[l * nqp + qp] = jxw * phi;
        }
        ++qp;
      }
  return GGEM_OK;
}

}  // extern "C"
