// host_util.cpp -- host-side input utilities of the C-ABI (no GPU needed):
// the reference's Brownian noise generator, a seeded random-sequential-
// addition bead generator, N(0,1) forces, and the HEX20 IF table.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
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

// The same generator as a stream that persists across calls, as the
// reference keeps one RandomGenerator for the whole run
// (brownian_system.h:292, copss.C:1389): each draw continues the engine.
struct ggem_noise_stream {
  std::default_random_engine gen;
};

int ggem_util_noise_create(uint64_t seed, ggem_noise_stream** out) {
  if (!out) return GGEM_ERR_INVALID_ARG;
  *out = new (std::nothrow) ggem_noise_stream;
  if (!*out) return GGEM_ERR_CAPACITY;
  (*out)->gen.seed(static_cast<std::default_random_engine::result_type>(seed));
  return GGEM_OK;
}

int ggem_util_noise_destroy(ggem_noise_stream* s) {
  delete s;
  return GGEM_OK;
}

int ggem_util_noise_draw(ggem_noise_stream* s, int64_t n, double dt, double* out) {
  if (!s || n < 0 || (n > 0 && !out) || !(dt > 0)) return GGEM_ERR_INVALID_ARG;
  std::normal_distribution<double> dist(0.0, std::sqrt(dt));
  dist.reset();
  const double s2 = std::sqrt(2.0);
  for (int64_t i = 0; i < n; ++i) out[i] = dist(s->gen) * s2;
  return GGEM_OK;
}

int ggem_util_noise_skip(ggem_noise_stream* s, int64_t n, int64_t n_vectors) {
  if (!s || n < 0 || n_vectors < 0) return GGEM_ERR_INVALID_ARG;
  for (int64_t v = 0; v < n_vectors; ++v) {
    std::normal_distribution<double> d(0.0, 1.0);
    d.reset();
    for (int64_t i = 0; i < n; ++i) (void)d(s->gen);
  }
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

// ---- FP32C pair coefficients as a piecewise polynomial in s = r^2 ----
//
// G_reg(x) = A(s) I + Bx(s) x x^T with A = eps (c1 + c2), Bx = eps (c1 - c2)/s,
// c1 = (erf(xi r) - erf(alpha r))/r, c2 = 2xi/sqrt(pi) e^{-xi^2 s} -
// 2alpha/sqrt(pi) e^{-alpha^2 s} (ggem_stokes.C:304-357).  Both are entire
// functions of s.  With k = max(alpha^2, xi^2) and m = min(alpha^2, xi^2),
// [0, s1) is cut into intervals of width d1 = 2^floor(log2(1.4/k)) up to
// s1 >= 36/k, where every term with the faster decay is below e^-36 of its
// peak, and [s1, s_cut) into intervals of width d2 = 2^floor(log2(1.4/m))
// up to s_cut >= 36/m; beyond, A = Bx = 0.  On each interval both
// functions are interpolated at the 8 Chebyshev points and stored as
// monomials in y = 2 (s - mid)/d in [-1, 1]: the interpolation error is at
// most 2 (d k / 4)^8 / 8! <= 1.1e-8 of the function's size on the interval
// (its 8th derivative is at most k^8 times that size), below FP32
// rounding.  Widths are powers of two, so the kernels form the interval
// and y exactly.
namespace {

const double kSqrtPi = 1.77245385090551602729816748334;

// (A, Bx) in FP64: a power series near s = 0, the closed form elsewhere
void coef_f64(double eps, double alpha, double ksi, double s, double* A, double* Bx) {
  const double a2 = alpha * alpha, k2 = ksi * ksi;
  if (s * std::max(a2, k2) < 0.25) {
    // c1 = 2/sqrt(pi) sum (-1)^n (xi^{2n+1} - alpha^{2n+1}) s^n / (n! (2n+1)),
    // c2 = 2/sqrt(pi) sum (-1)^n (xi^{2n+1} - alpha^{2n+1}) s^n / n!
    // so c1 + c2 = 2/sqrt(pi) sum (-1)^n D_n s^n/n! (1 + 1/(2n+1)) and
    // (c1 - c2)/s = 2/sqrt(pi) sum_{n>=1} (-1)^{n+1} D_n s^{n-1}/n! 2n/(2n+1),
    // D_n = xi^{2n+1} - alpha^{2n+1}
    double sa = 0.0, sb = 0.0, pk = ksi, pa = alpha, sn = 1.0, fact = 1.0;
    for (int n = 0; n < 30; ++n) {
      const double sign = n % 2 ? -1.0 : 1.0;
      sa += sign * (pk - pa) * sn / fact * (1.0 + 1.0 / (2 * n + 1));
      pk *= k2;
      pa *= a2;
      fact *= (n + 1);
      // term n + 1 of (c1 - c2)/s: s^n / (n+1)!
      sb -= sign * (pk - pa) * sn / fact * (2.0 * (n + 1) / (2 * n + 3));
      sn *= s;
    }
    *A = eps * 2.0 / kSqrtPi * sa;
    *Bx = -eps * 2.0 / kSqrtPi * sb;
    return;
  }
  const double r = std::sqrt(s);
  const double c1 = (std::erfc(alpha * r) - std::erfc(ksi * r)) / r;
  const double c2 = 2.0 * ksi / kSqrtPi * std::exp(-k2 * s) - 2.0 * alpha / kSqrtPi * std::exp(-a2 * s);
  *A = eps * (c1 + c2);
  *Bx = eps * (c1 - c2) / s;
}

double pow2_floor(double v) { return std::ldexp(1.0, int(std::floor(std::log2(v)))); }

// monomial coefficients in y of sum_k c_k T_k(y), degree 7
void cheb_to_mono(const double* c, double* out) {
  double T[8][8] = {};
  T[0][0] = 1.0;
  T[1][1] = 1.0;
  for (int k = 2; k < 8; ++k)
    for (int j = 0; j < 8; ++j) T[k][j] = (j ? 2.0 * T[k - 1][j - 1] : 0.0) - T[k - 2][j];
  for (int j = 0; j < 8; ++j) {
    out[j] = 0.0;
    for (int k = 0; k < 8; ++k) out[j] += c[k] * T[k][j];
  }
}

}  // namespace

extern "C" int ggem_util_fp32c_table(double eps, double alpha, double ksi, int32_t cap, float* table,
                                     float* meta, int32_t* n_intervals) {
  if (!(eps > 0) || !(alpha > 0) || !(ksi > 0) || !meta || !n_intervals) return GGEM_ERR_INVALID_ARG;
  const double k = std::max(alpha * alpha, ksi * ksi), m = std::min(alpha * alpha, ksi * ksi);
  const double d1 = pow2_floor(1.4 / k), d2 = pow2_floor(1.4 / m);
  // zero from s m >= 20 on: both coefficients are below e^-20 of their size
  // there (erfc and the Gaussians decay as e^{-m s}), far below the FP32
  // rounding of the pairs that remain
  const double s_zero = 20.0 / m;
  const int n1 = int(std::ceil(std::min(36.0 / k, s_zero) / d1));
  const double s1 = n1 * d1;
  const int n2 = std::max(0, int(std::ceil((s_zero - s1) / d2)));
  const int nint = n1 + n2 + 1;  // the last interval is zero
  *n_intervals = nint;
  meta[0] = float(s1);
  meta[1] = float(1.0 / d1);
  meta[2] = float(1.0 / d2);
  meta[3] = float(d1);
  meta[4] = float(d2);
  meta[5] = float(n1);
  meta[6] = float(n2);
  meta[7] = float(s1 + n2 * d2);  // start of the zero interval
  if (!table) return GGEM_OK;  // size query
  if (cap < nint * 16) return GGEM_ERR_INVALID_ARG;
  const double PI = 3.14159265358979323846;
  for (int i = 0; i < nint; ++i) {
    float* o = table + 16 * i;
    if (i == nint - 1) {
      for (int j = 0; j < 16; ++j) o[j] = 0.f;
      break;
    }
    const double lo = i < n1 ? i * d1 : s1 + (i - n1) * d2;
    const double d = i < n1 ? d1 : d2;
    double fa[8], fb[8];
    for (int j = 0; j < 8; ++j) {
      const double y = std::cos((2 * j + 1) * PI / 16.0);
      coef_f64(eps, alpha, ksi, lo + 0.5 * d * (y + 1.0), &fa[j], &fb[j]);
    }
    double ca[8], cb[8];
    for (int kk = 0; kk < 8; ++kk) {
      double sa = 0.0, sb = 0.0;
      for (int j = 0; j < 8; ++j) {
        const double t = std::cos(kk * (2 * j + 1) * PI / 16.0);
        sa += fa[j] * t;
        sb += fb[j] * t;
      }
      ca[kk] = (kk ? 2.0 : 1.0) / 8.0 * sa;
      cb[kk] = (kk ? 2.0 : 1.0) / 8.0 * sb;
    }
    double ma[8], mb[8];
    cheb_to_mono(ca, ma);
    cheb_to_mono(cb, mb);
    for (int j = 0; j < 8; ++j) {
      o[j] = float(ma[j]);
      o[8 + j] = float(mb[j]);
    }
  }
  return GGEM_OK;
}
