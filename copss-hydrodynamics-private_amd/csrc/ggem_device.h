// ggem_device.h -- device-side building blocks shared by the kernels:
// minimum-image displacement, the element-stencil walk over the cell-sorted
// particle store, and the regularized-Stokeslet pair term in FP64 and in
// FP32 with compensated accumulation.
#pragma once

#include <hip/hip_runtime.h>

#include "ggem_internal.h"

namespace ggem {

// PMPeriodicBoundary::point_vector (pm_periodic_boundary.C:350-374), one axis.
__device__ __forceinline__ double min_image(double d, double L, double half_L, int periodic) {
  if (periodic) {
    if (d <= -half_L) d += L;
    if (d > half_L) d -= L;
  }
  return d;
}

__device__ __forceinline__ void cell_ijk(const Grid& g, int64_t e, int& i, int& j, int& k) {
  const int64_t nx = g.n[0], nxy = int64_t(g.n[0]) * g.n[1];
  k = int(e / nxy);
  const int64_t r = e - int64_t(k) * nxy;
  j = int(r / nx);
  i = int(r - int64_t(j) * nx);
}

// Walk the particle slots of every element of the stencil of element
// (ci, cj, ck): calls seg(a, b) for each contiguous slot range [a, b).  The
// rows are wave-uniform, so their loads are scalar.  Ranges come row by row,
// x-contiguous inside a row; on a periodic x axis a row can split in two.
template <class Seg>
__device__ __forceinline__ void for_each_segment(const Grid& g, const StencilRow* __restrict__ rows,
                                                 int nrows, const int32_t* __restrict__ cell_start,
                                                 int ci, int cj, int ck, Seg&& seg) {
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  for (int r = 0; r < nrows; ++r) {
    const StencilRow R = rows[r];
    int jj = cj + R.dj, kk = ck + R.dk;
    if (g.periodic[1]) {
      jj += (jj < 0) ? ny : 0;
      jj -= (jj >= ny) ? ny : 0;
    } else if (jj < 0 || jj >= ny) {
      continue;
    }
    if (g.periodic[2]) {
      kk += (kk < 0) ? nz : 0;
      kk -= (kk >= nz) ? nz : 0;
    } else if (kk < 0 || kk >= nz) {
      continue;
    }
    const int64_t base = int64_t(nx) * (jj + int64_t(ny) * kk);
    int lo = ci + R.di_lo, hi = ci + R.di_hi;
    if (g.periodic[0]) {
      if (lo < 0) {
        seg(cell_start[base + lo + nx], cell_start[base + nx]);
        lo = 0;
      }
      if (hi >= nx) {
        seg(cell_start[base], cell_start[base + hi - nx + 1]);
        hi = nx - 1;
      }
    } else {
      lo = lo < 0 ? 0 : lo;
      hi = hi >= nx ? nx - 1 : hi;
    }
    if (lo <= hi) seg(cell_start[base + lo], cell_start[base + hi + 1]);
  }
}

// G_reg(x) f accumulated into u, FP64, the reference's formula
// (ggem_stokes.C:304-357): G f = eps[(c1 + c2) f + (c1 - c2) (x.f) x / r^2]
// with c1 = (erf(xi r) - erf(alpha r))/r and
// c2 = 2xi/sqrt(pi) e^{-xi^2 r^2} - 2alpha/sqrt(pi) e^{-alpha^2 r^2};
// r < 1e-6 gives coeff_2 f.
__device__ __forceinline__ void green_acc_f64(const GGEMConst& k, double dx, double dy, double dz,
                                              double fx, double fy, double fz, double& ux,
                                              double& uy, double& uz) {
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (r < 1e-6) {
    ux += k.coeff_2 * fx;
    uy += k.coeff_2 * fy;
    uz += k.coeff_2 * fz;
    return;
  }
  const double r2 = r * r;
  const double kr = k.ksi * r, ar = k.alpha * r;
  const double c1 = (erf(kr) - erf(ar)) / r;
  const double c2 = k.two_ksi_sqrt_pi * exp(-kr * kr) - k.two_alpha_sqrt_pi * exp(-ar * ar);
  const double A = k.eps * (c1 + c2);
  const double B = k.eps * (c1 - c2) * ((dx * fx + dy * fy + dz * fz) / r2);
  ux += A * fx + B * dx;
  uy += A * fy + B * dy;
  uz += A * fz + B * dz;
}

// Kahan-compensated FP32 accumulator.
struct KahanF {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return double(s) - double(c); }
};

// FP32 constants of the pair term.
struct GGEMConst32 {
  float eps, alpha, ksi, two_alpha_sqrt_pi, two_ksi_sqrt_pi, coeff_2;
  float erf_switch;  // alpha r below this: erf difference; above: erfc difference
};

// Pair coefficients A, B of G f = A f + B (x.f) x in FP32.  The erf
// difference is taken as erfc(alpha r) - erfc(xi r) once alpha r passes the
// switch, which keeps its relative error at a few ulp where both erf values
// approach 1.
__device__ __forceinline__ void green_coef_f32(const GGEMConst32& k, float r2, float& A,
                                               float& Bx) {
  const float r = __builtin_sqrtf(r2);
  const float kr = k.ksi * r, ar = k.alpha * r;
  float d;
  if (ar < k.erf_switch) d = erff(kr) - erff(ar);
  else d = erfcf(ar) - erfcf(kr);
  const float c1 = d / r;
  const float c2 = k.two_ksi_sqrt_pi * __expf(-kr * kr) - k.two_alpha_sqrt_pi * __expf(-ar * ar);
  A = k.eps * (c1 + c2);
  Bx = k.eps * (c1 - c2) / r2;
}

}  // namespace ggem
