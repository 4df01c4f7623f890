// ggem_device.h -- device-side building blocks shared by the kernels:
// minimum-image displacement, the element-stencil walk over the cell-sorted
// particle store, XCD-aware block numbering, and the regularized-Stokeslet
// pair coefficients in FP64 and FP32.
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

// Single periodic shift of a neighbour entry (code: 2 bits per axis).
__device__ __forceinline__ double apply_shift(double d, uint32_t code2, double L) {
  const int w = int(code2 & 1u) - int((code2 >> 1) & 1u);
  return w ? __builtin_fma(double(w), L, d) : d;
}

// Reference element id e = i + nx (j + ny k) -> (i, j, k).
__device__ __forceinline__ void elem_ijk(const Grid& g, int64_t e, int& i, int& j, int& k) {
  const int64_t nx = g.n[0], nxy = int64_t(g.n[0]) * g.n[1];
  k = int(e / nxy);
  const int64_t r = e - int64_t(k) * nxy;
  j = int(r / nx);
  i = int(r - int64_t(j) * nx);
}

// Brick-order cell index of cell (i, j, k) and its inverse.
__device__ __forceinline__ int64_t cell_index_nb(int nb0, int nb1, int i, int j, int k) {
  constexpr int B = kBrickBits, M = (1 << kBrickBits) - 1;
  const int64_t brick = (i >> B) + int64_t(nb0) * ((j >> B) + int64_t(nb1) * (k >> B));
  return (brick << (3 * B)) | ((k & M) << (2 * B)) | ((j & M) << B) | (i & M);
}

__device__ __forceinline__ int64_t cell_index(const Grid& g, int i, int j, int k) {
  return cell_index_nb(g.nb[0], g.nb[1], i, j, k);
}

__device__ __forceinline__ void cell_ijk(const Grid& g, int64_t c, int& i, int& j, int& k) {
  constexpr int B = kBrickBits, M = (1 << kBrickBits) - 1;
  const int64_t brick = c >> (3 * B);
  const int l = int(c & ((1 << (3 * B)) - 1));
  const int64_t bxy = int64_t(g.nb[0]) * g.nb[1];
  const int bk = int(brick / bxy);
  const int64_t r = brick - int64_t(bk) * bxy;
  const int bj = int(r / g.nb[0]);
  const int bi = int(r - int64_t(bj) * g.nb[0]);
  i = (bi << B) | (l & M);
  j = (bj << B) | ((l >> B) & M);
  k = (bk << B) | (l >> (2 * B));
}

// Workgroups are dealt to the 8 XCDs round-robin; renumber them so that each
// XCD works on one contiguous range of slots and neighbouring targets share
// its L2.  Bijective on [0, nb).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, rmd = nb % 8;
  const int64_t x = b % 8, i = b / 8;
  return x * q + (x < rmd ? x : rmd) + i;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// x mod n for x a few periods at most outside [0, n) (no integer division)
__device__ __forceinline__ int wrap_axis(int x, int n) {
  while (x < 0) x += n;
  while (x >= n) x -= n;
  return x;
}

// q / d for 0 <= q < 2^14, 1 <= d <= 64: (q + 1/2) / d is at least 1/128
// away from an integer, far more than the FP32 rounding of the product
__device__ __forceinline__ int div_small(int q, float inv_d) { return int((float(q) + 0.5f) * inv_d); }

constexpr int kBrickCells = 1 << (3 * kBrickBits);
constexpr int kBoxMaxCells = 8192;  // box cells whose tables a brick kernel keeps in LDS
constexpr int kBoxBits = 6;         // box coordinates: below 2^6 per axis
constexpr int kBoxField = 8;        // ... packed one byte per axis (x, y, z from the low byte)

// Box-cell differences t - o per axis of two packed box cells, given kb =
// 0x808080 - o: each byte of t + kb holds its axis' difference + 128 (the
// coordinates are below 64, so no byte carries into the next), and the
// byte-to-float conversions are exact.
__device__ __forceinline__ void box_diff(uint32_t t, uint32_t kb, float& dx, float& dy, float& dz) {
  const uint32_t d = t + kb;
  dx = float(d & 0xffu) - 128.f;
  dy = float((d >> 8) & 0xffu) - 128.f;
  dz = float((d >> 16) & 0xffu) - 128.f;
}
__device__ __forceinline__ uint32_t box_pack(int x, int y, int z) {
  return uint32_t(x) | (uint32_t(y) << kBoxField) | (uint32_t(z) << (2 * kBoxField));
}

// Box of one brick: the brick widened by the stencil reach on every side,
// origin (ox, oy, oz) and extent (ex, ey, ez) in cells of the brick's grid,
// periodic images unwrapped; box cells are numbered x-major through the
// box.  The grid fields are copied (no pointer into the kernel arguments,
// which would force a private copy of them).
struct BoxGeom {
  int64_t coff;
  int n[3], per[3], nb0, nb1;
  float h[3];
  int ox, oy, oz, ex, ey, ez, nbc;
  int bx0, by0, bz0;  // the brick's first cell
  float inv_ex, inv_ey;

  // coff: first brick-order cell of the grid's system
  __device__ __forceinline__ void init(const Grid& g, int64_t cell_off, const int32_t* reach, int64_t brick) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      n[d] = g.n[d];
      per[d] = g.periodic[d];
      h[d] = float(g.h[d]);
    }
    nb0 = g.nb[0];
    nb1 = g.nb[1];
    coff = cell_off;
    const int rx = reach[0], ry = reach[1], rz = reach[2];
    const int64_t lb = brick - (coff >> (3 * kBrickBits));
    const int64_t bxy = int64_t(nb0) * nb1;
    const int bk = int(lb / bxy);
    const int64_t rem = lb - int64_t(bk) * bxy;
    const int bj = int(rem / nb0);
    const int bi = int(rem - int64_t(bj) * nb0);
    bx0 = bi << kBrickBits;
    by0 = bj << kBrickBits;
    bz0 = bk << kBrickBits;
    ox = bx0 - rx;
    oy = by0 - ry;
    oz = bz0 - rz;
    ex = (1 << kBrickBits) + 2 * rx;
    ey = (1 << kBrickBits) + 2 * ry;
    ez = (1 << kBrickBits) + 2 * rz;
    nbc = ex * ey * ez;
    inv_ex = 1.0f / float(ex);
    inv_ey = 1.0f / float(ey);
  }
  __device__ __forceinline__ bool fits() const {
    return nbc <= kBoxMaxCells && ex <= (1 << kBoxBits) && ey <= (1 << kBoxBits) && ez <= (1 << kBoxBits);
  }
  __device__ __forceinline__ void xyz(int q, int& x, int& y, int& z) const {
    const int yz = div_small(q, inv_ex);
    x = q - yz * ex;
    z = div_small(yz, inv_ey);
    y = yz - z * ey;
  }
  // brick-order index of grid cell (i, j, k) of the system
  __device__ __forceinline__ int64_t index(int i, int j, int k) const {
    return coff + cell_index_nb(nb0, nb1, i, j, k);
  }
  // global cell of box cell q; false outside a non-periodic domain
  __device__ __forceinline__ bool cell(int q, int64_t& c) const {
    int x, y, z;
    xyz(q, x, y, z);
    int X = ox + x, Y = oy + y, Z = oz + z;
    if (X < 0 || X >= n[0]) {
      if (!per[0]) return false;
      X = wrap_axis(X, n[0]);
    }
    if (Y < 0 || Y >= n[1]) {
      if (!per[1]) return false;
      Y = wrap_axis(Y, n[1]);
    }
    if (Z < 0 || Z >= n[2]) {
      if (!per[2]) return false;
      Z = wrap_axis(Z, n[2]);
    }
    c = index(X, Y, Z);
    return true;
  }
};

// The box cells of a brick in LDS, whole block of NT threads: first[q] =
// first slot of box cell q's grid cell (when first is not null), cst[q] =
// particles of the box cells before q (cst[nbc] = all of them).  Cells
// outside a non-periodic domain are empty.  s_w holds NT/64 ints.  Ends
// with a barrier.
template <int NT>
__device__ __forceinline__ void stage_box_cells(const BoxGeom& B, const int32_t* __restrict__ cs,
                                                int32_t* first, int32_t* cst, int32_t* s_w) {
  const int per = (B.nbc + NT - 1) / NT;
  const int q0 = min(B.nbc, int(threadIdx.x) * per), q1 = min(B.nbc, q0 + per);
  int mine = 0;
  // eight cells at a time, their cell-start loads all in flight (cells
  // outside the domain or past the chunk read cell 0 and count nothing)
  constexpr int G = 8;
  for (int q = q0; q < q1; q += G) {
    int64_t c[G];
    bool in[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      in[u] = q + u < q1 && B.cell(q + u, c[u]);
      if (!in[u]) c[u] = 0;
    }
    int f[G], e[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      f[u] = cs[c[u]];
      e[u] = cs[c[u] + 1];
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (q + u >= q1) break;
      const int k = in[u] ? e[u] - f[u] : 0;
      cst[q + u] = k;
      if (first) first[q + u] = in[u] ? f[u] : 0;
      mine += k;
    }
  }
  int incl = mine;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if ((threadIdx.x & 63) >= o) incl += y;
  }
  if ((threadIdx.x & 63) == 63) s_w[threadIdx.x >> 6] = incl;
  __syncthreads();
  int run = incl - mine;
  for (int w = 0; w < int(threadIdx.x >> 6); ++w) run += s_w[w];
  for (int q = q0; q < q1; ++q) {
    const int k = cst[q];
    cst[q] = run;
    run += k;
  }
  if (threadIdx.x == NT - 1) cst[B.nbc] = run;  // the last thread ends at the total
  __syncthreads();
}

// Element l (0..511, x fastest) of the brick of box B: its grid cell and
// reference element id; false for the padding cells of the last bricks of
// an axis, which are no elements.
__device__ __forceinline__ bool brick_elem(const BoxGeom& B, int l, int& ci, int& cj, int& ck, int64_t& e) {
  constexpr int M = (1 << kBrickBits) - 1;
  ci = B.bx0 + (l & M);
  cj = B.by0 + ((l >> kBrickBits) & M);
  ck = B.bz0 + (l >> (2 * kBrickBits));
  e = ci + int64_t(B.n[0]) * (cj + int64_t(B.n[1]) * ck);
  return ci < B.n[0] && cj < B.n[1] && ck < B.n[2];
}

// Stencil rows r0 .. r1 - 1 of the element in grid cell (ci, cj, ck) of the brick as runs
// of box cells: calls run(q0, q1, code) for box cells q0..q1 of one row
// with the periodic shift code of that run, in the order for_each_segment
// visits the same cells (rows in stencil order; within a row the part past
// the low face, then the part past the high face, then the rest, x
// ascending), so the particles come out in the same order.
template <class Run>
__device__ __forceinline__ void box_walk_rows(const BoxGeom& B, const StencilRow* __restrict__ rows, int r0, int r1,
                                              int ci, int cj, int ck, Run&& run) {
  for (int r = r0; r < r1; ++r) {
    const StencilRow R = rows[r];
    const int jj = cj + R.dj, kk = ck + R.dk;
    uint32_t cy = 0, cz = 0;
    if (jj < 0 || jj >= B.n[1]) {
      if (!B.per[1]) continue;
      cy = jj < 0 ? 2u : 1u;
    }
    if (kk < 0 || kk >= B.n[2]) {
      if (!B.per[2]) continue;
      cz = kk < 0 ? 2u : 1u;
    }
    const uint32_t cyz = (cy << 2) | (cz << 4);
    const int rowq = ((kk - B.oz) * B.ey + (jj - B.oy)) * B.ex - B.ox;  // box cell of unwrapped x = rowq + x
    int lo = ci + R.di_lo, hi = ci + R.di_hi;
    if (B.per[0]) {
      if (lo < 0) {
        run(rowq + lo, rowq - 1, cyz | 2u);
        lo = 0;
      }
      if (hi >= B.n[0]) {
        run(rowq + B.n[0], rowq + hi, cyz | 1u);
        hi = B.n[0] - 1;
      }
    } else {
      lo = lo < 0 ? 0 : lo;
      hi = hi >= B.n[0] ? B.n[0] - 1 : hi;
    }
    if (lo <= hi) run(rowq + lo, rowq + hi, cyz);
  }
}

// box_walk_rows over the whole stencil
template <class Run>
__device__ __forceinline__ void box_walk(const BoxGeom& B, const StencilRow* __restrict__ rows, int nrows,
                                         int ci, int cj, int ck, Run&& run) {
  box_walk_rows(B, rows, 0, nrows, ci, cj, ck, run);
}

// Walk the particle slots of every element of the stencil of element
// (ci, cj, ck): calls seg(a, b, code) for each contiguous slot range [a, b)
// with the periodic shift code of that range (a row's x-run is one range
// per brick it crosses), rows in stencil order, x ascending within a row.
template <class Seg>
__device__ __forceinline__ void for_each_segment(const Grid& g, const StencilRow* __restrict__ rows,
                                                 int nrows, const int32_t* __restrict__ cell_start,
                                                 int ci, int cj, int ck, Seg&& seg) {
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  for (int r = 0; r < nrows; ++r) {
    const StencilRow R = rows[r];
    int jj = cj + R.dj, kk = ck + R.dk;
    uint32_t cy = 0, cz = 0;
    if (jj < 0 || jj >= ny) {
      if (!g.periodic[1]) continue;
      // a neighbour cell reached across the low face sits near the top:
      // d ~ +L, shift -L (code 2); across the high face: d ~ -L, +L (code 1)
      if (jj < 0) {
        jj += ny;
        cy = 2;
      } else {
        jj -= ny;
        cy = 1;
      }
    }
    if (kk < 0 || kk >= nz) {
      if (!g.periodic[2]) continue;
      if (kk < 0) {
        kk += nz;
        cz = 2;
      } else {
        kk -= nz;
        cz = 1;
      }
    }
    const uint32_t cyz = (cy << 2) | (cz << 4);
    // x-run [x0, x1] of row (jj, kk): contiguous slots within each brick
    auto run = [&](int x0, int x1, uint32_t code) {
      constexpr int M = (1 << kBrickBits) - 1;
      for (int p = x0; p <= x1;) {
        const int q = min(x1, p | M);
        const int64_t c0 = cell_index(g, p, jj, kk);
        seg(cell_start[c0], cell_start[c0 + (q - p) + 1], code);
        p = q + 1;
      }
    };
    int lo = ci + R.di_lo, hi = ci + R.di_hi;
    if (g.periodic[0]) {
      if (lo < 0) {
        run(lo + nx, nx - 1, cyz | 2u);
        lo = 0;
      }
      if (hi >= nx) {
        run(0, hi - nx, cyz | 1u);
        hi = nx - 1;
      }
    } else {
      lo = lo < 0 ? 0 : lo;
      hi = hi >= nx ? nx - 1 : hi;
    }
    if (lo <= hi) run(lo, hi, cyz);
  }
}

// Displacement x_t - x_s of a neighbour entry: the raw difference, then
// either the entry's periodic shift or, in exact mode, the full
// minimum-image test.  Both are the reference's arithmetic: one subtraction,
// at most one shift by the box length.
template <bool EXACT>
__device__ __forceinline__ void pair_disp(const Grid& g, const double4& s, const double4& t,
                                          uint32_t code, double& dx, double& dy, double& dz) {
  dx = t.x - s.x;
  dy = t.y - s.y;
  dz = t.z - s.z;
  if (EXACT) {
    dx = min_image(dx, g.L[0], g.half_L[0], g.periodic[0]);
    dy = min_image(dy, g.L[1], g.half_L[1], g.periodic[1]);
    dz = min_image(dz, g.L[2], g.half_L[2], g.periodic[2]);
  } else if (code) {
    dx = apply_shift(dx, code, g.L[0]);
    dy = apply_shift(dy, code >> 2, g.L[1]);
    dz = apply_shift(dz, code >> 4, g.L[2]);
  }
}

// G_reg(x) f accumulated into u, FP64, the reference's formula
// (ggem_stokes.C:304-357): G f = eps[(c1 + c2) f + (c1 - c2) (x.f) x / r^2]
// with c1 = (erf(xi r) - erf(alpha r))/r and
// c2 = 2xi/sqrt(pi) e^{-xi^2 r^2} - 2alpha/sqrt(pi) e^{-alpha^2 r^2};
// r < 1e-6 gives coeff_2 f.
__device__ __forceinline__ void green_coef_f64(const GGEMConst& k, double r2, double& A,
                                               double& Bx) {
  const double r = sqrt(r2);
  if (r < 1e-6) {
    A = k.coeff_2;
    Bx = 0.0;
    return;
  }
  const double rr = r * r;
  const double kr = k.ksi * r, ar = k.alpha * r;
  const double c1 = (erf(kr) - erf(ar)) / r;
  const double c2 = k.two_ksi_sqrt_pi * exp(-kr * kr) - k.two_alpha_sqrt_pi * exp(-ar * ar);
  A = k.eps * (c1 + c2);
  Bx = k.eps * (c1 - c2) / rr;
}

__device__ __forceinline__ void green_acc_f64(const GGEMConst& k, double dx, double dy, double dz,
                                              double fx, double fy, double fz, double& ux,
                                              double& uy, double& uz) {
  double A, Bx;
  green_coef_f64(k, dx * dx + dy * dy + dz * dz, A, Bx);
  const double B = Bx * (dx * fx + dy * fy + dz * fz);
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
constexpr int kErfcxDeg = 10;
struct GGEMConst32 {
  float eps, alpha, ksi, two_alpha_sqrt_pi, two_ksi_sqrt_pi, coeff_2;
  float erf_switch;   // alpha r below this: erf difference; above: erfc difference
  float erfcx_scale;  // 2 / t_max of the erfcx fit
  float erfcx_c[kErfcxDeg + 1];
};

// e^{-z^2} in FP32 to about one ulp: z^2 and z^2 log2(e) are carried with
// their rounding errors (FMA) into a first-order correction of 2^y.
__device__ __forceinline__ float exp_neg_sq(float z) {
  constexpr float L2E = 1.44269502163f;        // log2(e) rounded to FP32
  constexpr float L2E_LO = 1.925963033e-08f;   // log2(e) - L2E
  constexpr float LN2 = 0.693147180560f;
  const float s = z * z;
  const float se = __builtin_fmaf(z, z, -s);
  const float y = -s * L2E;
  const float ye = __builtin_fmaf(-s, L2E, -y) - __builtin_fmaf(s, L2E_LO, se * L2E);
  return __builtin_amdgcn_exp2f(y) * __builtin_fmaf(ye, LN2, 1.0f);
}

__device__ __forceinline__ float rcp_nr(float a) {
  const float t0 = __builtin_amdgcn_rcpf(a);
  const float e = __builtin_fmaf(-a, t0, 1.0f);
  return __builtin_fmaf(t0, e, t0);
}

// Scaled complementary error function e^{z^2} erfc(z) for z >= 0.45:
// t = 1/(1 + z/2), erfcx = t Q(2t/t_max - 1), Q a degree-10 least-squares
// fit of erfcx(z)(1 + z/2) (relative error ~1e-7 in FP32).
__device__ __forceinline__ float erfcx_f32(const GGEMConst32& k, float z) {
  const float t = rcp_nr(__builtin_fmaf(0.5f, z, 1.0f));
  const float x = __builtin_fmaf(t, k.erfcx_scale, -1.0f);
  float q = k.erfcx_c[kErfcxDeg];
#pragma unroll
  for (int i = kErfcxDeg - 1; i >= 0; --i) q = __builtin_fmaf(q, x, k.erfcx_c[i]);
  return t * q;
}

// Pair coefficients A, Bx of G f = A f + Bx (x.f) x in FP32.  Beyond the
// switch the erf difference is taken as erfc(alpha r) - erfc(xi r) =
// e^{-alpha^2 r^2} erfcx(alpha r) - e^{-xi^2 r^2} erfcx(xi r), sharing the two
// exponentials with c2; this keeps a few-ulp relative error where both erf
// values approach 1.
__device__ __forceinline__ void green_coef_f32(const GGEMConst32& k, float r2, float& A, float& Bx) {
  if (r2 < 1e-12f) {  // r < r_eps (ggem_stokes.C:312-318)
    A = k.coeff_2;
    Bx = 0.f;
    return;
  }
  // 1/r from the hardware reciprocal square root with one Newton step
  float ir = __builtin_amdgcn_rsqf(r2);
  ir = ir * __builtin_fmaf(-0.5f * r2, ir * ir, 1.5f);
  const float r = r2 * ir;
  const float ar = k.alpha * r, kr = k.ksi * r;
  const float ea = exp_neg_sq(ar), ek = exp_neg_sq(kr);
  float d;
  if (ar >= k.erf_switch) d = ea * erfcx_f32(k, ar) - ek * erfcx_f32(k, kr);
  else d = erff(kr) - erff(ar);
  const float c1 = d * ir;
  const float c2 = k.two_ksi_sqrt_pi * ek - k.two_alpha_sqrt_pi * ea;
  A = k.eps * (c1 + c2);
  Bx = k.eps * (c1 - c2) * (ir * ir);
}

}  // namespace ggem
