// kernels_pair.hip -- (b) the local regularized-Stokeslet pair sum
// u_i = sum_{j in nbr(i)} G_reg(x_j - x_i) f_j over the element stencil.
//
// Reference: GGEMStokes::local_solution_bead (ggem_stokes.C:427-503),
// local_solution_field (:360-424), green_tensor_local_regularized
// (:304-357), global_self_exclusion (:506-523) and the velocity combination
// of PMSystemStokes::compute_point_velocity (pm_system_stokes.C:441-457).
//
// One thread per target in cell order.  A wave holds 64 consecutive slots,
// i.e. neighbouring cells of one x-row, so the lanes walk the same stencil
// rows (wave-uniform row loads) over overlapping slot ranges (cache reuse of
// the neighbours' positions and forces).  Displacements are formed in FP64
// (minimum image) in both precisions; the FP32 form evaluates the tensor in
// FP32 and sums with Kahan compensation, and the combination is FP64.
#include "ggem_device.h"

namespace ggem {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(int64_t n, int b = kBlock) { return unsigned((n + b - 1) / b); }

struct PairArgs {
  Grid g;
  GGEMConst k;
  GGEMConst32 k32;
  const StencilRow* rows;
  int nrows;
  const int32_t* cell_start;
  const int32_t* slot_cell;
  const int32_t* perm;
  const double* xs;
  int64_t n;
};

// Sum of G_reg(x_t - x_s) f_t over the stencil of slot s's element, t != s.
template <int P>
__device__ __forceinline__ void pair_sum_slot(const PairArgs& a, int64_t s,
                                              const double* __restrict__ fs,
                                              const float* __restrict__ fs32, double& ux,
                                              double& uy, double& uz) {
  const int64_t n = a.n;
  int ci, cj, ck;
  cell_ijk(a.g, a.slot_cell[s], ci, cj, ck);
  const double x0 = a.xs[s], y0 = a.xs[n + s], z0 = a.xs[2 * n + s];
  if (P == GGEM_FP64) {
    double sx = 0, sy = 0, sz = 0;
    for_each_segment(a.g, a.rows, a.nrows, a.cell_start, ci, cj, ck, [&](int32_t lo, int32_t hi) {
      for (int32_t t = lo; t < hi; ++t) {
        if (t == s) continue;
        const double dx = min_image(a.xs[t] - x0, a.g.L[0], a.g.half_L[0], a.g.periodic[0]);
        const double dy = min_image(a.xs[n + t] - y0, a.g.L[1], a.g.half_L[1], a.g.periodic[1]);
        const double dz = min_image(a.xs[2 * n + t] - z0, a.g.L[2], a.g.half_L[2], a.g.periodic[2]);
        green_acc_f64(a.k, dx, dy, dz, fs[t], fs[n + t], fs[2 * n + t], sx, sy, sz);
      }
    });
    ux = sx;
    uy = sy;
    uz = sz;
  } else {
    KahanF kx, ky, kz;
    for_each_segment(a.g, a.rows, a.nrows, a.cell_start, ci, cj, ck, [&](int32_t lo, int32_t hi) {
      for (int32_t t = lo; t < hi; ++t) {
        if (t == s) continue;
        const float dx = float(min_image(a.xs[t] - x0, a.g.L[0], a.g.half_L[0], a.g.periodic[0]));
        const float dy = float(min_image(a.xs[n + t] - y0, a.g.L[1], a.g.half_L[1], a.g.periodic[1]));
        const float dz =
            float(min_image(a.xs[2 * n + t] - z0, a.g.L[2], a.g.half_L[2], a.g.periodic[2]));
        const float fx = fs32[t], fy = fs32[n + t], fz = fs32[2 * n + t];
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < 1e-12f) {
          kx.add(a.k32.coeff_2 * fx);
          ky.add(a.k32.coeff_2 * fy);
          kz.add(a.k32.coeff_2 * fz);
          continue;
        }
        float A, Bx;
        green_coef_f32(a.k32, r2, A, Bx);
        const float B = Bx * (dx * fx + dy * fy + dz * fz);
        kx.add(A * fx + B * dx);
        ky.add(A * fy + B * dy);
        kz.add(A * fz + B * dz);
      }
    });
    ux = kx.value();
    uy = ky.value();
    uz = kz.value();
  }
}

// Velocity of every bead, written in particle order (interleaved).
//   EPI_LOCAL:    U_loc (+ coeff_2 f for Lagrangian points, ggem_stokes.C:485-499)
//   EPI_MOBILITY: compute_point_velocity("disturbed") with U_glob from u_glob
template <int P>
__global__ void __launch_bounds__(kBlock) k_pair_beads(PairArgs a, const double* __restrict__ fs,
                                                      const float* __restrict__ fs32, int epi,
                                                      const double* __restrict__ u_glob,
                                                      double* __restrict__ out) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  double ux, uy, uz;
  pair_sum_slot<P>(a, s, fs, fs32, ux, uy, uz);
  const int64_t n = a.n;
  const double fx = fs[s], fy = fs[n + s], fz = fs[2 * n + s];
  if (a.k.point_type == GGEM_LAGRANGIAN_POINT) {
    ux += a.k.coeff_2 * fx;
    uy += a.k.coeff_2 * fy;
    uz += a.k.coeff_2 * fz;
  }
  const int64_t p = a.perm[s];
  if (epi == EPI_MOBILITY) {
    const double gx = u_glob ? u_glob[3 * p] : 0.0;
    const double gy = u_glob ? u_glob[3 * p + 1] : 0.0;
    const double gz = u_glob ? u_glob[3 * p + 2] : 0.0;
    if (a.k.point_type == GGEM_POLYMER_BEAD) {
      ux = gx + ux - a.k.coeff_1 * fx + fx;
      uy = gy + uy - a.k.coeff_1 * fy + fy;
      uz = gz + uz - a.k.coeff_1 * fz + fz;
    } else if (a.k.point_type == GGEM_LAGRANGIAN_POINT) {
      ux = gx + ux;
      uy = gy + uy;
      uz = gz + uz;
    } else {
      ux = uy = uz = 0.0;
    }
  }
  out[3 * p] = ux;
  out[3 * p + 1] = uy;
  out[3 * p + 2] = uz;
}

// Field points: sum over the element's particle list, own element
// included, no exclusion (ggem_stokes.C:396-420).
template <int P>
__global__ void __launch_bounds__(kBlock) k_pair_points(PairArgs a, const double* __restrict__ fs,
                                                       const float* __restrict__ fs32, int64_t m,
                                                       const double* __restrict__ pts,
                                                       const int32_t* __restrict__ elem,
                                                       double* __restrict__ u) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const int64_t n = a.n;
  int ci, cj, ck;
  cell_ijk(a.g, elem[q], ci, cj, ck);
  const double x0 = pts[3 * q], y0 = pts[3 * q + 1], z0 = pts[3 * q + 2];
  double ux, uy, uz;
  if (P == GGEM_FP64) {
    double sx = 0, sy = 0, sz = 0;
    for_each_segment(a.g, a.rows, a.nrows, a.cell_start, ci, cj, ck, [&](int32_t lo, int32_t hi) {
      for (int32_t t = lo; t < hi; ++t) {
        const double dx = min_image(a.xs[t] - x0, a.g.L[0], a.g.half_L[0], a.g.periodic[0]);
        const double dy = min_image(a.xs[n + t] - y0, a.g.L[1], a.g.half_L[1], a.g.periodic[1]);
        const double dz = min_image(a.xs[2 * n + t] - z0, a.g.L[2], a.g.half_L[2], a.g.periodic[2]);
        green_acc_f64(a.k, dx, dy, dz, fs[t], fs[n + t], fs[2 * n + t], sx, sy, sz);
      }
    });
    ux = sx;
    uy = sy;
    uz = sz;
  } else {
    KahanF kx, ky, kz;
    for_each_segment(a.g, a.rows, a.nrows, a.cell_start, ci, cj, ck, [&](int32_t lo, int32_t hi) {
      for (int32_t t = lo; t < hi; ++t) {
        const double ddx = min_image(a.xs[t] - x0, a.g.L[0], a.g.half_L[0], a.g.periodic[0]);
        const double ddy = min_image(a.xs[n + t] - y0, a.g.L[1], a.g.half_L[1], a.g.periodic[1]);
        const double ddz = min_image(a.xs[2 * n + t] - z0, a.g.L[2], a.g.half_L[2], a.g.periodic[2]);
        const float fx = fs32[t], fy = fs32[n + t], fz = fs32[2 * n + t];
        if (ddx * ddx + ddy * ddy + ddz * ddz < 1e-12) {
          kx.add(a.k32.coeff_2 * fx);
          ky.add(a.k32.coeff_2 * fy);
          kz.add(a.k32.coeff_2 * fz);
          continue;
        }
        const float dx = float(ddx), dy = float(ddy), dz = float(ddz);
        const float r2 = dx * dx + dy * dy + dz * dz;
        float A, Bx;
        green_coef_f32(a.k32, r2, A, Bx);
        const float B = Bx * (dx * fx + dy * fy + dz * fz);
        kx.add(A * fx + B * dx);
        ky.add(A * fy + B * dy);
        kz.add(A * fz + B * dz);
      }
    });
    ux = kx.value();
    uy = ky.value();
    uz = kz.value();
  }
  u[3 * q] = ux;
  u[3 * q + 1] = uy;
  u[3 * q + 2] = uz;
}

// Mobility on slot-ordered SoA vectors for nrhs right-hand sides sharing
// every tensor evaluation, with the Chebyshev recurrence fused into the
// epilogue (brownian_system.C:1037-1041, :1093-1101).
template <int P, int NR>
__global__ void __launch_bounds__(kBlock) k_apply_slots(PairArgs a, const double* __restrict__ x,
                                                       double* v, ChebEpi epi, int64_t vstride) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const int64_t n = a.n;
  int ci, cj, ck;
  cell_ijk(a.g, a.slot_cell[s], ci, cj, ck);
  const double x0 = a.xs[s], y0 = a.xs[n + s], z0 = a.xs[2 * n + s];
  double acc[NR][3];
  if (P == GGEM_FP64) {
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r][0] = acc[r][1] = acc[r][2] = 0.0;
    for_each_segment(a.g, a.rows, a.nrows, a.cell_start, ci, cj, ck, [&](int32_t lo, int32_t hi) {
      for (int32_t t = lo; t < hi; ++t) {
        if (t == s) continue;
        const double dx = min_image(a.xs[t] - x0, a.g.L[0], a.g.half_L[0], a.g.periodic[0]);
        const double dy = min_image(a.xs[n + t] - y0, a.g.L[1], a.g.half_L[1], a.g.periodic[1]);
        const double dz = min_image(a.xs[2 * n + t] - z0, a.g.L[2], a.g.half_L[2], a.g.periodic[2]);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        double A, Bx;
        if (r < 1e-6) {
          A = a.k.coeff_2;
          Bx = 0.0;
        } else {
          const double r2 = r * r;
          const double kr = a.k.ksi * r, ar = a.k.alpha * r;
          const double c1 = (erf(kr) - erf(ar)) / r;
          const double c2 =
              a.k.two_ksi_sqrt_pi * exp(-kr * kr) - a.k.two_alpha_sqrt_pi * exp(-ar * ar);
          A = a.k.eps * (c1 + c2);
          Bx = a.k.eps * (c1 - c2) / r2;
        }
#pragma unroll
        for (int r_ = 0; r_ < NR; ++r_) {
          const double* xr = x + r_ * vstride;
          const double fx = xr[t], fy = xr[n + t], fz = xr[2 * n + t];
          const double B = Bx * (dx * fx + dy * fy + dz * fz);
          acc[r_][0] += A * fx + B * dx;
          acc[r_][1] += A * fy + B * dy;
          acc[r_][2] += A * fz + B * dz;
        }
      }
    });
  } else {
    KahanF kk[NR][3];
    for_each_segment(a.g, a.rows, a.nrows, a.cell_start, ci, cj, ck, [&](int32_t lo, int32_t hi) {
      for (int32_t t = lo; t < hi; ++t) {
        if (t == s) continue;
        const float dx = float(min_image(a.xs[t] - x0, a.g.L[0], a.g.half_L[0], a.g.periodic[0]));
        const float dy = float(min_image(a.xs[n + t] - y0, a.g.L[1], a.g.half_L[1], a.g.periodic[1]));
        const float dz =
            float(min_image(a.xs[2 * n + t] - z0, a.g.L[2], a.g.half_L[2], a.g.periodic[2]));
        const float r2 = dx * dx + dy * dy + dz * dz;
        float A, Bx;
        if (r2 < 1e-12f) {
          A = a.k32.coeff_2;
          Bx = 0.f;
        } else {
          green_coef_f32(a.k32, r2, A, Bx);
        }
#pragma unroll
        for (int r_ = 0; r_ < NR; ++r_) {
          const double* xr = x + r_ * vstride;
          const float fx = float(xr[t]), fy = float(xr[n + t]), fz = float(xr[2 * n + t]);
          const float B = Bx * (dx * fx + dy * fy + dz * fz);
          kk[r_][0].add(A * fx + B * dx);
          kk[r_][1].add(A * fy + B * dy);
          kk[r_][2].add(A * fz + B * dz);
        }
      }
    });
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[r][c] = kk[r][c].value();
  }
  const double c_k = epi.coef ? *epi.coef : epi.c_host;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t idx = r * vstride + c * n + s;
      const double xv = x[idx];
      double m = acc[r][c];
      if (a.k.point_type == GGEM_POLYMER_BEAD) m = m - a.k.coeff_1 * xv + xv;
      else if (a.k.point_type == GGEM_LAGRANGIAN_POINT) m = m + a.k.coeff_2 * xv;
      else m = 0.0;
      if (epi.kind == 0) {
        v[idx] = m;
      } else if (epi.kind == 1) {
        v[idx] = epi.db * xv + epi.da * m;
      } else {
        double t = 2. * epi.db * xv + 2. * epi.da * m;
        t += -1.0 * epi.xm1[idx];
        v[idx] = t;
        epi.y[idx] += c_k * t;
      }
    }
  }
}

PairArgs make_args(const ggem_ctx* ctx) {
  PairArgs a;
  a.g = ctx->grid;
  a.k = ctx->gc;
  a.k32.eps = float(ctx->gc.eps);
  a.k32.alpha = float(ctx->gc.alpha);
  a.k32.ksi = float(ctx->gc.ksi);
  a.k32.two_alpha_sqrt_pi = float(ctx->gc.two_alpha_sqrt_pi);
  a.k32.two_ksi_sqrt_pi = float(ctx->gc.two_ksi_sqrt_pi);
  a.k32.coeff_2 = float(ctx->gc.coeff_2);
  a.k32.erf_switch = 0.5f;
  a.rows = ctx->rows_d;
  a.nrows = ctx->n_rows;
  a.cell_start = ctx->cell_start;
  a.slot_cell = ctx->slot_cell;
  a.perm = ctx->perm;
  a.xs = ctx->xs;
  a.n = ctx->n;
  return a;
}

template <int P, int NR>
void launch_apply_nr(ggem_ctx* ctx, const PairArgs& a, const double* x, double* v, ChebEpi epi) {
  hipLaunchKernelGGL((k_apply_slots<P, NR>), dim3(blocks_for(a.n)), dim3(kBlock), 0, ctx->stream, a, x,
                     v, epi, int64_t(3) * a.n);
}

template <int P>
void launch_apply_chunk(ggem_ctx* ctx, const PairArgs& a, int nr, const double* x, double* v,
                        const ChebEpi& epi) {
  switch (nr) {
    case 1: launch_apply_nr<P, 1>(ctx, a, x, v, epi); break;
    case 2: launch_apply_nr<P, 2>(ctx, a, x, v, epi); break;
    case 4: launch_apply_nr<P, 4>(ctx, a, x, v, epi); break;
    case 8: launch_apply_nr<P, 8>(ctx, a, x, v, epi); break;
    default: launch_apply_nr<P, 16>(ctx, a, x, v, epi); break;
  }
}

}  // namespace

hipError_t launch_pair_beads(ggem_ctx* ctx, int precision, int epi, const double* u_glob,
                             double* out) {
  if (ctx->n == 0) return hipSuccess;
  PairArgs a = make_args(ctx);
  ProfScope ps(ctx, precision == GGEM_FP64 ? "pair_beads_fp64" : "pair_beads_fp32c");
  if (precision == GGEM_FP64)
    hipLaunchKernelGGL(k_pair_beads<GGEM_FP64>, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream,
                       a, ctx->fs, ctx->fs32, epi, u_glob, out);
  else
    hipLaunchKernelGGL(k_pair_beads<GGEM_FP32C>, dim3(blocks_for(ctx->n)), dim3(kBlock), 0,
                       ctx->stream, a, ctx->fs, ctx->fs32, epi, u_glob, out);
  return hipGetLastError();
}

hipError_t launch_pair_points(ggem_ctx* ctx, int precision, int64_t m, const double* pts,
                              const int32_t* elem, double* u) {
  if (m == 0) return hipSuccess;
  PairArgs a = make_args(ctx);
  ProfScope ps(ctx, precision == GGEM_FP64 ? "pair_points_fp64" : "pair_points_fp32c");
  if (precision == GGEM_FP64)
    hipLaunchKernelGGL(k_pair_points<GGEM_FP64>, dim3(blocks_for(m)), dim3(kBlock), 0, ctx->stream, a,
                       ctx->fs, ctx->fs32, m, pts, elem, u);
  else
    hipLaunchKernelGGL(k_pair_points<GGEM_FP32C>, dim3(blocks_for(m)), dim3(kBlock), 0, ctx->stream, a,
                       ctx->fs, ctx->fs32, m, pts, elem, u);
  return hipGetLastError();
}

hipError_t launch_apply_slots(ggem_ctx* ctx, int precision, int nrhs, const double* x, double* v,
                              const ChebEpi& epi) {
  if (ctx->n == 0) return hipSuccess;
  PairArgs a = make_args(ctx);
  ProfScope ps(ctx, precision == GGEM_FP64 ? "apply_fp64" : "apply_fp32c");
  const int64_t stride = 3 * ctx->n;
  int done = 0;
  while (done < nrhs) {
    const int left = nrhs - done;
    const int nr = left >= 16 ? 16 : left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
    ChebEpi e = epi;
    if (e.xm1) e.xm1 += done * stride;
    if (e.y) e.y += done * stride;
    if (precision == GGEM_FP64)
      launch_apply_chunk<GGEM_FP64>(ctx, a, nr, x + done * stride, v + done * stride, e);
    else
      launch_apply_chunk<GGEM_FP32C>(ctx, a, nr, x + done * stride, v + done * stride, e);
    done += nr;
  }
  return hipGetLastError();
}

}  // namespace ggem
