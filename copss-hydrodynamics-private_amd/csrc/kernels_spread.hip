// kernels_spread.hip -- (c) Gaussian force spreading onto the Gauss points
// of every element and the contraction with the element's IF table.
//
// Reference: AssembleStokes::compute_element_rhs, point-force branch
// (assemble_stokes.C:643-672) for every element of assemble_global_F
// (:473-516), GGEMStokes::smoothed_force_exp (ggem_stokes.C:117-128) and
// the IF table of assemble_int_force (:733-757):
//   rho(qp) = sum_{p in list(e)} g(x_p - q) f_p,
//   g(r) = alpha^3/pi^{3/2} e^{-alpha^2 r^2} (2.5 - alpha^2 r^2),
//   Fe[j*n_u + l] = sum_qp IF[l*nqp + qp] rho_j(qp).
//
// Gather form, no atomics: a block owns EPB elements x nqp Gauss points,
// one thread per (element, Gauss point).  The nqp threads of an element walk
// the same particle list (broadcast loads); rho stays in LDS for the IF
// contraction, which runs in FP64 in both precisions.
#include "ggem_device.h"

namespace ggem {

namespace {

struct SpreadArgs {
  Grid g;
  double alpha2, alpha3_pi_23;
  float alpha2_f, alpha3_pi_23_f;
  double gp[3];  // 1D Gauss points, ascending
  const StencilRow* rows;
  int nrows;
  const int32_t* cell_start;
  const double* xs;
  const double* fs;
  const float* fs32;
  int64_t n;
};

constexpr int kMaxNu = 27;

template <int P, int NQ1>
__global__ void __launch_bounds__(256) k_spread(SpreadArgs a, const double* __restrict__ IF, int n_u,
                                               double* __restrict__ rho, double* __restrict__ Fe) {
  constexpr int NQP = NQ1 * NQ1 * NQ1;
  constexpr int EPB = 256 / NQP;
  __shared__ double s_rho[EPB * NQP * 3];
  __shared__ double s_if[kMaxNu * NQP];
  for (int q = threadIdx.x; q < n_u * NQP; q += blockDim.x) s_if[q] = IF[q];

  const int le = threadIdx.x / NQP;
  const int qp = threadIdx.x - le * NQP;
  const int64_t e = int64_t(blockIdx.x) * EPB + le;
  const bool active = le < EPB && e < a.g.n_elem;
  if (active) {
    int ci, cj, ck;
    cell_ijk(a.g, e, ci, cj, ck);
    const int qa = qp % NQ1, qb = (qp / NQ1) % NQ1, qc = qp / (NQ1 * NQ1);
    const double qx = (a.g.bmin[0] + (double(ci) + 0.5) * a.g.h[0]) + a.gp[qa] * (0.5 * a.g.h[0]);
    const double qy = (a.g.bmin[1] + (double(cj) + 0.5) * a.g.h[1]) + a.gp[qb] * (0.5 * a.g.h[1]);
    const double qz = (a.g.bmin[2] + (double(ck) + 0.5) * a.g.h[2]) + a.gp[qc] * (0.5 * a.g.h[2]);
    const int64_t n = a.n;
    double r0, r1, r2v;
    if (P == GGEM_FP64) {
      double s0 = 0, s1 = 0, s2 = 0;
      for_each_segment(a.g, a.rows, a.nrows, a.cell_start, ci, cj, ck, [&](int32_t lo, int32_t hi) {
        for (int32_t t = lo; t < hi; ++t) {
          const double dx = min_image(a.xs[t] - qx, a.g.L[0], a.g.half_L[0], a.g.periodic[0]);
          const double dy = min_image(a.xs[n + t] - qy, a.g.L[1], a.g.half_L[1], a.g.periodic[1]);
          const double dz = min_image(a.xs[2 * n + t] - qz, a.g.L[2], a.g.half_L[2], a.g.periodic[2]);
          const double a2r2 = a.alpha2 * (dx * dx + dy * dy + dz * dz);
          const double gv = a.alpha3_pi_23 * exp(-a2r2) * (2.5 - a2r2);
          s0 += gv * a.fs[t];
          s1 += gv * a.fs[n + t];
          s2 += gv * a.fs[2 * n + t];
        }
      });
      r0 = s0;
      r1 = s1;
      r2v = s2;
    } else {
      KahanF k0, k1, k2;
      for_each_segment(a.g, a.rows, a.nrows, a.cell_start, ci, cj, ck, [&](int32_t lo, int32_t hi) {
        for (int32_t t = lo; t < hi; ++t) {
          const float dx = float(min_image(a.xs[t] - qx, a.g.L[0], a.g.half_L[0], a.g.periodic[0]));
          const float dy = float(min_image(a.xs[n + t] - qy, a.g.L[1], a.g.half_L[1], a.g.periodic[1]));
          const float dz =
              float(min_image(a.xs[2 * n + t] - qz, a.g.L[2], a.g.half_L[2], a.g.periodic[2]));
          const float a2r2 = a.alpha2_f * (dx * dx + dy * dy + dz * dz);
          const float gv = a.alpha3_pi_23_f * __expf(-a2r2) * (2.5f - a2r2);
          k0.add(gv * a.fs32[t]);
          k1.add(gv * a.fs32[n + t]);
          k2.add(gv * a.fs32[2 * n + t]);
        }
      });
      r0 = k0.value();
      r1 = k1.value();
      r2v = k2.value();
    }
    s_rho[(le * NQP + qp) * 3 + 0] = r0;
    s_rho[(le * NQP + qp) * 3 + 1] = r1;
    s_rho[(le * NQP + qp) * 3 + 2] = r2v;
    if (rho) {
      rho[(e * NQP + qp) * 3 + 0] = r0;
      rho[(e * NQP + qp) * 3 + 1] = r1;
      rho[(e * NQP + qp) * 3 + 2] = r2v;
    }
  }
  __syncthreads();
  if (!Fe) return;
  const int per = 3 * n_u;
  for (int idx = threadIdx.x; idx < EPB * per; idx += blockDim.x) {
    const int l2 = idx / per;
    const int rem = idx - l2 * per;
    const int j = rem / n_u;
    const int l = rem - j * n_u;
    const int64_t e2 = int64_t(blockIdx.x) * EPB + l2;
    if (e2 >= a.g.n_elem) continue;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < NQP; ++q) acc += s_if[l * NQP + q] * s_rho[(l2 * NQP + q) * 3 + j];
    Fe[e2 * per + rem] = acc;
  }
}

}  // namespace

hipError_t launch_spread(ggem_ctx* ctx, int precision, int nqp_1d, int n_u, const double* IF_d,
                         double* rho, double* Fe) {
  SpreadArgs a;
  a.g = ctx->grid;
  a.alpha2 = ctx->gc.alpha2;
  a.alpha3_pi_23 = ctx->gc.alpha3_pi_23;
  a.alpha2_f = float(ctx->gc.alpha2);
  a.alpha3_pi_23_f = float(ctx->gc.alpha3_pi_23);
  if (nqp_1d == 2) {
    a.gp[0] = -std::sqrt(1. / 3.);
    a.gp[1] = std::sqrt(1. / 3.);
    a.gp[2] = 0.0;
  } else {
    a.gp[0] = -std::sqrt(3. / 5.);
    a.gp[1] = 0.0;
    a.gp[2] = std::sqrt(3. / 5.);
  }
  a.rows = ctx->rows_d;
  a.nrows = ctx->n_rows;
  a.cell_start = ctx->cell_start;
  a.xs = ctx->xs;
  a.fs = ctx->fs;
  a.fs32 = ctx->fs32;
  a.n = ctx->n;
  const int nqp = nqp_1d * nqp_1d * nqp_1d;
  const int epb = 256 / nqp;
  const unsigned nb = unsigned((ctx->grid.n_elem + epb - 1) / epb);
  ProfScope ps(ctx, precision == GGEM_FP64 ? "spread_fp64" : "spread_fp32c");
  if (nqp_1d == 2) {
    if (precision == GGEM_FP64)
      hipLaunchKernelGGL((k_spread<GGEM_FP64, 2>), dim3(nb), dim3(256), 0, ctx->stream, a, IF_d, n_u,
                         rho, Fe);
    else
      hipLaunchKernelGGL((k_spread<GGEM_FP32C, 2>), dim3(nb), dim3(256), 0, ctx->stream, a, IF_d, n_u,
                         rho, Fe);
  } else {
    if (precision == GGEM_FP64)
      hipLaunchKernelGGL((k_spread<GGEM_FP64, 3>), dim3(nb), dim3(256), 0, ctx->stream, a, IF_d, n_u,
                         rho, Fe);
    else
      hipLaunchKernelGGL((k_spread<GGEM_FP32C, 3>), dim3(nb), dim3(256), 0, ctx->stream, a, IF_d, n_u,
                         rho, Fe);
  }
  return hipGetLastError();
}

}  // namespace ggem
