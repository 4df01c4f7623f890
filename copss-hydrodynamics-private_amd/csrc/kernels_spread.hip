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
// Gather form, no atomics.  The Gauss points of an element form a tensor
// grid q = c + (g_a, g_b, g_c) h/2 and the minimum image is taken per axis,
// so with d = x_p - c and u = d - g h/2 per axis,
//   g = alpha^3/pi^{3/2} E_x[a] E_y[b] E_z[c] (2.5 - S_x[a] - S_y[b] - S_z[c]),
//   S = alpha^2 u^2, E = e^{-S}.
// One thread owns the NQ1^2 Gauss points of one x-coordinate a of one
// element: per particle it forms the FP64 displacement once, the 1 + 2 NQ1
// factors E (instead of one exponential per Gauss point) and updates
// NQ1^2 x 3 accumulators.  The NQ1 threads of an element read the same
// list entries (broadcast loads).  rho goes to LDS for the IF contraction
// (FP64 arithmetic in both precisions).  FP32C: factors and kernel values
// in FP32 after the FP64 displacement, partial sums merged every 32
// particles into Kahan accumulators, rho staged in LDS as FP32.
#include "ggem_device.h"

namespace ggem {

namespace {

constexpr int kMergeEvery = 32;
// elements per block, candidates of the autotuner (index 0 is the default);
// the per-element arithmetic does not depend on the choice
constexpr int kElemChoices[] = {64, 32, 16};
constexpr int kNElem = sizeof(kElemChoices) / sizeof(kElemChoices[0]);

template <class T>
struct SpreadArgs {
  Grid g;
  T alpha2, alpha3_pi_23;
  T goff[3][3];  // g_a h_d / 2 per axis d and Gauss index a
  const int64_t* off;
  const uint32_t* elist;
  const double4* x4;
  const float4* fg32;
  const double4* fg64;
};

// e^{-a} for a >= 0 in FP32 from the hardware exp2: the rounding of a log2(e)
// perturbs the result by at most |a| 2^-24 relative, and a e^{-a} <= 1/e
// keeps the absolute error below 2^-25 of the peak weight.
__device__ __forceinline__ float exp_neg(float a) {
  constexpr float L2E = 1.44269502163f;
  return __builtin_amdgcn_exp2f(-a * L2E);
}
__device__ __forceinline__ double exp_neg(double a) { return exp(-a); }

__device__ __forceinline__ float4 load_f(const float4* f32, const double4*, uint32_t t, float) {
  return f32[t];
}
__device__ __forceinline__ double4 load_f(const float4*, const double4* f64, uint32_t t, double) {
  return f64[t];
}

template <class T, int NQ1, bool EXACT, int kElems>
__global__ void __launch_bounds__(kElems * NQ1) k_spread(SpreadArgs<T> a, const double* __restrict__ IF,
                                                        int n_u, double* __restrict__ rho,
                                                        double* __restrict__ Fe) {
  constexpr int NQP = NQ1 * NQ1 * NQ1;
  constexpr int NB = NQ1 * NQ1;  // Gauss points per thread
  using R = typename std::conditional<sizeof(T) == 4, float, double>::type;  // LDS rho type
  using V4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
  __shared__ R s_rho[kElems * NQP * 3];

  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int le = threadIdx.x / NQ1;
  const int qa = threadIdx.x - le * NQ1;
  const int64_t e = blk * kElems + le;
  const bool active = e < a.g.n_elem;
  if (active) {
    int ci, cj, ck;
    elem_ijk(a.g, e, ci, cj, ck);
    const double4 cen = make_double4(a.g.bmin[0] + (double(ci) + 0.5) * a.g.h[0],
                                     a.g.bmin[1] + (double(cj) + 0.5) * a.g.h[1],
                                     a.g.bmin[2] + (double(ck) + 0.5) * a.g.h[2], 0.0);
    const T gx = a.goff[0][qa];
    const int64_t b0 = a.off[e], b1 = a.off[e + 1];
    T p[NB][3];
    KahanF kh[sizeof(T) == 4 ? NB : 1][3];
#pragma unroll
    for (int b = 0; b < NB; ++b) p[b][0] = p[b][1] = p[b][2] = T(0);
    // the next particle's position and force records are loaded while the
    // current one is summed, its list entry one step earlier
    const int64_t lastk = b1 > b0 ? b1 - 1 : b0;
    uint32_t e1 = b1 > b0 ? a.elist[b0] : 0u;
    uint32_t e2 = b1 > b0 ? a.elist[min(b0 + 1, lastk)] : 0u;
    double4 x1 = make_double4(0.0, 0.0, 0.0, 0.0);
    V4 f1{};
    if (b1 > b0) {  // an empty list reads nothing (a context may hold no particles)
      x1 = a.x4[e1 & kSlotMask];
      f1 = load_f(a.fg32, a.fg64, e1 & kSlotMask, T(0));
    }
    for (int64_t k = b0; k < b1; ++k) {
      const uint32_t en = e1;
      const double4 xt = x1;
      const V4 f = f1;
      e1 = e2;
      x1 = a.x4[e1 & kSlotMask];
      f1 = load_f(a.fg32, a.fg64, e1 & kSlotMask, T(0));
      e2 = a.elist[min(k + 2, lastk)];
      double d0, d1, d2;
      pair_disp<EXACT>(a.g, cen, xt, en >> kSlotBits, d0, d1, d2);
      // x_p - q per axis; with the explicit minimum image it is taken again
      // relative to each Gauss point, as point_vector(q, x_p) does
      // (assemble_stokes.C:643-672): in a box barely wider than the
      // stencil, x_p - c and x_p - q may lie on different images
      T ux, uyv[NQ1], uzv[NQ1];
      if constexpr (EXACT) {
        ux = T(min_image(d0 - double(gx), a.g.L[0], a.g.half_L[0], a.g.periodic[0]));
#pragma unroll
        for (int g = 0; g < NQ1; ++g) {
          uyv[g] = T(min_image(d1 - double(a.goff[1][g]), a.g.L[1], a.g.half_L[1], a.g.periodic[1]));
          uzv[g] = T(min_image(d2 - double(a.goff[2][g]), a.g.L[2], a.g.half_L[2], a.g.periodic[2]));
        }
      } else {
        ux = T(d0) - gx;
#pragma unroll
        for (int g = 0; g < NQ1; ++g) {
          uyv[g] = T(d1) - a.goff[1][g];
          uzv[g] = T(d2) - a.goff[2][g];
        }
      }
      const T sx = a.alpha2 * (ux * ux);
      const T ex = a.alpha3_pi_23 * exp_neg(sx);
      T ey[NQ1], sy[NQ1], ez[NQ1], sz[NQ1];
#pragma unroll
      for (int g = 0; g < NQ1; ++g) {
        const T uy = uyv[g];
        const T uz = uzv[g];
        sy[g] = a.alpha2 * (uy * uy);
        sz[g] = a.alpha2 * (uz * uz);
        ey[g] = ex * exp_neg(sy[g]);
        ez[g] = exp_neg(sz[g]);
      }
      // g = E_y E_z (T0 - S_y - S_z) = E_z P_b - E_y Q_c with
      // T0 = 2.5 - S_x, P_b = E_y (T0 - S_y), Q_c = E_z S_z: two
      // operations per Gauss point
      const T t0 = T(2.5) - sx;
      T pb[NQ1], qc[NQ1];
#pragma unroll
      for (int g = 0; g < NQ1; ++g) {
        pb[g] = ey[g] * (t0 - sy[g]);
        qc[g] = ez[g] * sz[g];
      }
#pragma unroll
      for (int c = 0; c < NQ1; ++c)
#pragma unroll
        for (int b = 0; b < NQ1; ++b) {
          const T gv = fma(ez[c], pb[b], -(ey[b] * qc[c]));
          p[c * NQ1 + b][0] += gv * T(f.x);
          p[c * NQ1 + b][1] += gv * T(f.y);
          p[c * NQ1 + b][2] += gv * T(f.z);
        }
      if (sizeof(T) == 4 && ((k - b0) & (kMergeEvery - 1)) == kMergeEvery - 1) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            kh[sizeof(T) == 4 ? b : 0][j].add(float(p[b][j]));
            p[b][j] = T(0);
          }
      }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int qp = (b / NQ1) * NQ1 * NQ1 + (b % NQ1) * NQ1 + qa;  // qp = a + n (b + n c)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double v;
        if (sizeof(T) == 4) {
          kh[sizeof(T) == 4 ? b : 0][j].add(float(p[b][j]));
          v = kh[sizeof(T) == 4 ? b : 0][j].value();
        } else {
          v = double(p[b][j]);
        }
        s_rho[(le * NQP + qp) * 3 + j] = R(v);
        if (rho) rho[(e * NQP + qp) * 3 + j] = v;
      }
    }
  }
  __syncthreads();
  if (!Fe) return;
  // IF contraction, one (element, component) pair per thread and step: the
  // element's NQP values of rho in registers, the table row of each shape
  // function read with scalar loads (uniform over the block), FP64 FMAs in
  // Gauss point order.  The results are staged in the rho buffer (free once
  // rho sits in registers) and written out as one contiguous range per
  // pass, so the stores are coalesced.
  constexpr int M = (3 + NQ1 - 1) / NQ1;  // pairs per thread (blockDim = kElems NQ1)
  const int per = 3 * n_u;
  double r[M][NQP];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int idx = threadIdx.x + m * blockDim.x;
    const int l2 = min(idx / 3, kElems - 1);
    const int j = idx - 3 * (idx / 3);
#pragma unroll
    for (int q = 0; q < NQP; ++q) r[m][q] = double(s_rho[(l2 * NQP + q) * 3 + min(j, 2)]);
  }
  __syncthreads();
  double* stage = reinterpret_cast<double*>(s_rho);
  constexpr int kStage = int(sizeof(s_rho) / (sizeof(double)));  // doubles the buffer holds
  const int step = kStage / per;  // elements per pass (n_u <= 27 keeps it >= 1)
  const int64_t e_first = blk * kElems;
  const int ne = int(min<int64_t>(kElems, a.g.n_elem - e_first));
  for (int p0 = 0; p0 < ne; p0 += step) {
    const int p1 = min(ne, p0 + step);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int idx = threadIdx.x + m * blockDim.x;
      const int l2 = idx / 3;
      const int j = idx - 3 * l2;
      if (l2 < p0 || l2 >= p1) continue;
      double* o = stage + (l2 - p0) * per + j * n_u;
      for (int l = 0; l < n_u; ++l) {
        const double* __restrict__ row = IF + l * NQP;
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < NQP; ++q) s = __builtin_fma(row[q], r[m][q], s);
        o[l] = s;
      }
    }
    __syncthreads();
    double* __restrict__ dst = Fe + (e_first + p0) * per;
    for (int i = threadIdx.x; i < (p1 - p0) * per; i += blockDim.x) dst[i] = stage[i];
    __syncthreads();
  }
}

template <class T, int NQ1, int KE>
void launch_spk(ggem_ctx* ctx, const SpreadArgs<T>& a, const double* IF_d, int n_u, double* rho, double* Fe) {
  const unsigned nb = unsigned((ctx->grid.n_elem + KE - 1) / KE);
  if (ctx->exact_minimg)
    hipLaunchKernelGGL((k_spread<T, NQ1, true, KE>), dim3(nb), dim3(KE * NQ1), 0, ctx->stream, a, IF_d, n_u, rho,
                       Fe);
  else
    hipLaunchKernelGGL((k_spread<T, NQ1, false, KE>), dim3(nb), dim3(KE * NQ1), 0, ctx->stream, a, IF_d, n_u,
                       rho, Fe);
}

template <class T, int NQ1>
hipError_t launch_sp(ggem_ctx* ctx, int c, const double* gp, const double* IF_d, int n_u, double* rho,
                     double* Fe) {
  SpreadArgs<T> a;
  a.g = ctx->grid;
  a.alpha2 = T(ctx->gc.alpha2);
  a.alpha3_pi_23 = T(ctx->gc.alpha3_pi_23);
  for (int d = 0; d < 3; ++d)
    for (int g = 0; g < 3; ++g) a.goff[d][g] = T(gp[g] * (0.5 * ctx->grid.h[d]));
  a.off = ctx->elist_off;
  a.elist = ctx->elist;
  a.x4 = ctx->x4;
  a.fg32 = ctx->fg32;
  a.fg64 = ctx->fg64;
  if (c < 0 || c >= kNElem) return hipErrorInvalidValue;
  switch (kElemChoices[c]) {
    case 32: launch_spk<T, NQ1, 32>(ctx, a, IF_d, n_u, rho, Fe); break;
    case 16: launch_spk<T, NQ1, 16>(ctx, a, IF_d, n_u, rho, Fe); break;
    default: launch_spk<T, NQ1, 64>(ctx, a, IF_d, n_u, rho, Fe); break;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_spread(ggem_ctx* ctx, int precision, int nqp_1d, int n_u, const double* IF_d,
                         double* rho, double* Fe) {
  double gp[3];
  if (nqp_1d == 2) {
    gp[0] = -std::sqrt(1. / 3.);
    gp[1] = std::sqrt(1. / 3.);
    gp[2] = 0.0;
  } else {
    gp[0] = -std::sqrt(3. / 5.);
    gp[1] = 0.0;
    gp[2] = std::sqrt(3. / 5.);
  }
  auto run = [&](int c) {
    if (nqp_1d == 2)
      return precision == GGEM_FP64 ? launch_sp<double, 2>(ctx, c, gp, IF_d, n_u, rho, Fe)
                                    : launch_sp<float, 2>(ctx, c, gp, IF_d, n_u, rho, Fe);
    return precision == GGEM_FP64 ? launch_sp<double, 3>(ctx, c, gp, IF_d, n_u, rho, Fe)
                                  : launch_sp<float, 3>(ctx, c, gp, IF_d, n_u, rho, Fe);
  };
  // rho and Fe are outputs only: idempotent
  const int c = tune_pick(ctx, tune_key("spread", precision, nqp_1d, ctx->grid.n_elem), kNElem, run);
  ProfScope ps(ctx, precision == GGEM_FP64 ? "spread_fp64" : "spread_fp32c");
  return run(c);
}

}  // namespace ggem
