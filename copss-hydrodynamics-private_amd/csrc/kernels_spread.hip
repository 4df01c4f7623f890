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
// element: per particle it forms the displacement once and its own
// factors E_x[a], E_y[a], E_z[a]; the NQ1 threads of an element sit in
// consecutive lanes of one wave (64 / NQ1 elements per wave, the rest of the
// lanes idle) and exchange their y and z factors with lane permutes, so
// each factor is evaluated once per element.  Each thread then updates
// NQ1^2 x 3 accumulators.  The NQ1 threads read the same list entries
// (broadcast loads).  rho goes to LDS for the IF contraction
// (FP64 arithmetic in both precisions).  FP32C: factors and kernel values
// in FP32 from the cell-relative position records, partial sums of 32
// particles added into the thread's rho entries in LDS (FP32).
#include "ggem_device.h"

namespace ggem {

namespace {

constexpr int kMergeEvery = 32;
// waves per block, candidates of the autotuner (index 0 is the default);
// the per-element arithmetic does not depend on the choice
constexpr int kWaveChoices[] = {3, 2, 1};
constexpr int kNElem = sizeof(kWaveChoices) / sizeof(kWaveChoices[0]);

template <class T>
struct SpreadArgs {
  Grid g;
  T alpha2, alpha3_pi_23;
  T goff[3][3];  // g_a h_d / 2 per axis d and Gauss index a
  const int64_t* off;
  const uint32_t* elist;
  const double4* x4;
  const float4* pc32;  // FP32C: position relative to the cell centre, packed cell
  const float4* fg32;
  const double4* fg64;
  int fe_node_major;  // Fe[e][l][j] (3 values per shape function) instead of Fe[e][j][l]
};

// e^{-a} for a >= 0 in FP32 from the hardware exp2: the rounding of a log2(e)
// perturbs the result by at most |a| 2^-24 relative, and a e^{-a} <= 1/e
// keeps the absolute error below 2^-25 of the peak weight.
__device__ __forceinline__ float exp_neg(float a) {
  constexpr float L2E = 1.44269502163f;
  return __builtin_amdgcn_exp2f(-a * L2E);
}
__device__ __forceinline__ double exp_neg(double a) { return exp(-a); }

// Gauss point of the thread's b-th point: qp = a + n (b + n c) for a = qa,
// b = b % n, c = b / n
template <int NQ1>
__device__ __forceinline__ int qp_of_n(int b, int qa) {
  return (b / NQ1) * NQ1 * NQ1 + (b % NQ1) * NQ1 + qa;
}

// Gaussian factors of one particle seen from one thread: its own x factor
// (with the normalisation) and the element's y and z factors
template <class T, int N>
struct Factors {
  T ex, sx;
  T sy[N], eyr[N], ez[N], qc[N];
};

// value of v in lane src of the wave (a lane permute through the LDS
// crossbar, no LDS storage)
__device__ __forceinline__ float lane_get(float v, int src) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ double lane_get(double v, int src) {
  const int2 w = *reinterpret_cast<int2*>(&v);
  int2 r;
  r.x = __builtin_amdgcn_ds_bpermute(src << 2, w.x);
  r.y = __builtin_amdgcn_ds_bpermute(src << 2, w.y);
  return *reinterpret_cast<double*>(&r);
}

__device__ __forceinline__ float4 load_pos(const double4*, const float4* pc, uint32_t t, float4) { return pc[t]; }
__device__ __forceinline__ double4 load_pos(const double4* x4, const float4*, uint32_t t, double4) { return x4[t]; }

__device__ __forceinline__ float4 load_f(const float4* f32, const double4*, uint32_t t, float) {
  return f32[t];
}
__device__ __forceinline__ double4 load_f(const float4*, const double4* f64, uint32_t t, double) {
  return f64[t];
}

template <class T, int NQ1, bool EXACT, int WAVES>
__global__ void __launch_bounds__(WAVES * 64) k_spread(SpreadArgs<T> a, const double* __restrict__ IF,
                                                      int n_u, double* __restrict__ rho,
                                                      double* __restrict__ Fe) {
  constexpr int NQP = NQ1 * NQ1 * NQ1;
  constexpr int NB = NQ1 * NQ1;     // Gauss points per thread
  constexpr int EPW = 64 / NQ1;     // elements per wave
  constexpr int kElems = WAVES * EPW;
  using R = typename std::conditional<sizeof(T) == 4, float, double>::type;  // LDS rho type
  using V4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
  __shared__ R s_rho[kElems * NQP * 3];

  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x & 63;
  const int lw = lane / NQ1;                       // element of the lane within its wave
  const int qa = lane - lw * NQ1;
  const int le = int(threadIdx.x >> 6) * EPW + lw;  // element within the block
  const int grp = lane - qa;                        // first lane of the element
  const int64_t e = blk * kElems + le;
  const bool active = lw < EPW && e < a.g.n_elem;
  if (active) {
    int ci, cj, ck;
    elem_ijk(a.g, e, ci, cj, ck);
    const double4 cen = make_double4(a.g.bmin[0] + (double(ci) + 0.5) * a.g.h[0],
                                     a.g.bmin[1] + (double(cj) + 0.5) * a.g.h[1],
                                     a.g.bmin[2] + (double(ck) + 0.5) * a.g.h[2], 0.0);
    const T gx = a.goff[0][qa], gy = a.goff[1][qa], gz = a.goff[2][qa];
    const int64_t b0 = a.off[e], b1 = a.off[e + 1];
    T p[NB][3];
    // FP32C: partial sums of kMergeEvery particles in registers, added into
    // the thread's own entries of the LDS rho buffer (two-level summation,
    // no registers held for the totals)
    R* const mine = s_rho + le * NQP * 3;
    if (sizeof(T) == 4) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int j = 0; j < 3; ++j) mine[qp_of_n<NQ1>(b, qa) * 3 + j] = R(0);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) p[b][0] = p[b][1] = p[b][2] = T(0);
    // the next particle's position and force records are loaded while the
    // current one is summed, its list entry one step earlier
    const int64_t lastk = b1 > b0 ? b1 - 1 : b0;
    uint32_t e1 = b1 > b0 ? a.elist[b0] : 0u;
    uint32_t e2 = b1 > b0 ? a.elist[min(b0 + 1, lastk)] : 0u;
    // FP32C (minimum image by the list's shift codes): the particle's
    // cell-relative record; the displacement from the element centre is
    // (cell difference) h + offset, every term small (16 bytes a particle
    // instead of the 32-byte FP64 position)
    constexpr bool REL = sizeof(T) == 4 && !EXACT;
    using P4 = typename std::conditional<REL, float4, double4>::type;
    const float hx = float(a.g.h[0]), hy = float(a.g.h[1]), hz = float(a.g.h[2]);
    P4 x1{};
    V4 f1{};
    if (b1 > b0) {  // an empty list reads nothing (a context may hold no particles)
      x1 = load_pos(a.x4, a.pc32, e1 & kSlotMask, P4{});
      f1 = load_f(a.fg32, a.fg64, e1 & kSlotMask, T(0));
    }
    // one particle's contribution to the thread's NQ1^2 Gauss points:
    // g = E_y E_z (T0 - S_y - S_z) = E_z P_b - E_y Q_c with T0 = 2.5 - S_x,
    // P_b = E_y (T0 - S_y), Q_c = E_z S_z (E_y with the E_x factor): two
    // operations per Gauss point; FP32C partial sums merge every
    // kMergeEvery particles
    auto sum_particle = [&](const Factors<T, NQ1>& F, const V4& f, int64_t j) {
      const T t0 = T(2.5) - F.sx;
      T ey[NQ1], pb[NQ1];
#pragma unroll
      for (int g = 0; g < NQ1; ++g) {
        ey[g] = F.ex * F.eyr[g];
        pb[g] = ey[g] * (t0 - F.sy[g]);
      }
#pragma unroll
      for (int c = 0; c < NQ1; ++c)
#pragma unroll
        for (int b = 0; b < NQ1; ++b) {
          const T gv = fma(F.ez[c], pb[b], -(ey[b] * F.qc[c]));
          p[c * NQ1 + b][0] += gv * T(f.x);
          p[c * NQ1 + b][1] += gv * T(f.y);
          p[c * NQ1 + b][2] += gv * T(f.z);
        }
      if (sizeof(T) == 4 && ((j - b0) & (kMergeEvery - 1)) == kMergeEvery - 1) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            mine[qp_of_n<NQ1>(b, qa) * 3 + q] += R(p[b][q]);
            p[b][q] = T(0);
          }
      }
    };
    for (int64_t k = b0; k < b1; ++k) {
      const uint32_t en = e1;
      const P4 xt = x1;
      const V4 f = f1;
      e1 = e2;
      x1 = load_pos(a.x4, a.pc32, e1 & kSlotMask, P4{});
      f1 = load_f(a.fg32, a.fg64, e1 & kSlotMask, T(0));
      e2 = a.elist[min(k + 2, lastk)];
      double d0 = 0.0, d1 = 0.0, d2 = 0.0;
      float r0 = 0.f, r1 = 0.f, r2 = 0.f;
      if constexpr (REL) {
        const uint32_t cp = __float_as_uint(xt.w), code = en >> kSlotBits;
        const int ui = int(cp & ((1u << kCellBitsX) - 1)) + (int(code & 1u) - int((code >> 1) & 1u)) * a.g.n[0];
        const int uj = int((cp >> kCellBitsX) & ((1u << kCellBitsY) - 1)) +
                       (int((code >> 2) & 1u) - int((code >> 3) & 1u)) * a.g.n[1];
        const int uk = int(cp >> (kCellBitsX + kCellBitsY)) + (int((code >> 4) & 1u) - int((code >> 5) & 1u)) * a.g.n[2];
        r0 = __builtin_fmaf(float(ui - ci), hx, float(xt.x));
        r1 = __builtin_fmaf(float(uj - cj), hy, float(xt.y));
        r2 = __builtin_fmaf(float(uk - ck), hz, float(xt.z));
      } else {
        pair_disp<EXACT>(a.g, cen, xt, en >> kSlotBits, d0, d1, d2);
      }
      // x_p - q per axis; with the explicit minimum image it is taken again
      // relative to each Gauss point, as point_vector(q, x_p) does
      // (assemble_stokes.C:643-672): in a box barely wider than the
      // stencil, x_p - c and x_p - q may lie on different images
      // (the lane's own Gauss index qa on every axis: its y and z factors
      // go to the element's other lanes)
      T ux, uy_o, uz_o;
      if constexpr (EXACT) {
        ux = T(min_image(d0 - double(gx), a.g.L[0], a.g.half_L[0], a.g.periodic[0]));
        uy_o = T(min_image(d1 - double(gy), a.g.L[1], a.g.half_L[1], a.g.periodic[1]));
        uz_o = T(min_image(d2 - double(gz), a.g.L[2], a.g.half_L[2], a.g.periodic[2]));
      } else if constexpr (REL) {
        ux = T(r0) - gx;
        uy_o = T(r1) - gy;
        uz_o = T(r2) - gz;
      } else {
        ux = T(d0) - gx;
        uy_o = T(d1) - gy;
        uz_o = T(d2) - gz;
      }
      const T sx = a.alpha2 * (ux * ux);
      const T ex = a.alpha3_pi_23 * exp_neg(sx);
      // the y and z factors of index qa here, the others from the element's
      // other lanes (same operations on the same displacement)
      const T sy_o = a.alpha2 * (uy_o * uy_o);
      const T sz_o = a.alpha2 * (uz_o * uz_o);
      const T eyr_o = exp_neg(sy_o);
      const T ez_o = exp_neg(sz_o);
      const T qc_o = ez_o * sz_o;
      Factors<T, NQ1> nx;
      nx.ex = ex;
      nx.sx = sx;
#pragma unroll
      for (int g = 0; g < NQ1; ++g) {
        nx.sy[g] = lane_get(sy_o, grp + g);
        nx.eyr[g] = lane_get(eyr_o, grp + g);
        nx.ez[g] = lane_get(ez_o, grp + g);
        nx.qc[g] = lane_get(qc_o, grp + g);
      }
      sum_particle(nx, f, k);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int qp = qp_of_n<NQ1>(b, qa);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        R v;
        if (sizeof(T) == 4) v = mine[qp * 3 + j] + R(p[b][j]);
        else v = R(p[b][j]);
        mine[qp * 3 + j] = v;
        if (rho) rho[(e * NQP + qp) * 3 + j] = double(v);
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
  constexpr int M = (3 * kElems + WAVES * 64 - 1) / (WAVES * 64);  // pairs per thread
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
      double* o = stage + (l2 - p0) * per + (a.fe_node_major ? j : j * n_u);
      const int ls = a.fe_node_major ? 3 : 1;
      for (int l = 0; l < n_u; ++l) {
        const double* __restrict__ row = IF + l * NQP;
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < NQP; ++q) s = __builtin_fma(row[q], r[m][q], s);
        o[l * ls] = s;
      }
    }
    __syncthreads();
    double* __restrict__ dst = Fe + (e_first + p0) * per;
    for (int i = threadIdx.x; i < (p1 - p0) * per; i += blockDim.x) dst[i] = stage[i];
    __syncthreads();
  }
}

template <class T, int NQ1, int WAVES>
void launch_spk(ggem_ctx* ctx, const SpreadArgs<T>& a, const double* IF_d, int n_u, double* rho, double* Fe) {
  constexpr int KE = WAVES * (64 / NQ1);  // elements per block
  const unsigned nb = unsigned((ctx->grid.n_elem + KE - 1) / KE);
  if (explicit_minimg(ctx))
    hipLaunchKernelGGL((k_spread<T, NQ1, true, WAVES>), dim3(nb), dim3(WAVES * 64), 0, ctx->stream, a, IF_d, n_u,
                       rho, Fe);
  else
    hipLaunchKernelGGL((k_spread<T, NQ1, false, WAVES>), dim3(nb), dim3(WAVES * 64), 0, ctx->stream, a, IF_d,
                       n_u, rho, Fe);
}

template <class T, int NQ1>
hipError_t launch_sp(ggem_ctx* ctx, int c, const double* gp, const double* IF_d, int n_u, double* rho,
                     double* Fe, bool node_major) {
  SpreadArgs<T> a;
  a.g = ctx->grid;
  a.alpha2 = T(ctx->gc.alpha2);
  a.alpha3_pi_23 = T(ctx->gc.alpha3_pi_23);
  for (int d = 0; d < 3; ++d)
    for (int g = 0; g < 3; ++g) a.goff[d][g] = T(gp[g] * (0.5 * ctx->grid.h[d]));
  a.off = ctx->elist_off;
  a.elist = ctx->elist;
  a.x4 = ctx->x4;
  a.pc32 = ctx->pc32;
  a.fg32 = ctx->fg32;
  a.fg64 = ctx->fg64;
  a.fe_node_major = node_major ? 1 : 0;
  if (c < 0 || c >= kNElem) return hipErrorInvalidValue;
  switch (kWaveChoices[c]) {
    case 2: launch_spk<T, NQ1, 2>(ctx, a, IF_d, n_u, rho, Fe); break;
    case 1: launch_spk<T, NQ1, 1>(ctx, a, IF_d, n_u, rho, Fe); break;
    default: launch_spk<T, NQ1, 3>(ctx, a, IF_d, n_u, rho, Fe); break;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_spread(ggem_ctx* ctx, int precision, int nqp_1d, int n_u, const double* IF_d,
                         double* rho, double* Fe, bool node_major) {
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
      return precision == GGEM_FP64 ? launch_sp<double, 2>(ctx, c, gp, IF_d, n_u, rho, Fe, node_major)
                                    : launch_sp<float, 2>(ctx, c, gp, IF_d, n_u, rho, Fe, node_major);
    return precision == GGEM_FP64 ? launch_sp<double, 3>(ctx, c, gp, IF_d, n_u, rho, Fe, node_major)
                                  : launch_sp<float, 3>(ctx, c, gp, IF_d, n_u, rho, Fe, node_major);
  };
  // rho and Fe are outputs only: idempotent
  const int c = tune_pick(ctx, tune_key("spread", precision, nqp_1d, ctx->grid.n_elem), kNElem, run);
  ProfScope ps(ctx, precision == GGEM_FP64 ? "spread_fp64" : "spread_fp32c");
  return run(c);
}

}  // namespace ggem
