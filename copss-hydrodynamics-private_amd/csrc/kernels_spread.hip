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

// The Gaussian factors of one particle at displacement (ux, uy, uz) from
// the thread's Gauss point (its own index on every axis): the x factor with
// the normalisation, and the y and z factors of the element's NQ1 lanes
// from lanes grp .. grp + NQ1 - 1 (each lane evaluates its own index; same
// operations on the same displacement)
template <class T, int NQ1>
__device__ __forceinline__ Factors<T, NQ1> particle_factors(T ux, T uy_o, T uz_o, T alpha2, T alpha3_pi_23,
                                                            int grp) {
  const T sx = alpha2 * (ux * ux);
  const T ex = alpha3_pi_23 * exp_neg(sx);
  const T sy_o = alpha2 * (uy_o * uy_o);
  const T sz_o = alpha2 * (uz_o * uz_o);
  const T eyr_o = exp_neg(sy_o);
  const T ez_o = exp_neg(sz_o);
  const T qc_o = ez_o * sz_o;
  Factors<T, NQ1> F;
  F.ex = ex;
  F.sx = sx;
#pragma unroll
  for (int g = 0; g < NQ1; ++g) {
    F.sy[g] = lane_get(sy_o, grp + g);
    F.eyr[g] = lane_get(eyr_o, grp + g);
    F.ez[g] = lane_get(ez_o, grp + g);
    F.qc[g] = lane_get(qc_o, grp + g);
  }
  return F;
}

// One particle's contribution to the thread's NQ1^2 Gauss points:
// g = E_y E_z (T0 - S_y - S_z) = E_z P_b - E_y Q_c with T0 = 2.5 - S_x,
// P_b = E_y (T0 - S_y), Q_c = E_z S_z (E_y with the E_x factor): two
// operations per Gauss point, then f g into the partial sums p
template <class T, int NQ1, class V4>
__device__ __forceinline__ void add_particle(const Factors<T, NQ1>& F, const V4& f, T (&p)[NQ1 * NQ1][3]) {
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
    auto sum_particle = [&](const Factors<T, NQ1>& F, const V4& f, int64_t j) {
      add_particle<T, NQ1>(F, f, p);
      // FP32C partial sums merge every kMergeEvery particles
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
      const Factors<T, NQ1> nx = particle_factors<T, NQ1>(ux, uy_o, uz_o, a.alpha2, a.alpha3_pi_23, grp);
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

// ---- Spreading by brick (FP32C) ----
//
// One block per brick of 8^3 elements.  The brick's box (the brick widened
// by the stencil reach, periodic images unwrapped) goes to LDS once: per
// box cell the records before it, then every particle of the box as two
// 16-byte records in box-cell order (offset from its cell centre with the
// packed box cell; force).  The waves then work independently, each taking
// the next group of 64 / NQ1 consecutive elements of the brick with
// k_spread's lane layout (NQ1 lanes an element, one per Gauss x-index):
// the element lists, the record numbers of each stencil in the element
// list's order (box_walk), go to the wave's own LDS buffer, the lanes of
// an element walking a third (half) of the stencil rows each; the lists
// are summed with k_spread's operations (the cell difference from the
// packed box cells equals the one k_spread forms from the list's shift
// codes, and the partial sums merge after the same particles, here into
// registers); the IF contraction then takes component j = the lane's
// Gauss x-index, gathers the 27 (8) values of rho it needs from the
// element's other lanes and sums them in k_spread's order.  The results
// equal k_spread's bit for bit.  A brick whose box holds more particles
// than the record table is flagged by that launch (k_spread_brick) and
// taken by a second one (k_spread_brick_hbm), which walks each element's
// whole stencil through the box cell table and reads the records from HBM
// (same order, same values; more registers, so it has a kernel of its own).
// No element lists in HBM.
constexpr int kBrickWaves = 8;  // waves per block
constexpr int kSortBuckets = 256;  // list lengths the element order tells apart (longer ones share the last)

struct BrickSpreadArgs {
  SpreadArgs<float> s;     // grid, constants, cell-relative records and forces
  const StencilRow* rows;
  int32_t nrows;
  int32_t reach[3];
  const int32_t* cs;       // cell starts, brick order
  int32_t recs_cap;        // records the LDS table holds
  int32_t cst_len;         // int32 words of the box cell table (box cells + 1, rounded up to 4)
  int32_t list_cap;        // 16-bit entries of each wave's list buffer (at least recs_cap)
  uint8_t* unstaged;       // [bricks] 1: the box exceeds the record table (written by k_spread_brick)
};

// LDS writes of a wave visible to its other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ float pick3(float a0, float a1, float a2, int k) { return k == 0 ? a0 : (k == 1 ? a1 : a2); }

// The IF contraction of one element from its lanes' rho values v (the
// thread's Gauss points qa + NQ1 b, three components): lane qa forms
// component qa (with NQ1 = 2, lane 0 also component 2) from all NQP Gauss
// points in order, FP64 FMAs as in k_spread, and stores its n_u values.
template <int NQ1>
__device__ __forceinline__ void brick_contract(const float (&v)[NQ1 * NQ1][3], int qa, int grp, const double* IF,
                                               int n_u, int node_major, double* __restrict__ fe) {
  constexpr int NB = NQ1 * NQ1, NQP = NQ1 * NB;
  auto chain = [&](const double (&r)[NQP], int j) {
    double* o = fe + (node_major ? j : j * n_u);
    const int ls = node_major ? 3 : 1;
    for (int l = 0; l < n_u; ++l) {
      const double* __restrict__ row = IF + l * NQP;
      double sum = 0.0;
#pragma unroll
      for (int q = 0; q < NQP; ++q) sum = __builtin_fma(row[q], r[q], sum);
      o[l * ls] = sum;
    }
  };
  if constexpr (NQ1 == 3) {
    // the value of Gauss point k + 3 b, component qa, comes from lane k:
    // lane qa + d sends its component (its qa) - d
    float own[NB], r1[NB], r2[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      own[b] = pick3(v[b][0], v[b][1], v[b][2], qa);
      r1[b] = lane_get(pick3(v[b][0], v[b][1], v[b][2], (qa + 2) % 3), grp + (qa + 1) % 3);
      r2[b] = lane_get(pick3(v[b][0], v[b][1], v[b][2], (qa + 1) % 3), grp + (qa + 2) % 3);
    }
    double r[NQP];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int k = 0; k < 3; ++k) r[k + 3 * b] = double(pick3(own[b], r1[b], r2[b], (k - qa + 3) % 3));
    chain(r, qa);
  } else {
    // two lanes: lane 0 forms components 0 and 2, lane 1 component 1
    float o1[NB], x1[NB], x2[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      o1[b] = qa == 0 ? v[b][0] : v[b][1];
      x1[b] = lane_get(qa == 0 ? v[b][1] : v[b][0], grp + 1 - qa);  // the other lane's component qa
      x2[b] = lane_get(v[b][2], grp + 1 - qa);
    }
    double r[NQP];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int k = 0; k < 2; ++k) r[k + 2 * b] = double(k == qa ? o1[b] : x1[b]);
    chain(r, qa);
    if (qa == 0) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int k = 0; k < 2; ++k) r[k + 2 * b] = double(k == 0 ? v[b][2] : x2[b]);
      chain(r, 2);
    }
  }
}

template <int NQ1, bool STAGED>
__device__ __forceinline__ void spread_brick(const BrickSpreadArgs& a, const double* __restrict__ IF, int n_u,
                                             double* __restrict__ rho, double* __restrict__ Fe) {
  constexpr int W = kBrickWaves, NT = W * 64;
  constexpr int NQP = NQ1 * NQ1 * NQ1;
  constexpr int NB = NQ1 * NQ1;
  constexpr int EPW = 64 / NQ1;                            // elements of a group
  constexpr int NG = (kBrickCells + EPW - 1) / EPW;        // groups of a brick
  extern __shared__ float4 dyn[];
  float4* rpos = dyn;                                            // [recs_cap]
  float4* rfrc = dyn + a.recs_cap;                               // [recs_cap]
  int32_t* cst = reinterpret_cast<int32_t*>(rfrc + a.recs_cap);  // [box cells + 1]
  uint16_t* lbuf = reinterpret_cast<uint16_t*>(cst + a.cst_len);  // [W][list_cap]
  __shared__ int32_t s_w[W];
  __shared__ int32_t s_next;

  const int64_t brick = xcd_block(blockIdx.x, gridDim.x);
  if (!STAGED && !a.unstaged[brick]) return;  // uniform over the block
  BoxGeom B;
  B.init(a.s.g, 0, a.reach, brick);
  // box cells: first slots (in the list buffers for now) and record starts
  int32_t* first_u = reinterpret_cast<int32_t*>(lbuf);
  stage_box_cells<NT>(B, a.cs, first_u, cst, s_w);
  if constexpr (STAGED) {
    const bool over = cst[B.nbc] > a.recs_cap;  // uniform over the block
    if (threadIdx.x == 0) a.unstaged[brick] = over ? 1 : 0;
    if (over) return;
    // record i lies in the last box cell whose start is <= i; four records
    // per thread and step, their searches and loads interleaved
    const int nrec = cst[B.nbc];
    int top = 1;
    while (2 * top < B.nbc) top *= 2;
    constexpr int G = 4;
    for (int i0 = int(threadIdx.x); i0 < nrec; i0 += G * NT) {
      int lo[G], ii[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        ii[u] = min(i0 + u * NT, nrec - 1);
        lo[u] = 0;
      }
      for (int step = top; step > 0; step >>= 1)
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int m = lo[u] + step;
          if (m < B.nbc && cst[m] <= ii[u]) lo[u] = m;
        }
      float4 pc[G], fc[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int64_t t = int64_t(first_u[lo[u]]) + (ii[u] - cst[lo[u]]);
        pc[u] = a.s.pc32[t];
        fc[u] = a.s.fg32[t];
      }
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int i = i0 + u * NT;
        if (i >= nrec) break;
        int x, y, z;
        B.xyz(lo[u], x, y, z);
        rpos[i] = make_float4(pc[u].x, pc[u].y, pc[u].z, __uint_as_float(box_pack(x, y, z)));
        rfrc[i] = fc[u];
      }
    }
  } else {
    // the first slots move to the record table's space (the host keeps it
    // at least 4 bytes a box cell), the records stay in HBM
    int32_t* fr = reinterpret_cast<int32_t*>(rpos);
    for (int q = int(threadIdx.x); q < B.nbc; q += NT) fr[q] = first_u[q];
  }
  if (threadIdx.x == 0) s_next = W;
  __syncthreads();
  const int32_t* first = reinterpret_cast<const int32_t*>(rpos);  // unstaged bricks

  const int lane = threadIdx.x & 63, wv = int(threadIdx.x >> 6);
  const int lw = lane / NQ1;  // element of the lane within its group
  const int qa = lane - lw * NQ1;
  const int grp = lane - qa;  // first lane of the element
  uint16_t* const mine = lbuf + wv * a.list_cap;
  const int rpl = (a.nrows + NQ1 - 1) / NQ1;  // stencil rows each lane of an element walks
  const int rlo = min(a.nrows, qa * rpl), rhi = min(a.nrows, rlo + rpl);
  // staged bricks: every element's list length first (each lane of an
  // element counting its share of the rows), then the elements by length,
  // longest first, so that a group's lists are alike (a group loops as
  // long as its longest list; where an element is summed does not change
  // its bits)
  __shared__ uint16_t s_cnt[STAGED ? kBrickCells * NQ1 : 1];
  __shared__ uint16_t s_order[STAGED ? kBrickCells : 1];
  __shared__ int32_t s_hist[STAGED ? kSortBuckets : 1];
  if constexpr (STAGED) {
    for (int g0 = wv; g0 < NG; g0 += W) {
      const int l = g0 * EPW + lw;
      int c = 0, ci, cj, ck;
      int64_t e;
      if (lw < EPW && l < kBrickCells && brick_elem(B, l, ci, cj, ck, e))
        box_walk_rows(B, a.rows, rlo, rhi, ci, cj, ck, [&](int q0, int q1, uint32_t) { c += cst[q1 + 1] - cst[q0]; });
      if (lw < EPW && l < kBrickCells) s_cnt[l * NQ1 + qa] = uint16_t(c);
    }
    for (int i = int(threadIdx.x); i < kSortBuckets; i += NT) s_hist[i] = 0;
    __syncthreads();
    auto bucket = [&](int l) {
      int t = 0;
#pragma unroll
      for (int k = 0; k < NQ1; ++k) t += s_cnt[l * NQ1 + k];
      return kSortBuckets - 1 - min(t, kSortBuckets - 1);
    };
    for (int l = int(threadIdx.x); l < kBrickCells; l += NT) atomicAdd(&s_hist[bucket(l)], 1);
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the buckets, kSortBuckets / 64 a lane
      constexpr int PL = kSortBuckets / 64;
      int v[PL], sum = 0;
#pragma unroll
      for (int k = 0; k < PL; ++k) {
        v[k] = s_hist[lane * PL + k];
        sum += v[k];
      }
      int incl = sum;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
      }
      int run = incl - sum;
#pragma unroll
      for (int k = 0; k < PL; ++k) {
        s_hist[lane * PL + k] = run;
        run += v[k];
      }
    }
    __syncthreads();
    for (int l = int(threadIdx.x); l < kBrickCells; l += NT) s_order[atomicAdd(&s_hist[bucket(l)], 1)] = uint16_t(l);
    __syncthreads();
  }

  // ---- per wave from here on ----
  const float gx = a.s.goff[0][qa], gy = a.s.goff[1][qa], gz = a.s.goff[2][qa];
  const float hx = B.h[0], hy = B.h[1], hz = B.h[2];
  const int per = 3 * n_u;
  int g = wv;
  while (g < NG) {  // uniform over the wave
    const int ng = min(EPW, kBrickCells - g * EPW);
    for (int s0 = 0; s0 < ng;) {  // passes over the group's elements, as many as fit the buffer each
      const int le = s0 + lw;
      int ci = 0, cj = 0, ck = 0;
      int64_t e = 0;
      const int l = lw < EPW && le < ng ? (STAGED ? int(s_order[g * EPW + le]) : g * EPW + le) : 0;
      const bool elem = lw < EPW && le < ng && brick_elem(B, l, ci, cj, ck, e);
      int take = ng - s0, n = 0, base = 0;
      if constexpr (STAGED) {
        // (1) the lane's share of the element's list; the exclusive scan
        // over the wave is where it starts (lists element by element, an
        // element's rows in lane order)
        const int c = elem ? int(s_cnt[l * NQ1 + qa]) : 0;
        int incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(incl, o, 64);
          if (lane >= o) incl += y;
        }
        const int off = incl - c;
        // (2) the elements whose lists fit (a prefix; the first always does,
        // its list being at most the box's records)
        const uint64_t fit = __ballot(lw < EPW && qa == NQ1 - 1 && incl <= a.list_cap);
        take = min(take, __popcll(fit));
        base = __shfl(off, grp, 64);
        n = __shfl(incl, grp + NQ1 - 1, 64) - base;
        // (3) the lists
        if (elem && lw < take) {
          int k = off;
          box_walk_rows(B, a.rows, rlo, rhi, ci, cj, ck, [&](int q0, int q1, uint32_t) {
            for (int r = cst[q0], re = cst[q1 + 1]; r < re; ++r) mine[k++] = uint16_t(r);
          });
        }
        wave_lds_sync();
      }
      const bool active = elem && lw < take;
      // (4) the sums, k_spread's arithmetic
      float p[NB][3], tot[NB][3];
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int j = 0; j < 3; ++j) p[b][j] = tot[b][j] = 0.f;
      if (active) {
        const uint32_t kb = 0x808080u - box_pack(ci - B.ox, cj - B.oy, ck - B.oz);
        auto one = [&](const float4& xt, const float4& f, int k) {
          float dx, dy, dz;
          box_diff(__float_as_uint(xt.w), kb, dx, dy, dz);
          const float r0 = __builtin_fmaf(dx, hx, xt.x);
          const float r1 = __builtin_fmaf(dy, hy, xt.y);
          const float r2 = __builtin_fmaf(dz, hz, xt.z);
          const Factors<float, NQ1> F =
              particle_factors<float, NQ1>(r0 - gx, r1 - gy, r2 - gz, a.s.alpha2, a.s.alpha3_pi_23, grp);
          add_particle<float, NQ1>(F, f, p);
          if ((k & (kMergeEvery - 1)) == kMergeEvery - 1) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
              for (int j = 0; j < 3; ++j) {
                tot[b][j] += p[b][j];
                p[b][j] = 0.f;
              }
          }
        };
        if constexpr (STAGED) {
          // the next record is read while the current one is summed, its
          // list entry one step earlier; two records in turn, so no copy
          const uint16_t* lst = mine + base;
          if (n > 0) {
            const int r0 = lst[0];
            int e2 = lst[min(1, n - 1)];
            float4 xa = rpos[r0], fa = rfrc[r0], xb, fb;
            int k = 0;
            for (; k + 1 < n; k += 2) {
              xb = rpos[e2];
              fb = rfrc[e2];
              e2 = lst[min(k + 2, n - 1)];
              one(xa, fa, k);
              xa = rpos[e2];
              fa = rfrc[e2];
              e2 = lst[min(k + 3, n - 1)];
              one(xb, fb, k + 1);
            }
            if (k < n) one(xa, fa, k);
          }
        } else {
          int k = 0;
          box_walk(B, a.rows, a.nrows, ci, cj, ck, [&](int q0, int q1, uint32_t) {
            for (int q = q0; q <= q1; ++q) {
              int x, y, z;
              B.xyz(q, x, y, z);
              const float pk = __uint_as_float(box_pack(x, y, z));
              for (int t = first[q], te = t + (cst[q + 1] - cst[q]); t < te; ++t) {
                float4 xt = a.s.pc32[t];
                xt.w = pk;
                one(xt, a.s.fg32[t], k);
                ++k;
              }
            }
          });
        }
      }
      // (5) rho, and the contraction (the exchange runs on every lane; an
      // element's lanes are active together)
      float v[NB][3];
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int j = 0; j < 3; ++j) v[b][j] = tot[b][j] + p[b][j];
      if (active && rho) {
        double* const rho_e = rho + (e * NQP + qa) * 3;  // Gauss point qa + NQ1 b
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int j = 0; j < 3; ++j) rho_e[NQ1 * b * 3 + j] = double(v[b][j]);
      }
      if (Fe) brick_contract<NQ1>(v, qa, grp, IF, active ? n_u : 0, a.s.fe_node_major, Fe + e * per);
      s0 += take;
      if constexpr (STAGED) wave_lds_sync();  // the lists are read before the next pass writes
    }
    int nx = 0;
    if (lane == 0) nx = atomicAdd(&s_next, 1);
    g = __shfl(nx, 0, 64);
  }
}

// the staged launch: 4 waves per SIMD (at most 128 VGPRs), two blocks per CU
template <int NQ1>
__global__ void __launch_bounds__(kBrickWaves * 64) __attribute__((amdgpu_waves_per_eu(4)))
k_spread_brick(BrickSpreadArgs a, const double* __restrict__ IF, int n_u, double* __restrict__ rho,
               double* __restrict__ Fe) {
  spread_brick<NQ1, true>(a, IF, n_u, rho, Fe);
}

// the flagged bricks, records from HBM (registers as the walk needs them)
template <int NQ1>
__global__ void __launch_bounds__(kBrickWaves * 64)
k_spread_brick_hbm(BrickSpreadArgs a, const double* __restrict__ IF, int n_u, double* __restrict__ rho,
                   double* __restrict__ Fe) {
  spread_brick<NQ1, false>(a, IF, n_u, rho, Fe);
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

template <class T>
SpreadArgs<T> spread_args(const ggem_ctx* ctx, const double* gp, bool node_major) {
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
  return a;
}

template <class T, int NQ1>
hipError_t launch_sp(ggem_ctx* ctx, int c, const double* gp, const double* IF_d, int n_u, double* rho,
                     double* Fe, bool node_major) {
  const SpreadArgs<T> a = spread_args<T>(ctx, gp, node_major);
  if (c < 0 || c >= kNElem) return hipErrorInvalidValue;
  switch (kWaveChoices[c]) {
    case 2: launch_spk<T, NQ1, 2>(ctx, a, IF_d, n_u, rho, Fe); break;
    case 1: launch_spk<T, NQ1, 1>(ctx, a, IF_d, n_u, rho, Fe); break;
    default: launch_spk<T, NQ1, 3>(ctx, a, IF_d, n_u, rho, Fe); break;
  }
  return hipGetLastError();
}

// LDS of one brick block: record table, box cell table, list buffers
struct BrickPlan {
  int32_t recs_cap, cst_len, list_cap;
  size_t lds;
};

constexpr int kBrickStaticLds = 5248;  // k_spread_brick's static LDS (list lengths, element order), rounded up
constexpr int kBrickListMin = 1536;  // list entries of a wave's buffer (a group's lists, 21 x ~40 at 1M beads)

int lds_per_block(ggem_ctx* ctx) {
  if (ctx->lds_block <= 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess || v <= 0)
      v = 65536;
    ctx->lds_block = v;
  }
  return ctx->lds_block;
}

bool brick_plan(ggem_ctx* ctx, int precision, int nqp_1d, BrickPlan& P) {
  if (precision != GGEM_FP32C || (nqp_1d != 2 && nqp_1d != 3) || ctx->nsys > 1 || explicit_minimg(ctx))
    return false;
  int nbc = 1;
  for (int d = 0; d < 3; ++d) {
    const int ext = (1 << kBrickBits) + 2 * ctx->reach[d];
    if (ext > (1 << kBoxBits)) return false;  // box cells are packed 6 bits an axis
    nbc *= ext;
  }
  if (nbc > kBoxMaxCells) return false;
  P.cst_len = (nbc + 1 + 3) & ~3;
  // records: the most crowded staged box of the pair sums' layout when it
  // describes the current lists, else 1.25 x the mean box population; a
  // brick with more reads its records from HBM
  int64_t want;
  if (ctx->box_built && ctx->box_valid && !ctx->box_fp64 && ctx->box_max > 0)
    want = ctx->box_max;
  else
    want = int64_t(1.25 * double(ctx->n) * nbc / double(std::max<int64_t>(ctx->grid.n_elem, 1))) + 64;
  // an unstaged brick keeps its first slots in the record table (4 bytes a
  // box cell); the list buffers hold them while the cells are staged and
  // take at least one whole list each (a list has at most the box's records)
  const int64_t min_recs = (nbc + 7) / 8;
  auto lists = [&](int64_t r) {
    return (std::max({r, int64_t(kBrickListMin), int64_t(nbc + kBrickWaves / 2 - 1) / (kBrickWaves / 2)}) + 7) & ~7;
  };
  auto bytes = [&](int64_t r) {
    return int64_t(2 * sizeof(float4)) * r + int64_t(sizeof(int32_t)) * P.cst_len +
           int64_t(sizeof(uint16_t)) * kBrickWaves * lists(r);
  };
  // the table takes what two blocks per CU leave (more LDS would not cost
  // occupancy), at least the wanted size when one block per CU holds it
  const int64_t limit = int64_t(lds_per_block(ctx)) - kBrickStaticLds;
  const int64_t two = std::min<int64_t>(limit, int64_t(lds_per_block(ctx)) / 2 - kBrickStaticLds);
  int64_t cap = std::min<int64_t>(std::max(want, min_recs), 65535);
  while (cap > min_recs && bytes(cap) > limit) cap = std::max(min_recs, cap - std::max<int64_t>(1, cap / 8));
  if (bytes(cap) > limit) return false;
  while (cap < 65535 && bytes(cap + 64) <= two) cap += 64;
  P.recs_cap = int32_t(cap);
  P.list_cap = int32_t(lists(cap));
  P.lds = size_t(bytes(cap));
  return true;
}

template <int NQ1>
hipError_t launch_brick(ggem_ctx* ctx, const BrickPlan& P, const double* gp, const double* IF_d, int n_u,
                        double* rho, double* Fe, bool node_major) {
  BrickSpreadArgs a;
  a.s = spread_args<float>(ctx, gp, node_major);
  a.rows = ctx->rows_d;
  a.nrows = ctx->n_rows;
  for (int d = 0; d < 3; ++d) a.reach[d] = ctx->reach[d];
  a.cs = ctx->cell_start;
  a.recs_cap = P.recs_cap;
  a.cst_len = P.cst_len;
  a.list_cap = P.list_cap;
  a.unstaged = ctx->spread_unstaged;
  const unsigned nb = unsigned(ctx->grid.n_cells >> (3 * kBrickBits));
  if (!a.unstaged || ctx->spread_unstaged_cap < nb) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_spread_brick<NQ1>), dim3(nb), dim3(kBrickWaves * 64), P.lds, ctx->stream, a, IF_d, n_u,
                     rho, Fe);
  hipLaunchKernelGGL((k_spread_brick_hbm<NQ1>), dim3(nb), dim3(kBrickWaves * 64), P.lds, ctx->stream, a, IF_d, n_u,
                     rho, Fe);
  return hipGetLastError();
}

}  // namespace

// The brick kernel walks every element's stencil rows twice per launch;
// the element lists walk them once per build.  At reach 2 (21 rows, 1M
// beads: 1.09 ms against 1.08 ms plus the 0.37 ms list build) the brick
// kernel is the default; at larger reach (the slit's 69 rows: 0.65 ms
// against 0.30 ms) the element-list kernel stays the default and the brick
// kernel an autotuner candidate.
static bool brick_default(const ggem_ctx* ctx) {
  for (int d = 0; d < 3; ++d)
    if (ctx->reach[d] > 2) return false;
  return true;
}

int spread_brick_mode(ggem_ctx* ctx, int precision, int nqp_1d) {
  BrickPlan P;
  if (!brick_plan(ctx, precision, nqp_1d, P)) return 0;
  return brick_default(ctx) ? 2 : 1;
}

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
  // candidates: the block shapes of k_spread, then (FP32C, when its box
  // fits) the brick kernel, the default at reach <= 2; all give the same
  // bits
  BrickPlan bp;
  const bool brick = brick_plan(ctx, precision, nqp_1d, bp);
  auto run = [&](int c) {
    if (c == kNElem) {
      if (!brick) return hipErrorInvalidValue;
      return nqp_1d == 2 ? launch_brick<2>(ctx, bp, gp, IF_d, n_u, rho, Fe, node_major)
                         : launch_brick<3>(ctx, bp, gp, IF_d, n_u, rho, Fe, node_major);
    }
    if (nqp_1d == 2)
      return precision == GGEM_FP64 ? launch_sp<double, 2>(ctx, c, gp, IF_d, n_u, rho, Fe, node_major)
                                    : launch_sp<float, 2>(ctx, c, gp, IF_d, n_u, rho, Fe, node_major);
    return precision == GGEM_FP64 ? launch_sp<double, 3>(ctx, c, gp, IF_d, n_u, rho, Fe, node_major)
                                  : launch_sp<float, 3>(ctx, c, gp, IF_d, n_u, rho, Fe, node_major);
  };
  // rho and Fe are outputs only: idempotent
  int c = tune_pick(ctx, tune_key("spread", precision, nqp_1d, ctx->grid.n_elem), brick ? kNElem + 1 : kNElem, run,
                    brick && brick_default(ctx) ? kNElem : 0);
  if (c == kNElem && !brick) c = 0;  // a pick recorded while the brick kernel applied
  ProfScope ps(ctx, precision == GGEM_FP64 ? "spread_fp64" : "spread_fp32c");
  return run(c);
}

}  // namespace ggem
