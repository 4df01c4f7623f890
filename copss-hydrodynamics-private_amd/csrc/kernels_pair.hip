// kernels_pair.hip -- (b) the local regularized-Stokeslet pair sum
// u_i = sum_{j in nbr(i)} G_reg(x_j - x_i) f_j, and the mobility and
// Chebyshev applications built on it.
//
// Reference: GGEMStokes::local_solution_bead (ggem_stokes.C:427-503),
// local_solution_field (:360-424), green_tensor_local_regularized
// (:304-357), global_self_exclusion (:506-523), the velocity combination of
// PMSystemStokes::compute_point_velocity (pm_system_stokes.C:441-457) and the
// Chebyshev vector updates (brownian_system.C:1037-1041, :1093-1101).
//
// The targets' neighbour lists are stored sliced by wave (entry k of lane l
// of slice q at slice_off[q] + 64k + l): every list access of a wave is one
// coalesced transaction.  Positions do not move between neighbour builds,
// so pair tensors are evaluated once per build and streamed by every later
// application (the local velocity, each Chebyshev term, M y).  FP64:
// k_tensors stores G = A I + sgn w w^T, w = sqrt(|Bx|) x_ij (32 bytes per
// entry) in list order and k_apply streams them, gathering one operand
// record per neighbour.  FP32C, one operand: per brick of cells the
// neighbours' records are staged in LDS and the entries stream a 2-byte
// record number and the 8-byte (A, Bx) (k_box_fill, k_apply_box, below).
// FP32C, several operands or the explicit minimum image: k_apply evaluates
// (A, Bx) per pair from the coefficient table in LDS.  Blocks are
// renumbered per XCD so a chip partition keeps a contiguous stretch of
// cells and their neighbours in its L2.
//
// FP64: the reference's formula with libm-class erf/exp.  FP32C: x_ij from
// the cell-relative FP32 position records ((cell difference) h + offset
// difference), (A, Bx) from a piecewise polynomial in r^2, partial sums in
// FP32 merged every 8 neighbours into a Kahan accumulator; the epilogue
// (combination, recurrence) runs in FP64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ggem_device.h"

namespace ggem {

namespace {

constexpr int kBlock = 256;
constexpr int kMergeEvery = 8;

inline unsigned blocks_for(int64_t n, int b = kBlock) { return unsigned((n + b - 1) / b); }

struct PairArgs {
  Grid g;
  GGEMConst k;
  GGEMConst32 k32;
  const double4* x4;
  const int32_t* cnt;
  const int64_t* slice_off;
  const uint32_t* ell;
  const int32_t* perm;
  const double* fs;  // own forces (FP64 SoA) for the mobility epilogue
  int64_t n;
  double bsign;
  const SysGrid* sys;       // batched contexts: grid of each system
  const int32_t* sys_slot;  // and system of each slot
  const float4* pc32;       // FP32C position records (cell-relative offsets, packed cell)
  int32_t xcd;              // renumber blocks per XCD (xcd_block)
  int32_t exact;            // explicit minimum image (no list shift codes)
  const float4* tab;        // FP32C (A, Bx) as a piecewise polynomial in r^2, 4 float4 per interval
  CoefTab ct;
};

// FP32C (A, Bx) at s = r^2 from the interval table (copied to LDS): the
// interval and y = 2 (s - mid)/d are formed exactly (widths are powers of
// two), then two degree-7 Horner chains; past the cut-off the last
// interval is zero.  One body for every FP32C kernel, so they all
// evaluate bit-identical tensors.
// The interval of s and y in it.  The last interval (index t.n - 1) is the
// zero one: its pairs are skipped by every FP32C pair sum (fp32c_kept),
// which add exactly the same terms in the same order whether they store the
// kept pairs or test each entry.
__device__ __forceinline__ int fp32c_interval(const CoefTab& t, float s, float& y) {
  const bool r2 = s >= t.s1;
  const float base = r2 ? t.s1 : 0.f;
  const float inv = r2 ? t.inv2 : t.inv1;
  const float d = r2 ? t.d2 : t.d1;
  const int j = min(int((s - base) * inv), r2 ? t.n2 : t.n1 - 1);
  const float mid = __builtin_fmaf(float(j) + 0.5f, d, base);
  y = (s - mid) * (2.f * inv);
  return (r2 ? t.n1 : 0) + j;
}
__device__ __forceinline__ bool fp32c_kept(const CoefTab& t, int i) { return i < t.n - 1; }
// fp32c_interval for a table of one width (n2 = 0, so s1 lies at or past
// the cut): the same arithmetic below s1; s >= s1 gives the zero interval
__device__ __forceinline__ int fp32c_interval1(const CoefTab& t, float s, float& y) {
  const int j = min(int(s * t.inv1), t.n1);
  const float mid = __builtin_fmaf(float(j) + 0.5f, t.d1, 0.f);
  y = (s - mid) * (2.f * t.inv1);
  return j;
}

// (A, Bx) at y in interval i of the table's LDS copy (load_tab: the four
// coefficient quadruples of an interval padded to kTabStride, so that lanes
// at different intervals spread over the LDS banks and the four reads
// share one address)
constexpr int kTabStride = 5;
__device__ __forceinline__ void fp32c_eval(const float4* __restrict__ tab, int i, float y, float& A, float& Bx) {
  const float4* t = tab + __mul24(i, kTabStride);  // (a full-rate multiply)
  const float4 a0 = t[0], a1 = t[1], b0 = t[2], b1 = t[3];
  float pa = __builtin_fmaf(a1.w, y, a1.z);
  pa = __builtin_fmaf(pa, y, a1.y);
  pa = __builtin_fmaf(pa, y, a1.x);
  pa = __builtin_fmaf(pa, y, a0.w);
  pa = __builtin_fmaf(pa, y, a0.z);
  pa = __builtin_fmaf(pa, y, a0.y);
  A = __builtin_fmaf(pa, y, a0.x);
  float pb = __builtin_fmaf(b1.w, y, b1.z);
  pb = __builtin_fmaf(pb, y, b1.y);
  pb = __builtin_fmaf(pb, y, b1.x);
  pb = __builtin_fmaf(pb, y, b0.w);
  pb = __builtin_fmaf(pb, y, b0.z);
  pb = __builtin_fmaf(pb, y, b0.y);
  Bx = __builtin_fmaf(pb, y, b0.x);
}

// r^2 of an FP32C displacement, one fixed evaluation order
__device__ __forceinline__ float fp32c_r2(float dx, float dy, float dz) {
  return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}

// p += A f + Bx (x . f) x with every product fused explicitly, so that no
// kernel is left a choice of contraction: all FP32C pair sums add
// bit-identical terms
__device__ __forceinline__ void fp32c_acc(float A, float Bx, float dx, float dy, float dz, float fx, float fy,
                                          float fz, float& p0, float& p1, float& p2) {
  const float b = Bx * __builtin_fmaf(dz, fz, __builtin_fmaf(dy, fy, dx * fx));
  p0 += __builtin_fmaf(A, fx, b * dx);
  p1 += __builtin_fmaf(A, fy, b * dy);
  p2 += __builtin_fmaf(A, fz, b * dz);
}

// G f = A f + sgn (w . f) w with the stored FP64 (A, w) = c, every product
// fused explicitly, so that all FP64 pair sums add bit-identical terms
__device__ __forceinline__ void fp64_acc(const double4& c, double sg, const double4& f, double& p0, double& p1,
                                         double& p2) {
  const double wf = sg * __builtin_fma(c.w, f.z, __builtin_fma(c.z, f.y, c.y * f.x));
  p0 += __builtin_fma(c.x, f.x, wf * c.y);
  p1 += __builtin_fma(c.x, f.y, wf * c.z);
  p2 += __builtin_fma(c.x, f.z, wf * c.w);
}

// the interval table into LDS: quadruple c of interval i (a.tab[4 i + c])
// to dst[kTabStride i + c] (the whole block takes part)
__device__ __forceinline__ void load_tab(float4* dst, const PairArgs& a) {
  for (int q = threadIdx.x; q < 4 * a.ct.n; q += blockDim.x) dst[kTabStride * (q >> 2) + (q & 3)] = a.tab[q];
}

// Launch shapes the autotuner chooses from (index 0 is the default); every
// shape computes the same per-slot arithmetic, so results do not depend on
// the choice.
struct Shape {
  int block, xcd;
};
constexpr Shape kApplyShapes[] = {{256, 1}, {128, 1}, {512, 1}, {64, 1}, {256, 0}};
constexpr int kNApply = sizeof(kApplyShapes) / sizeof(kApplyShapes[0]);
constexpr Shape kTensorShapes[] = {{256, 1}, {128, 1}, {512, 1}, {64, 1}};
constexpr int kNTensor = sizeof(kTensorShapes) / sizeof(kTensorShapes[0]);

__device__ __forceinline__ int64_t block_index(int xcd) {
  return xcd ? xcd_block(blockIdx.x, gridDim.x) : int64_t(blockIdx.x);
}

// Grid of a slot's system.
template <bool BATCH>
__device__ __forceinline__ const Grid& slot_grid(const PairArgs& a, int64_t s) {
  if constexpr (BATCH) return a.sys[a.sys_slot[s]].g;
  else return a.g;
}

template <class T>
struct V4;
template <>
struct V4<float> {
  using type = float4;
};
template <>
struct V4<double> {
  using type = double4;
};

// FP64 pair tensors of all list entries: (A, w) with G = A I + sgn w w^T.
template <bool EXACT, bool BATCH, int BLK>
__global__ void __launch_bounds__(BLK) k_tensors(PairArgs a, double4* __restrict__ c64) {
  const int64_t s = block_index(a.xcd) * BLK + threadIdx.x;
  if (s >= a.n) return;
  const Grid& g = slot_grid<BATCH>(a, s);
  const int cnt = a.cnt[s];
  const int64_t base = a.slice_off[s >> 6] + (s & 63);
  const uint32_t* lst = a.ell + base;
  const double4 xs = a.x4[s];
  const int last = cnt > 0 ? cnt - 1 : 0;
  uint32_t e1 = cnt > 0 ? lst[0] : 0u;
  uint32_t e2 = cnt > 1 ? lst[kSlice] : e1;
  double4 xa = a.x4[e1 & kSlotMask];
  for (int k = 0; k < cnt; ++k) {
    const uint32_t e = e1;
    const double4 xt = xa;
    e1 = e2;
    xa = a.x4[e1 & kSlotMask];
    e2 = lst[uint32_t(min(k + 2, last)) * uint32_t(kSlice)];
    double dx, dy, dz;
    pair_disp<EXACT>(g, xs, xt, e >> kSlotBits, dx, dy, dz);
    const int64_t o = base + int64_t(k) * kSlice;
    double A, Bx;
    green_coef_f64(a.k, dx * dx + dy * dy + dz * dz, A, Bx);
    const double sb = sqrt(fabs(Bx));
    c64[o] = make_double4(A, sb * dx, sb * dy, sb * dz);
  }
}

// Epilogue of the pair sum of target slot s, m = acc (see Epi).
template <int NR, int EPI>
__device__ __forceinline__ void apply_epilogue(const PairArgs& a, int64_t s, const double (&acc)[NR][3],
                                               const double* __restrict__ xsoa, const SlotVec& v,
                                               const ChebEpi& epi, const double* __restrict__ u_glob,
                                               double* __restrict__ out) {
  const int64_t n = a.n;
  if (EPI == EPI_LOCAL || EPI == EPI_MOBILITY) {
    // forces are the operand: own force from the FP64 store
    const double fo[3] = {a.fs[s], a.fs[n + s], a.fs[2 * n + s]};
    double u[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      u[q] = acc[0][q];
      if (a.k.point_type == GGEM_LAGRANGIAN_POINT) u[q] += a.k.coeff_2 * fo[q];  // self term
    }
    const int64_t p = a.perm[s];
    if (EPI == EPI_MOBILITY) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double ug = u_glob ? u_glob[3 * p + q] : 0.0;
        if (a.k.point_type == GGEM_POLYMER_BEAD) u[q] = ug + u[q] - a.k.coeff_1 * fo[q] + fo[q];
        else if (a.k.point_type == GGEM_LAGRANGIAN_POINT) u[q] = ug + u[q];
        else u[q] = 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) out[3 * p + q] = u[q];
  } else {
    double da = epi.da, db = epi.db, ck = epi.c, cm1 = epi.cm1, c0 = epi.c0;
    if (epi.dpar) {
      const double* pp = epi.dpar + int64_t(epi.sys ? epi.sys[s] : 0) * epi.pstride;
      da = pp[0];
      db = pp[1];
      ck = pp[2 + epi.term];
      if (epi.ymask & 1) cm1 = pp[epi.term];
      if (epi.ymask & 2) c0 = pp[1 + epi.term];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double nv[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int64_t idx = int64_t(r) * 3 * n + int64_t(q) * n + s;
        const double xv = xsoa[idx];
        double m = acc[r][q];
        if (a.k.point_type == GGEM_POLYMER_BEAD) m = m - a.k.coeff_1 * xv + xv;
        else if (a.k.point_type == GGEM_LAGRANGIAN_POINT) m = m + a.k.coeff_2 * xv;
        else m = 0.0;
        if (epi.kind == 0) {
          nv[q] = m;
        } else if (epi.kind == 1) {
          nv[q] = db * xv + da * m;
        } else {
          const double xm = epi.xm1.soa[idx];
          double tv = 2. * db * xv + 2. * da * m;
          tv += -1.0 * xm;
          nv[q] = tv;
          if (epi.ymask == 4) {
            epi.y[idx] += ck * tv;
          } else if (epi.ymask) {  // x_{k-1}, x_k and x_{k+1} terms in one update of y
            double add = 0.0;
            if (epi.ymask & 1) add += cm1 * xm;
            if (epi.ymask & 2) add += c0 * xv;
            if (epi.ymask & 4) add += ck * tv;
            epi.y[idx] += add;
          }
        }
        v.soa[idx] = nv[q];
      }
      if (v.g32) v.g32[r * n + s] = make_float4(float(nv[0]), float(nv[1]), float(nv[2]), 0.f);
      if (v.g64) v.g64[r * n + s] = make_double4(nv[0], nv[1], nv[2], 0.0);
    }
  }
}

// Output of one target: m = sum over its list of G_ij x_j for NR operand
// vectors, then the epilogue.
template <int P, int NR, int EPI, int BLK>
__global__ void __launch_bounds__(BLK) k_apply(PairArgs a, const void* __restrict__ coef_raw,
                                                 const double4* __restrict__ xg64,
                                                 const float4* __restrict__ xg32,
                                                 const double* __restrict__ xsoa, SlotVec v,
                                                 ChebEpi epi, const double* __restrict__ u_glob,
                                                 double* __restrict__ out) {
  // (the tensor pointer is untyped in the signature: a type computed from
  // the header's unnamed precision enum would be mangled differently by the
  // host and device compilations; FP32C keeps no tensors)
  if (epi.skip && *epi.skip) return;  // uniform over the grid
  const int64_t s = block_index(a.xcd) * BLK + threadIdx.x;
  const float4* tab = nullptr;
  if constexpr (P != GGEM_FP64) {
    __shared__ float4 s_tab[kTabStride * kTabMaxIntervals];
    load_tab(s_tab, a);
    __syncthreads();
    tab = s_tab;
  }
  if (s >= a.n) return;
  const int64_t n = a.n;
  const int cnt = a.cnt[s];
  const int64_t base = a.slice_off[s >> 6] + (s & 63);
  const uint32_t* lst = a.ell + base;
  double acc[NR][3];
  // Groups of U entries: the entries of the next group are loaded one group
  // ahead, and the U tensors and operand records of a group are all
  // requested before the first is used (memory-level parallelism for the
  // dependent gathers).  The summation order is the list order.
  constexpr int U = NR >= 4 ? 1 : (NR == 2 ? 2 : 4);
  const int last = cnt > 0 ? cnt - 1 : 0;

  if constexpr (P == GGEM_FP64) {
    uint32_t en[U];
#pragma unroll
    for (int i = 0; i < U; ++i) en[i] = cnt > 0 ? lst[uint32_t(min(i, last)) * uint32_t(kSlice)] : 0u;
    const double4* __restrict__ cf = static_cast<const double4*>(coef_raw) + base;
    const double4* __restrict__ xg = xg64;
    const double sg = a.bsign;
    double ps[NR][3];
#pragma unroll
    for (int r = 0; r < NR; ++r) ps[r][0] = ps[r][1] = ps[r][2] = 0.0;
    // G f = A f + sgn (w . f) w with the stored (A, w)
    auto pair_acc = [&](const double4& c, const double4* f) {
#pragma unroll
      for (int r = 0; r < NR; ++r) fp64_acc(c, sg, f[r], ps[r][0], ps[r][1], ps[r][2]);
    };
    int k = 0;
    for (; k + U <= cnt; k += U) {
      uint32_t e[U];
      double4 c[U];
#pragma unroll
      for (int i = 0; i < U; ++i) {
        e[i] = en[i];
        en[i] = lst[uint32_t(min(k + U + i, last)) * uint32_t(kSlice)];
        c[i] = cf[int64_t(k + i) * kSlice];
      }
      double4 f[U][NR];
#pragma unroll
      for (int i = 0; i < U; ++i)
#pragma unroll
        for (int r = 0; r < NR; ++r) f[i][r] = xg[r * n + (e[i] & kSlotMask)];
#pragma unroll
      for (int i = 0; i < U; ++i) pair_acc(c[i], f[i]);
    }
#pragma unroll
    for (int i = 0; i < U - 1; ++i) {
      if (k + i < cnt) {  // entries en[0 .. cnt - k)
        double4 f[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) f[r] = xg[r * n + (en[i] & kSlotMask)];
        pair_acc(cf[int64_t(k + i) * kSlice], f);
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) acc[r][q] = ps[r][q];
  } else {
    // FP32C: G f = A f + Bx (x . f) x with x_ij from the cell-relative FP32
    // positions, x_ij = (cell_t - cell_s + shift) h + (offset_t - offset_s)
    // per axis, and (A, Bx) from the table at r^2
    const float4* __restrict__ xg = xg32;
    const float4* __restrict__ pc = a.pc32;
    const Grid& g = a.sys ? a.sys[a.sys_slot[s]].g : a.g;
    uint32_t en[U];  // entries of the next group
#pragma unroll
    for (int i = 0; i < U; ++i) en[i] = cnt > 0 ? lst[uint32_t(min(i, last)) * uint32_t(kSlice)] : 0u;
    const float hx = float(g.h[0]), hy = float(g.h[1]), hz = float(g.h[2]);
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const float Lx = float(g.L[0]), Ly = float(g.L[1]), Lz = float(g.L[2]);
    const int px = g.periodic[0], py = g.periodic[1], pz = g.periodic[2];
    const float4 po = pc[s];
    const uint32_t co = __float_as_uint(po.w);
    const int io = int(co & ((1u << kCellBitsX) - 1));
    const int jo = int((co >> kCellBitsX) & ((1u << kCellBitsY) - 1));
    const int ko = int(co >> (kCellBitsX + kCellBitsY));
    float ps[NR][3];
    KahanF kh[NR][3];
#pragma unroll
    for (int r = 0; r < NR; ++r) ps[r][0] = ps[r][1] = ps[r][2] = 0.f;
    auto disp = [&](uint32_t e, const float4& pt, float& dx, float& dy, float& dz) {
      const uint32_t ct = __float_as_uint(pt.w);
      int di = int(ct & ((1u << kCellBitsX) - 1)) - io;
      int dj = int((ct >> kCellBitsX) & ((1u << kCellBitsY) - 1)) - jo;
      int dk = int(ct >> (kCellBitsX + kCellBitsY)) - ko;
      const uint32_t code = e >> kSlotBits;
      if (!a.exact) {  // the entry's single periodic shift, in cells
        di += (int(code & 1u) - int((code >> 1) & 1u)) * nx;
        dj += (int((code >> 2) & 1u) - int((code >> 3) & 1u)) * ny;
        dk += (int((code >> 4) & 1u) - int((code >> 5) & 1u)) * nz;
      } else {  // the nearest image of the cell difference, exact in integers
        if (px) di = 2 * di > nx ? di - nx : (2 * di < -nx ? di + nx : di);
        if (py) dj = 2 * dj > ny ? dj - ny : (2 * dj < -ny ? dj + ny : dj);
        if (pz) dk = 2 * dk > nz ? dk - nz : (2 * dk < -nz ? dk + nz : dk);
      }
      dx = __builtin_fmaf(float(di), hx, pt.x - po.x);
      dy = __builtin_fmaf(float(dj), hy, pt.y - po.y);
      dz = __builtin_fmaf(float(dk), hz, pt.z - po.z);
      if (a.exact) {  // minimum image (pm_periodic_boundary.C:350-374)
        if (px) dx = dx <= -0.5f * Lx ? dx + Lx : (dx > 0.5f * Lx ? dx - Lx : dx);
        if (py) dy = dy <= -0.5f * Ly ? dy + Ly : (dy > 0.5f * Ly ? dy - Ly : dy);
        if (pz) dz = dz <= -0.5f * Lz ? dz + Lz : (dz > 0.5f * Lz ? dz - Lz : dz);
      }
    };
    // kept pairs are summed and merged every kMergeEvery of them; the zero
    // interval's pairs are skipped (nk counts the kept ones)
    int nk = 0;
    auto merge = [&]() {
      if ((nk & (kMergeEvery - 1)) == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            kh[r][q].add(ps[r][q]);
            ps[r][q] = 0.f;
          }
      }
    };
    auto pair_acc = [&](uint32_t e, const float4& pt, const float4* f) {
      float dx, dy, dz;
      disp(e, pt, dx, dy, dz);
      float y;
      const int it = fp32c_interval(a.ct, fp32c_r2(dx, dy, dz), y);
      if (!fp32c_kept(a.ct, it)) return;
      float A, Bx;
      fp32c_eval(tab, it, y, A, Bx);
#pragma unroll
      for (int r = 0; r < NR; ++r) fp32c_acc(A, Bx, dx, dy, dz, f[r].x, f[r].y, f[r].z, ps[r][0], ps[r][1], ps[r][2]);
      ++nk;
      merge();
    };
    int k = 0;
    for (; k + U <= cnt; k += U) {
      uint32_t c[U];
      float4 pt[U];
#pragma unroll
      for (int i = 0; i < U; ++i) {
        c[i] = en[i];
        en[i] = lst[uint32_t(min(k + U + i, last)) * uint32_t(kSlice)];
      }
      float4 f[U][NR];
#pragma unroll
      for (int i = 0; i < U; ++i) {
        const uint32_t t = c[i] & kSlotMask;
        pt[i] = pc[t];
#pragma unroll
        for (int r = 0; r < NR; ++r) f[i][r] = xg[r * n + t];
      }
#pragma unroll
      for (int i = 0; i < U; ++i) pair_acc(c[i], pt[i], f[i]);
    }
#pragma unroll
    for (int i = 0; i < U - 1; ++i) {
      if (k + i < cnt) {  // entries en[0 .. cnt - k)
        const uint32_t t = en[i] & kSlotMask;
        float4 f[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) f[r] = xg[r * n + t];
        pair_acc(en[i], pc[t], f);
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        kh[r][q].add(ps[r][q]);
        acc[r][q] = kh[r][q].value();
      }
  }

  apply_epilogue<NR, EPI>(a, s, acc, xsoa, v, epi, u_glob, out);
}


// ---- FP32C pair sum with the neighbour records staged in LDS ----
//
// One workgroup per brick of 8 x 8 x 8 cells.  The brick's box is the brick
// widened by the stencil reach (periodic images unwrapped); each particle of
// the box is one record, in box order: box cells x-major through the box,
// particles of a cell in slot order, so a stencil run of box cells is one
// range of records.  Once per list build k_box_fill writes every staged
// brick's records (slot, cell-relative FP32 position with the packed box
// cell) and walks each target's stencil through them in list order
// (box_walk): per kept entry the neighbour's record number (16 bits).
// k_apply_box copies the records, their operands and the coefficient table
// to LDS and streams the record numbers (2 bytes per pair): per pair the
// neighbour from LDS and (A, Bx) from the table at the displacement, a
// group of entries at a time (records first, then intervals, then tensors
// and sums) so that the LDS reads of a group overlap.  The summation order
// (list order) and the arithmetic are k_apply's; the displacement (box_t -
// box_s) h + (offset_t - offset_s) has the same integer cell difference as
// the list entry's periodic shift gives, so the results are bit-identical.
// A brick whose box holds more than kBoxMaxRecs records reads its list and
// gathers the neighbours' records from global memory.
constexpr int kBoxThreads = 256;
constexpr int kBoxMaxRecs = 2048;   // records of a staged brick (64 KiB of LDS)
// A target's packed entry base: the offset of its entry 0 (below 2^40), its
// own record number (12 bits) and, in staged bricks, its count of stored
// (kept) entries (12 bits; a staged box holds at most kBoxMaxRecs records).
constexpr int kBaseBits = 40;
constexpr int64_t kBaseMask = (int64_t(1) << kBaseBits) - 1;
constexpr int kSelfShift = kBaseBits, kKeptShift = kBaseBits + 12;
constexpr uint64_t kField12 = 0xFFFu;

struct BoxArgs {
  const int32_t* cs;     // cell starts, brick order
  int32_t reach[3];      // stencil reach of a plain context
  int64_t* cnt;          // [n_bricks] records of each staged brick (k_box_count)
  int64_t* ecnt;         // [n_bricks] entries of each brick's layout (k_box_count)
  int32_t* maxrec;       // [3] largest staged record count, most targets of a brick, unstaged bricks
  const int32_t* recs;   // [n_bricks] particles in each brick's box (from the list count pass)
  const int64_t* off;    // [n_bricks + 1] first record of each brick
  const int64_t* eoff;   // [n_bricks + 1] first entry of each brick
  int32_t* tgt;          // [n] slots of each brick, longest lists first (staged bricks)
  int64_t* base;         // [n] entry 0 of the target at each position of tgt (low kBaseBits),
                         // its own record in a staged brick (above)
  uint32_t* slot;        // [records] slot of each record
  float4* pos;           // [records] offset xyz, packed box cell
  uint16_t* idx;         // [entries] neighbour's record number (staged bricks): FP32C brick layout,
                         // FP64 the lists' slice layout
  int32_t cells_cap;     // k_box_fill: box cells, waves and records of a brick its LDS tables hold
  int32_t waves_cap;
  int32_t recs_cap;
  int32_t fp64;          // FP64: every pair stored, at its list position
  const StencilRow* rows;  // stencil rows (a batched context: all systems', SysGrid::row_off)
  int32_t nrows;
};

// What the staged pair sums read of the layout (their kernel arguments
// stay small)
struct BoxView {
  const int32_t* cs;
  int32_t reach[3];
  const int64_t* off;
  const int32_t* tgt;
  const int64_t* base;
  const uint32_t* slot;
  const float4* pos;
  const uint16_t* idx;
};

// Entry layout of a brick: its targets in the order tgt, in waves of 64;
// entry k of the target at position i = 64 w + l lies at
// eoff + 64 (W_0 + .. + W_{w-1}) + 64 k + l, W_w being the longest list of
// wave w.  Staged bricks order their targets by list length, longest first
// (ties by slot), so the lanes of a wave do similar work.
constexpr int kBoxMaxWaves = kBoxMaxTargets / 64;  // waves of one brick
constexpr int kBoxSortMax = 2048;   // targets of a brick ordered in LDS

// Box of the brick holding first_slot (the slot picks the system of a
// batched context).
__device__ __forceinline__ void box_init(BoxGeom& B, const PairArgs& a, const int32_t* reach, int64_t brick,
                                         int64_t first_slot) {
  if (a.sys) {
    const SysGrid* S = a.sys + a.sys_slot[first_slot];
    B.init(S->g, S->cell_off, S->reach, brick);
  } else {
    B.init(a.g, 0, reach, brick);
  }
}

// Records of every brick's box (a brick is staged when it has targets, its
// box fits the packing and holds at most kBoxMaxRecs records), the order of
// its targets and the size of its entry layout.
__global__ void __launch_bounds__(kBoxThreads) k_box_count(PairArgs a, BoxArgs b) {
  __shared__ int32_t keys[kBoxSortMax];   // list lengths in slot order
  __shared__ int32_t skeys[kBoxSortMax];  // the same, longest first
  __shared__ int64_t s_e[kBoxThreads / 64];
  const int64_t brick = blockIdx.x;
  const int64_t c0 = brick * kBrickCells;
  const int32_t sb = b.cs[c0], se = b.cs[c0 + kBrickCells];
  const int nt = se - sb;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int total = 0;
  if (nt > 0) {  // uniform over the block
    BoxGeom B;
    box_init(B, a, b.reach, brick, sb);
    // the list count pass staged exactly the boxes that fit and counted
    // their particles
    if (B.fits()) total = b.recs[brick];
  }
  const bool staged = total > 0 && total <= kBoxMaxRecs;
  const bool sorted = staged && nt <= kBoxSortMax;
  if (sorted) {
    for (int i = int(threadIdx.x); i < nt; i += kBoxThreads) keys[i] = a.cnt[sb + i];
    __syncthreads();
    for (int i = int(threadIdx.x); i < nt; i += kBoxThreads) {
      const int ki = keys[i];
      int r = 0;
      for (int j = 0; j < nt; ++j) {
        const int kj = keys[j];
        r += (kj > ki || (kj == ki && j < i)) ? 1 : 0;
      }
      skeys[r] = ki;
      b.tgt[sb + r] = sb + i;
    }
    __syncthreads();
  } else {
    for (int i = int(threadIdx.x); i < nt; i += kBoxThreads) b.tgt[sb + i] = sb + i;
  }
  // entries: 64 x the longest list of every wave of the order
  int64_t esum = 0;
  const int nw = (nt + 63) / 64;
  for (int w = wv; w < nw; w += kBoxThreads / 64) {
    const int i = 64 * w + lane;
    const int c = i < nt ? (sorted ? skeys[i] : a.cnt[sb + i]) : 0;
    esum += 64 * int64_t(wave_max(c));
  }
  if (lane == 0) s_e[wv] = esum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t e = 0;
#pragma unroll
    for (int w = 0; w < kBoxThreads / 64; ++w) e += s_e[w];
    b.ecnt[brick] = e;
    b.cnt[brick] = staged ? total : 0;
  }
}

// Largest staged record count and most targets of any brick (one block; a
// reduction instead of an atomic per brick on the same two words).
constexpr int kMaxThreads = 1024;
__global__ void __launch_bounds__(kMaxThreads) k_box_max(BoxArgs b, int64_t nbricks) {
  __shared__ int32_t s_m[3][kMaxThreads / 64];
  int mr = 0, mt = 0, nu = 0;  // nu: bricks with targets that are not staged
  for (int64_t k = threadIdx.x; k < nbricks; k += kMaxThreads) {
    const int nt = b.cs[(k + 1) * kBrickCells] - b.cs[k * kBrickCells];
    mr = max(mr, int(b.cnt[k]));
    mt = max(mt, nt);
    nu += nt > 0 && b.cnt[k] == 0 ? 1 : 0;
  }
  mr = wave_max(mr);
  mt = wave_max(mt);
  nu = wave_sum(nu);
  if ((threadIdx.x & 63) == 0) {
    s_m[0][threadIdx.x >> 6] = mr;
    s_m[1][threadIdx.x >> 6] = mt;
    s_m[2][threadIdx.x >> 6] = nu;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kMaxThreads / 64; ++w) {
      mr = max(mr, s_m[0][w]);
      mt = max(mt, s_m[1][w]);
      nu += s_m[2][w];
    }
    b.maxrec[0] = mr;
    b.maxrec[1] = mt;
    b.maxrec[2] = nu;
  }
}

// Records of staged bricks, the entry layout of every brick's targets, and
// in staged bricks each entry's neighbour record number.
template <bool BATCH>
__global__ void __launch_bounds__(kBoxThreads) k_box_fill(PairArgs a, BoxArgs b) {
  extern __shared__ float4 dyn[];
  float4* rpos = dyn;                                                  // [recs_cap] records of a staged brick
  int32_t* cfirst = reinterpret_cast<int32_t*>(rpos + b.recs_cap);     // [cells_cap] first slot of each box cell
  int32_t* cst = cfirst + b.cells_cap;                                 // [cells_cap + 1] record start of each box cell
  int32_t* wpre = cst + b.cells_cap + 1;                               // [waves_cap + 1] entry rows before each wave
  __shared__ int32_t s_w[kBoxThreads / 64];
  const int64_t brick = blockIdx.x;
  const int64_t c0 = brick * kBrickCells;
  const int32_t sb = b.cs[c0], se = b.cs[c0 + kBrickCells];
  if (sb >= se) return;
  const int nt = se - sb;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t r0 = b.off[brick];
  const int nrec = int(b.off[brick + 1] - r0);
  BoxGeom B;
  box_init(B, a, b.reach, brick, sb);
  if (nrec > 0) {
    // (1) first slots and record starts of the box cells (the total is nrec),
    // (2) the records: record i lies in the last box cell whose start is <= i
    stage_box_cells<kBoxThreads>(B, b.cs, cfirst, cst, s_w);
    int top = 1;  // largest power of two below nbc
    while (2 * top < B.nbc) top *= 2;
    constexpr int G = 4;  // records per thread and step, their searches and loads interleaved
    for (int i0 = int(threadIdx.x); i0 < nrec; i0 += G * kBoxThreads) {
      int lo[G], ii[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        ii[u] = min(i0 + u * kBoxThreads, nrec - 1);
        lo[u] = 0;
      }
      for (int step = top; step > 0; step >>= 1)
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const int m = lo[u] + step;
          if (m < B.nbc && cst[m] <= ii[u]) lo[u] = m;
        }
      int64_t t[G];
      float4 p[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        t[u] = int64_t(cfirst[lo[u]]) + (ii[u] - cst[lo[u]]);
        p[u] = a.pc32[t[u]];
      }
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const int i = i0 + u * kBoxThreads;
        if (i >= nrec) break;
        int x, y, z;
        B.xyz(lo[u], x, y, z);
        const uint32_t packed = box_pack(x, y, z);
        const float4 rec = make_float4(p[u].x, p[u].y, p[u].z, __uint_as_float(packed));
        b.slot[r0 + i] = uint32_t(t[u]);
        b.pos[r0 + i] = rec;
        rpos[i] = rec;
      }
    }
  }
  // (3) the entry layout: the longest list of every wave of the order,
  // exclusive prefix (the host keeps nt within kBoxMaxWaves waves)
  const int nw = (nt + 63) / 64;
  for (int w = wv; w < nw; w += kBoxThreads / 64) {
    const int i = 64 * w + lane;
    const int m = wave_max(i < nt ? a.cnt[b.tgt[sb + i]] : 0);
    if (lane == 0) wpre[w] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < nw; ++w) {
      const int m = wpre[w];
      wpre[w] = run;
      run += m;
    }
  }
  __syncthreads();
  // (4) the entries of the staged bricks' targets: the neighbour's record
  // number of every kept pair (outside the table's zero interval at the
  // FP32C displacement, formed exactly as the list-order kernel forms it),
  // walking the stencil through the box records in LDS (list order,
  // box_walk); the targets of other bricks keep their entry offset only.
  const float hx = B.h[0], hy = B.h[1], hz = B.h[2];
  const StencilRow* rows = b.rows;
  int nrows = b.nrows;
  if constexpr (BATCH) {
    const SysGrid& S = a.sys[a.sys_slot[sb]];
    rows += S.row_off;
    nrows = S.nrows;
  }
  for (int i = int(threadIdx.x); i < nt; i += kBoxThreads) {
    const int64_t s = b.tgt[sb + i];
    const int64_t bb = b.eoff[brick] + 64 * int64_t(wpre[i >> 6]) + (i & 63);
    const float4 po = a.pc32[s];
    const uint32_t co = __float_as_uint(po.w);
    const int io = int(co & ((1u << kCellBitsX) - 1));
    const int jo = int((co >> kCellBitsX) & ((1u << kCellBitsY) - 1));
    const int ko = int(co >> (kCellBitsX + kCellBitsY));
    // staged bricks: the target's own record, packed above the entry offset
    const int bx = io - B.ox, by = jo - B.oy, bz = ko - B.oz;
    const uint32_t kb = 0x808080u - box_pack(bx, by, bz);
    int self = 0;
    if (nrec > 0) {
      const int qs = (bz * B.ey + by) * B.ex + bx;
      self = cst[qs] + int(s - cfirst[qs]);
    }
    const int cnt = a.cnt[s];
    if (nrec > 0 && b.fp64) {
      // FP64: every pair, at its position in the lists' slice layout
      int64_t o = a.slice_off[s >> 6] + (s & 63);
      if (cnt > 0)
        box_walk(B, rows, nrows, io, jo, ko, [&](int q0, int q1, uint32_t) {
          for (int r = cst[q0], re = cst[q1 + 1]; r < re; ++r) {
            if (r == self) continue;
            b.idx[o] = uint16_t(r);
            o += kSlice;
          }
        });
      b.base[sb + i] = bb;
      continue;
    }
    if (nrec > 0) {
      // only the kept pairs (not in the table's zero interval) are stored
      int64_t o = bb;
      int kept = 0;
      if (cnt > 0)
        box_walk(B, rows, nrows, io, jo, ko, [&](int q0, int q1, uint32_t) {
          for (int r = cst[q0], re = cst[q1 + 1]; r < re; ++r) {
            if (r == self) continue;
            const float4 pt = rpos[r];
            float di, dj, dk;
            box_diff(__float_as_uint(pt.w), kb, di, dj, dk);
            const float dx = __builtin_fmaf(di, hx, pt.x - po.x);
            const float dy = __builtin_fmaf(dj, hy, pt.y - po.y);
            const float dz = __builtin_fmaf(dk, hz, pt.z - po.z);
            float y;
            if (!fp32c_kept(a.ct, fp32c_interval(a.ct, fp32c_r2(dx, dy, dz), y))) continue;
            b.idx[o] = uint16_t(r);
            o += kSlice;
            ++kept;
          }
        });
      b.base[sb + i] = bb | (int64_t(self) << kSelfShift) | (int64_t(kept) << kKeptShift);
      continue;
    }
    b.base[sb + i] = bb;  // unstaged brick: the pair sum reads its list
  }
}

// pair sum of slot s through global gathers (FP32C, one operand; the
// coefficient table in LDS)
__device__ __forceinline__ void fp32c_pairs_global(const PairArgs& a, int64_t s, const float4* __restrict__ tab,
                                                   const float4* __restrict__ xg, double (&acc)[1][3]) {
  const int cnt = a.cnt[s];
  const int64_t base = a.slice_off[s >> 6] + (s & 63);
  const Grid& g = a.sys ? a.sys[a.sys_slot[s]].g : a.g;
  const float hx = float(g.h[0]), hy = float(g.h[1]), hz = float(g.h[2]);
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const float4 po = a.pc32[s];
  const uint32_t co = __float_as_uint(po.w);
  const int io = int(co & ((1u << kCellBitsX) - 1));
  const int jo = int((co >> kCellBitsX) & ((1u << kCellBitsY) - 1));
  const int ko = int(co >> (kCellBitsX + kCellBitsY));
  float ps[3] = {0.f, 0.f, 0.f};
  KahanF kh[3];
  int nk = 0;  // kept pairs so far
  for (int k = 0; k < cnt; ++k) {
    const uint32_t e = a.ell[base + int64_t(k) * kSlice];
    const uint32_t t = e & kSlotMask;
    const float4 pt = a.pc32[t];
    const float4 f = xg[t];
    const uint32_t ct = __float_as_uint(pt.w);
    const uint32_t code = e >> kSlotBits;
    const int di = int(ct & ((1u << kCellBitsX) - 1)) - io + (int(code & 1u) - int((code >> 1) & 1u)) * nx;
    const int dj = int((ct >> kCellBitsX) & ((1u << kCellBitsY) - 1)) - jo +
                   (int((code >> 2) & 1u) - int((code >> 3) & 1u)) * ny;
    const int dk = int(ct >> (kCellBitsX + kCellBitsY)) - ko + (int((code >> 4) & 1u) - int((code >> 5) & 1u)) * nz;
    const float dx = __builtin_fmaf(float(di), hx, pt.x - po.x);
    const float dy = __builtin_fmaf(float(dj), hy, pt.y - po.y);
    const float dz = __builtin_fmaf(float(dk), hz, pt.z - po.z);
    float y;
    const int it = fp32c_interval(a.ct, fp32c_r2(dx, dy, dz), y);
    if (!fp32c_kept(a.ct, it)) continue;  // the zero interval
    float A, Bx;
    fp32c_eval(tab, it, y, A, Bx);
    fp32c_acc(A, Bx, dx, dy, dz, f.x, f.y, f.z, ps[0], ps[1], ps[2]);
    if ((++nk & (kMergeEvery - 1)) == 0) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        kh[q].add(ps[q]);
        ps[q] = 0.f;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    kh[q].add(ps[q]);
    acc[0][q] = kh[q].value();
  }
}

template <int EPI, bool ONE>  // ONE: the table has one interval width
__global__ void __launch_bounds__(kBoxThreads) k_apply_box(PairArgs a, BoxView b, const float4* __restrict__ xg,
                                                          const double* __restrict__ xsoa, SlotVec v, ChebEpi epi,
                                                          const double* __restrict__ u_glob,
                                                          double* __restrict__ out) {
  extern __shared__ float4 lds[];
  if (epi.skip && *epi.skip) return;  // uniform over the grid
  const int64_t brick = block_index(a.xcd);
  const int64_t c0 = brick * kBrickCells;
  const int32_t sb = b.cs[c0], se = b.cs[c0 + kBrickCells];
  if (sb >= se) return;  // no targets (uniform over the block)
  const int64_t r0 = b.off[brick];
  const int nrec = int(b.off[brick + 1] - r0);
  const int nt = se - sb;
  float4* s_tab = lds;                         // coefficient table
  float4* rpos = lds + kTabStride * a.ct.n;    // [nrec] offset xyz, packed box cell
  float4* rval = rpos + nrec;         // [nrec] operand
  load_tab(s_tab, a);
  if (nrec == 0) {
    __syncthreads();
    for (int i = int(threadIdx.x); i < nt; i += kBoxThreads) {
      const int64_t s = b.tgt[sb + i];
      double acc[1][3];
      fp32c_pairs_global(a, s, s_tab, xg, acc);
      apply_epilogue<1, EPI>(a, s, acc, xsoa, v, epi, u_glob, out);
    }
    return;
  }
  {
    // four records in flight per thread, then the remainder one at a time
    int i = int(threadIdx.x);
    for (; i + 3 * kBoxThreads < nrec; i += 4 * kBoxThreads) {
      const uint32_t t0 = b.slot[r0 + i], t1 = b.slot[r0 + i + kBoxThreads];
      const uint32_t t2 = b.slot[r0 + i + 2 * kBoxThreads], t3 = b.slot[r0 + i + 3 * kBoxThreads];
      const float4 p0 = b.pos[r0 + i], p1 = b.pos[r0 + i + kBoxThreads];
      const float4 p2 = b.pos[r0 + i + 2 * kBoxThreads], p3 = b.pos[r0 + i + 3 * kBoxThreads];
      const float4 f0 = xg[t0], f1 = xg[t1], f2 = xg[t2], f3 = xg[t3];
      rpos[i] = p0;
      rpos[i + kBoxThreads] = p1;
      rpos[i + 2 * kBoxThreads] = p2;
      rpos[i + 3 * kBoxThreads] = p3;
      rval[i] = f0;
      rval[i + kBoxThreads] = f1;
      rval[i + 2 * kBoxThreads] = f2;
      rval[i + 3 * kBoxThreads] = f3;
    }
    for (; i < nrec; i += kBoxThreads) {
      rpos[i] = b.pos[r0 + i];
      rval[i] = xg[b.slot[r0 + i]];
    }
  }
  __syncthreads();
  BoxGeom B;
  box_init(B, a, b.reach, brick, sb);
  const float hx = B.h[0], hy = B.h[1], hz = B.h[2];
  constexpr int U = 4;  // entries per group, the next group's record numbers loaded one ahead
  for (int i = int(threadIdx.x); i < nt; i += kBoxThreads) {
    const int64_t s = b.tgt[sb + i];
    const int64_t bpk = b.base[sb + i];
    const int64_t base = bpk & kBaseMask;
    const float4 po = rpos[int((uint64_t(bpk) >> kSelfShift) & kField12)];  // the target's own record
    const uint32_t kb = 0x808080u - __float_as_uint(po.w);
    const int cnt = int((uint64_t(bpk) >> kKeptShift) & kField12);  // stored (kept) entries
    float ps[3] = {0.f, 0.f, 0.f};
    KahanF kh[3];
    // displacement to record rec and its interval in the table (every
    // stored entry is a kept pair)
    auto disp = [&](const float4& pt, float& dx, float& dy, float& dz, float& y) {
      float di, dj, dk;
      box_diff(__float_as_uint(pt.w), kb, di, dj, dk);
      dx = __builtin_fmaf(di, hx, pt.x - po.x);
      dy = __builtin_fmaf(dj, hy, pt.y - po.y);
      dz = __builtin_fmaf(dk, hz, pt.z - po.z);
      return ONE ? fp32c_interval1(a.ct, fp32c_r2(dx, dy, dz), y) : fp32c_interval(a.ct, fp32c_r2(dx, dy, dz), y);
    };
    auto merge = [&](int k) {
      if ((k & (kMergeEvery - 1)) == kMergeEvery - 1) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          kh[q].add(ps[q]);
          ps[q] = 0.f;
        }
      }
    };
    auto one = [&](uint32_t rec, int k) {
      const float4 pt = rpos[rec];
      const float4 f = rval[rec];
      float dx, dy, dz, y, A, Bx;
      const int it = disp(pt, dx, dy, dz, y);
      fp32c_eval(s_tab, it, y, A, Bx);
      fp32c_acc(A, Bx, dx, dy, dz, f.x, f.y, f.z, ps[0], ps[1], ps[2]);
      merge(k);
    };
    if (cnt > 0) {  // an empty list has no storage
      const uint16_t* __restrict__ ci = b.idx + base;
      const int last = cnt - 1;
      uint32_t en[U];  // record numbers of the next group
#pragma unroll
      for (int u = 0; u < U; ++u) en[u] = ci[int64_t(min(u, last)) * kSlice];
      int k = 0;
      for (; k + U <= cnt; k += U) {
        uint32_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          c[u] = en[u];
          en[u] = ci[int64_t(min(k + U + u, last)) * kSlice];
        }
        // a group: its records, then displacements and intervals, then the
        // tensors from the table and the sums in list order
        float4 pt[U], f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          pt[u] = rpos[c[u]];
          f[u] = rval[c[u]];
        }
        float dx[U], dy[U], dz[U], y[U];
        int it[U];
#pragma unroll
        for (int u = 0; u < U; ++u) it[u] = disp(pt[u], dx[u], dy[u], dz[u], y[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          float A, Bx;
          fp32c_eval(s_tab, it[u], y[u], A, Bx);
          fp32c_acc(A, Bx, dx[u], dy[u], dz[u], f[u].x, f[u].y, f[u].z, ps[0], ps[1], ps[2]);
          merge(k + u);
        }
      }
#pragma unroll
      for (int u = 0; u < U - 1; ++u)
        if (k + u < cnt) one(en[u], k + u);  // entries en[0 .. cnt - k)
    }
    double acc[1][3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      kh[q].add(ps[q]);
      acc[0][q] = kh[q].value();
    }
    apply_epilogue<1, EPI>(a, s, acc, xsoa, v, epi, u_glob, out);
  }
}


// k_apply_box for NR > 1 operands (the fixed-order step of several noise
// vectors): the box's positions and the NR operands' records staged in LDS
// ([NR][records] after the positions), each kept pair's table evaluated
// once for all NR; per operand the sums, their merge points and the
// epilogue are k_apply's for NR operands.  Every brick with targets is
// staged (the host launches it only then).
template <int EPI, bool ONE, int NR>
__global__ void __launch_bounds__(kBoxThreads) k_apply_box_nr(PairArgs a, BoxView b, const float4* __restrict__ xg,
                                                             const double* __restrict__ xsoa, SlotVec v, ChebEpi epi,
                                                             const double* __restrict__ u_glob,
                                                             double* __restrict__ out) {
  static_assert(NR > 1, "one operand: k_apply_box");
  extern __shared__ float4 lds[];
  if (epi.skip && *epi.skip) return;  // uniform over the grid
  const int64_t brick = block_index(a.xcd);
  const int64_t c0 = brick * kBrickCells;
  const int32_t sb = b.cs[c0], se = b.cs[c0 + kBrickCells];
  if (sb >= se) return;  // no targets (uniform over the block)
  const int64_t r0 = b.off[brick];
  const int nrec = int(b.off[brick + 1] - r0);
  const int nt = se - sb;
  if (nrec == 0) return;  // not staged: never launched so (uniform over the block)
  const int64_t n = a.n;
  float4* s_tab = lds;                       // coefficient table
  float4* rpos = lds + kTabStride * a.ct.n;  // [nrec] offset xyz, packed box cell
  float4* rval = rpos + nrec;                // [NR][nrec] operands
  load_tab(s_tab, a);
  for (int i = int(threadIdx.x); i < nrec; i += kBoxThreads) {
    const uint32_t t = b.slot[r0 + i];
    rpos[i] = b.pos[r0 + i];
#pragma unroll
    for (int r = 0; r < NR; ++r) rval[r * nrec + i] = xg[r * n + t];
  }
  __syncthreads();
  BoxGeom B;
  box_init(B, a, b.reach, brick, sb);
  const float hx = B.h[0], hy = B.h[1], hz = B.h[2];
  constexpr int U = 2;  // entries per group, the next group's record numbers loaded one ahead
  for (int i = int(threadIdx.x); i < nt; i += kBoxThreads) {
    const int64_t s = b.tgt[sb + i];
    const int64_t bpk = b.base[sb + i];
    const int64_t base = bpk & kBaseMask;
    const float4 po = rpos[int((uint64_t(bpk) >> kSelfShift) & kField12)];  // the target's own record
    const uint32_t kb = 0x808080u - __float_as_uint(po.w);
    const int cnt = int((uint64_t(bpk) >> kKeptShift) & kField12);  // stored (kept) entries
    float ps[NR][3];
    KahanF kh[NR][3];
#pragma unroll
    for (int r = 0; r < NR; ++r) ps[r][0] = ps[r][1] = ps[r][2] = 0.f;
    auto disp = [&](const float4& pt, float& dx, float& dy, float& dz, float& y) {
      float di, dj, dk;
      box_diff(__float_as_uint(pt.w), kb, di, dj, dk);
      dx = __builtin_fmaf(di, hx, pt.x - po.x);
      dy = __builtin_fmaf(dj, hy, pt.y - po.y);
      dz = __builtin_fmaf(dk, hz, pt.z - po.z);
      return ONE ? fp32c_interval1(a.ct, fp32c_r2(dx, dy, dz), y) : fp32c_interval(a.ct, fp32c_r2(dx, dy, dz), y);
    };
    auto merge = [&](int k) {
      if ((k & (kMergeEvery - 1)) == kMergeEvery - 1) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            kh[r][q].add(ps[r][q]);
            ps[r][q] = 0.f;
          }
      }
    };
    auto one = [&](uint32_t rec, int k) {
      const float4 pt = rpos[rec];
      float dx, dy, dz, y, A, Bx;
      const int it = disp(pt, dx, dy, dz, y);
      fp32c_eval(s_tab, it, y, A, Bx);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const float4 f = rval[r * nrec + rec];
        fp32c_acc(A, Bx, dx, dy, dz, f.x, f.y, f.z, ps[r][0], ps[r][1], ps[r][2]);
      }
      merge(k);
    };
    if (cnt > 0) {
      const uint16_t* __restrict__ ci = b.idx + base;
      const int last = cnt - 1;
      uint32_t en[U];
#pragma unroll
      for (int u = 0; u < U; ++u) en[u] = ci[int64_t(min(u, last)) * kSlice];
      int k = 0;
      for (; k + U <= cnt; k += U) {
        uint32_t c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          c[u] = en[u];
          en[u] = ci[int64_t(min(k + U + u, last)) * kSlice];
        }
        float4 pt[U], f[U][NR];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          pt[u] = rpos[c[u]];
#pragma unroll
          for (int r = 0; r < NR; ++r) f[u][r] = rval[r * nrec + c[u]];
        }
        float dx[U], dy[U], dz[U], y[U];
        int it[U];
#pragma unroll
        for (int u = 0; u < U; ++u) it[u] = disp(pt[u], dx[u], dy[u], dz[u], y[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
          float A, Bx;
          fp32c_eval(s_tab, it[u], y[u], A, Bx);
#pragma unroll
          for (int r = 0; r < NR; ++r)
            fp32c_acc(A, Bx, dx[u], dy[u], dz[u], f[u][r].x, f[u][r].y, f[u][r].z, ps[r][0], ps[r][1], ps[r][2]);
          merge(k + u);
        }
      }
#pragma unroll
      for (int u = 0; u < U - 1; ++u)
        if (k + u < cnt) one(en[u], k + u);
    }
    double acc[NR][3];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        kh[r][q].add(ps[r][q]);
        acc[r][q] = kh[r][q].value();
      }
    apply_epilogue<NR, EPI>(a, s, acc, xsoa, v, epi, u_glob, out);
  }
}

// FP64 pair sum with the operand records staged in LDS: one workgroup per
// brick, its targets in slot order; per entry the stored tensor (list
// layout) and the neighbour's record number in the brick's box (the same
// layout), the operand read from LDS.  List order and fp64_acc as in
// k_apply, so the sums are bit-identical.  A brick that is not staged
// gathers the operands from global memory.
template <int EPI>
__global__ void __launch_bounds__(kBoxThreads) k_apply_box64(PairArgs a, BoxView b, const double4* __restrict__ c64,
                                                            const double4* __restrict__ xg,
                                                            const double* __restrict__ xsoa, SlotVec v, ChebEpi epi,
                                                            const double* __restrict__ u_glob,
                                                            double* __restrict__ out) {
  extern __shared__ double4 lds64[];
  if (epi.skip && *epi.skip) return;  // uniform over the grid
  const int64_t brick = block_index(a.xcd);
  const int64_t c0 = brick * kBrickCells;
  const int32_t sb = b.cs[c0], se = b.cs[c0 + kBrickCells];
  if (sb >= se) return;  // no targets (uniform over the block)
  const int64_t r0 = b.off[brick];
  const int nrec = int(b.off[brick + 1] - r0);
  const int nt = se - sb;
  double4* rval = lds64;  // [nrec] operand
  if (nrec > 0) {
    // four records in flight per thread, then the remainder one at a time
    int i = int(threadIdx.x);
    for (; i + 3 * kBoxThreads < nrec; i += 4 * kBoxThreads) {
      const uint32_t t0 = b.slot[r0 + i], t1 = b.slot[r0 + i + kBoxThreads];
      const uint32_t t2 = b.slot[r0 + i + 2 * kBoxThreads], t3 = b.slot[r0 + i + 3 * kBoxThreads];
      const double4 f0 = xg[t0], f1 = xg[t1], f2 = xg[t2], f3 = xg[t3];
      rval[i] = f0;
      rval[i + kBoxThreads] = f1;
      rval[i + 2 * kBoxThreads] = f2;
      rval[i + 3 * kBoxThreads] = f3;
    }
    for (; i < nrec; i += kBoxThreads) rval[i] = xg[b.slot[r0 + i]];
    __syncthreads();
  }
  const double sg = a.bsign;
  constexpr int U = 2;  // entries per group, the next group's record numbers loaded one ahead (2 measured best of 2, 4, 8)
  for (int i = int(threadIdx.x); i < nt; i += kBoxThreads) {
    const int64_t s = sb + i;
    const int cnt = a.cnt[s];
    const int64_t base = a.slice_off[s >> 6] + (s & 63);
    const double4* __restrict__ cf = c64 + base;
    double ps[3] = {0.0, 0.0, 0.0};
    if (cnt > 0 && nrec > 0) {
      const uint16_t* __restrict__ ci = b.idx + base;
      const int last = cnt - 1;
      uint32_t en[U];
#pragma unroll
      for (int u = 0; u < U; ++u) en[u] = ci[int64_t(min(u, last)) * kSlice];
      int k = 0;
      for (; k + U <= cnt; k += U) {
        uint32_t e[U];
        double4 c[U], f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          e[u] = en[u];
          en[u] = ci[int64_t(min(k + U + u, last)) * kSlice];
          c[u] = cf[int64_t(k + u) * kSlice];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) f[u] = rval[e[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) fp64_acc(c[u], sg, f[u], ps[0], ps[1], ps[2]);
      }
#pragma unroll
      for (int u = 0; u < U - 1; ++u)
        if (k + u < cnt) fp64_acc(cf[int64_t(k + u) * kSlice], sg, rval[en[u]], ps[0], ps[1], ps[2]);
    } else if (cnt > 0) {
      const uint32_t* __restrict__ lst = a.ell + base;
      for (int k = 0; k < cnt; ++k)
        fp64_acc(cf[int64_t(k) * kSlice], sg, xg[lst[int64_t(k) * kSlice] & kSlotMask], ps[0], ps[1], ps[2]);
    }
    const double acc[1][3] = {{ps[0], ps[1], ps[2]}};
    apply_epilogue<1, EPI>(a, s, acc, xsoa, v, epi, u_glob, out);
  }
}

// Field points: sum over the element's particle list, own element
// included, no exclusion (ggem_stokes.C:396-420), full minimum image.
template <int P>
__global__ void __launch_bounds__(kBlock) k_pair_points(PairArgs a, const StencilRow* __restrict__ rows,
                                                       int nrows, const int32_t* __restrict__ cell_start,
                                                       const float4* __restrict__ fg32,
                                                       const double4* __restrict__ fg64, int64_t m,
                                                       const double* __restrict__ pts,
                                                       const int32_t* __restrict__ elem,
                                                       double* __restrict__ u) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m) return;
  int ci, cj, ck;
  elem_ijk(a.g, elem[q], ci, cj, ck);
  const double4 xq = make_double4(pts[3 * q], pts[3 * q + 1], pts[3 * q + 2], 0.0);
  double ux = 0, uy = 0, uz = 0;
  KahanF kx, ky, kz;
  for_each_segment(a.g, rows, nrows, cell_start, ci, cj, ck, [&](int32_t lo, int32_t hi, uint32_t) {
    for (int32_t t = lo; t < hi; ++t) {
      double dx, dy, dz;
      pair_disp<true>(a.g, xq, a.x4[t], 0u, dx, dy, dz);
      if (P == GGEM_FP64) {
        const double4 f = fg64[t];
        green_acc_f64(a.k, dx, dy, dz, f.x, f.y, f.z, ux, uy, uz);
      } else {
        const float4 f = fg32[t];
        const float fx = float(dx), fy = float(dy), fz = float(dz);
        float A, Bx;
        green_coef_f32(a.k32, fx * fx + fy * fy + fz * fz, A, Bx);
        const float B = Bx * (fx * f.x + fy * f.y + fz * f.z);
        kx.add(A * f.x + B * fx);
        ky.add(A * f.y + B * fy);
        kz.add(A * f.z + B * fz);
      }
    }
  });
  if (P != GGEM_FP64) {
    ux = kx.value();
    uy = ky.value();
    uz = kz.value();
  }
  u[3 * q] = ux;
  u[3 * q + 1] = uy;
  u[3 * q + 2] = uz;
}

// erfcx fit (see erfcx_f32): coefficients of Q in x = 2t/t_max - 1,
// t_max = 1/(1 + 0.45/2), least squares with relative weights on
// z in [0.45, 1e4].
constexpr float kErfcxTmax = 0.816326530612f;
constexpr float kErfcxC[kErfcxDeg + 1] = {
    4.52274475e-01f,  2.40276895e-01f, 8.51269019e-02f,  1.35228369e-02f,
    -3.15813629e-03f, -1.64298252e-03f, 2.09066292e-04f, 1.94463725e-04f,
    -2.89952557e-05f, -1.94542224e-05f, 3.23959705e-06f};

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
  a.k32.erfcx_scale = 2.0f / kErfcxTmax;
  for (int i = 0; i <= kErfcxDeg; ++i) a.k32.erfcx_c[i] = kErfcxC[i];
  a.x4 = ctx->x4;
  a.cnt = ctx->nbr_cnt;
  a.slice_off = ctx->slice_off;
  a.ell = ctx->ell;
  a.perm = ctx->perm;
  a.fs = ctx->fs;
  a.n = ctx->n;
  a.bsign = double(ctx->bsign);
  a.sys = ctx->sys_d;
  a.sys_slot = ctx->sys_slot;
  a.pc32 = ctx->pc32;
  a.xcd = 1;
  a.exact = explicit_minimg(ctx);
  a.tab = ctx->tab32;
  a.ct = ctx->tab;
  return a;
}

template <int P, int NR, int EPI, int BLK>
void launch_kb(ggem_ctx* ctx, const PairArgs& a, const double4* g64, const float4* g32, const double* soa,
               const SlotVec& v, const ChebEpi& epi, const double* u_glob, double* out) {
  if constexpr (P == GGEM_FP64)
    hipLaunchKernelGGL((k_apply<P, NR, EPI, BLK>), dim3(blocks_for(a.n, BLK)), dim3(BLK), 0, ctx->stream, a,
                       ctx->coef64, g64, g32, soa, v, epi, u_glob, out);
  else
    hipLaunchKernelGGL((k_apply<P, NR, EPI, BLK>), dim3(blocks_for(a.n, BLK)), dim3(BLK), 0, ctx->stream, a,
                       nullptr, g64, g32, soa, v, epi, u_glob, out);
}

BoxArgs box_args(const ggem_ctx* ctx) {
  BoxArgs b;
  b.cs = ctx->cell_start;
  for (int d = 0; d < 3; ++d) b.reach[d] = ctx->reach[d];
  b.rows = ctx->rows_d;
  b.nrows = ctx->n_rows;
  b.cells_cap = b.waves_cap = b.recs_cap = 0;
  b.cnt = ctx->box_cnt;
  b.ecnt = ctx->box_ecnt;
  b.maxrec = ctx->box_max_d;
  b.recs = ctx->brick_recs;
  b.off = ctx->box_off;
  b.eoff = ctx->box_eoff;
  b.tgt = ctx->box_tgt;
  b.base = ctx->box_base;
  b.slot = ctx->box_slot;
  b.pos = ctx->box_pos;
  b.idx = ctx->box_idx;
  b.fp64 = ctx->box_fp64 ? 1 : 0;
  return b;
}

BoxView box_view(const ggem_ctx* ctx) {
  BoxView b;
  b.cs = ctx->cell_start;
  for (int d = 0; d < 3; ++d) b.reach[d] = ctx->reach[d];
  b.off = ctx->box_off;
  b.tgt = ctx->box_tgt;
  b.base = ctx->box_base;
  b.slot = ctx->box_slot;
  b.pos = ctx->box_pos;
  b.idx = ctx->box_idx;
  return b;
}

unsigned n_bricks(const ggem_ctx* ctx) { return unsigned(ctx->grid.n_cells >> (3 * kBrickBits)); }

template <int EPI>
hipError_t launch_box(ggem_ctx* ctx, PairArgs a, const float4* xg, const double* soa, const SlotVec& v,
                      const ChebEpi& epi, const double* u_glob, double* out) {
  if (!ctx->box_valid || ctx->box_fp64 || ctx->box_max <= 0 || ctx->box_max > kBoxMaxRecs)
    return hipErrorInvalidConfiguration;
  a.xcd = 1;
  const size_t lds = sizeof(float4) * (kTabStride * size_t(ctx->tab.n) + size_t(ctx->box_max) * 2);
  if (ctx->tab.n2 == 0)
    hipLaunchKernelGGL((k_apply_box<EPI, true>), dim3(n_bricks(ctx)), dim3(kBoxThreads), lds, ctx->stream, a,
                       box_view(ctx), xg, soa, v, epi, u_glob, out);
  else
    hipLaunchKernelGGL((k_apply_box<EPI, false>), dim3(n_bricks(ctx)), dim3(kBoxThreads), lds, ctx->stream, a,
                       box_view(ctx), xg, soa, v, epi, u_glob, out);
  return hipGetLastError();
}

// staged FP32C pair sum of NR > 1 operands: every brick with targets staged
// and the box's positions and NR operand tables within the block's LDS
constexpr size_t kBoxNrLds = 81920;  // two blocks per CU
size_t box_nr_lds(const ggem_ctx* ctx, int nr) {
  return sizeof(float4) * (kTabStride * size_t(ctx->tab.n) + size_t(ctx->box_max) * size_t(1 + nr));
}
bool box_nr_ready(const ggem_ctx* ctx, int nr) {
  return nr > 1 && nr <= 4 && ctx->box_valid && !ctx->box_fp64 && !ctx->box_unstaged && ctx->box_max > 0 &&
         ctx->box_max <= kBoxMaxRecs && box_nr_lds(ctx, nr) <= kBoxNrLds;
}

template <int EPI, int NR>
hipError_t launch_box_nr(ggem_ctx* ctx, PairArgs a, const float4* xg, const double* soa, const SlotVec& v,
                         const ChebEpi& epi, const double* u_glob, double* out) {
  if (!box_nr_ready(ctx, NR)) return hipErrorInvalidConfiguration;
  a.xcd = 1;
  const size_t lds = box_nr_lds(ctx, NR);
  if (ctx->tab.n2 == 0)
    hipLaunchKernelGGL((k_apply_box_nr<EPI, true, NR>), dim3(n_bricks(ctx)), dim3(kBoxThreads), lds, ctx->stream, a,
                       box_view(ctx), xg, soa, v, epi, u_glob, out);
  else
    hipLaunchKernelGGL((k_apply_box_nr<EPI, false, NR>), dim3(n_bricks(ctx)), dim3(kBoxThreads), lds, ctx->stream,
                       a, box_view(ctx), xg, soa, v, epi, u_glob, out);
  return hipGetLastError();
}

template <int EPI>
hipError_t launch_box64(ggem_ctx* ctx, PairArgs a, const double4* xg, const double* soa, const SlotVec& v,
                        const ChebEpi& epi, const double* u_glob, double* out) {
  if (!ctx->box_valid || !ctx->box_fp64 || ctx->box_max <= 0 || ctx->box_max > kBoxMaxRecs || !ctx->coef64_valid)
    return hipErrorInvalidConfiguration;
  a.xcd = 1;
  const size_t lds = size_t(ctx->box_max) * sizeof(double4);
  hipLaunchKernelGGL((k_apply_box64<EPI>), dim3(n_bricks(ctx)), dim3(kBoxThreads), lds, ctx->stream, a,
                     box_view(ctx), ctx->coef64, xg, soa, v, epi, u_glob, out);
  return hipGetLastError();
}

// candidate index of the staged-record kernels (one operand, after the
// block shapes of k_apply); it is the default where some brick is staged
constexpr int kBoxShape = kNApply;
bool box_ready(const ggem_ctx* ctx, int precision, int nr) {
  if (nr > 1) return precision == GGEM_FP32C && box_nr_ready(ctx, nr);
  if (!ctx->box_valid || ctx->box_max <= 0) return false;
  return precision == GGEM_FP32C ? !ctx->box_fp64 : (ctx->box_fp64 && ctx->coef64_valid);
}
// the staged kernel by default for one operand; for two or four it is an
// autotuner candidate only (256 systems x 16 operands in chunks of 4:
// 3.53 ms per apply against 2.25 ms for the list-order kernel in chunks of 8)
int apply_default(const ggem_ctx* ctx, int precision, int nr) {
  return nr == 1 && box_ready(ctx, precision, nr) ? kBoxShape : 0;
}
int apply_candidates(const ggem_ctx* ctx, int precision, int nr) {
  return box_ready(ctx, precision, nr) ? kNApply + 1 : kNApply;
}

// one apply launch in candidate shape c
template <int P, int NR, int EPI>
hipError_t launch_k(ggem_ctx* ctx, PairArgs a, int c, const double4* g64, const float4* g32, const double* soa,
                    const SlotVec& v, const ChebEpi& epi, const double* u_glob, double* out) {
  if (c == kBoxShape) {
    if constexpr (P == GGEM_FP32C && NR == 1) return launch_box<EPI>(ctx, a, g32, soa, v, epi, u_glob, out);
    if constexpr (P == GGEM_FP32C && (NR == 2 || NR == 4))
      return launch_box_nr<EPI, NR>(ctx, a, g32, soa, v, epi, u_glob, out);
    if constexpr (P == GGEM_FP64 && NR == 1) return launch_box64<EPI>(ctx, a, g64, soa, v, epi, u_glob, out);
    return hipErrorInvalidConfiguration;
  }
  if (c < 0 || c >= kNApply) return hipErrorInvalidValue;
  const Shape sh = kApplyShapes[c];
  a.xcd = sh.xcd;
  switch (sh.block) {
    case 64: launch_kb<P, NR, EPI, 64>(ctx, a, g64, g32, soa, v, epi, u_glob, out); break;
    case 128: launch_kb<P, NR, EPI, 128>(ctx, a, g64, g32, soa, v, epi, u_glob, out); break;
    case 512: launch_kb<P, NR, EPI, 512>(ctx, a, g64, g32, soa, v, epi, u_glob, out); break;
    default: launch_kb<P, NR, EPI, 256>(ctx, a, g64, g32, soa, v, epi, u_glob, out); break;
  }
  return hipGetLastError();
}

template <int P>
hipError_t launch_cheb_nr(ggem_ctx* ctx, const PairArgs& a, int c, int nr, const double4* g64, const float4* g32,
                          const double* soa, const SlotVec& v, const ChebEpi& epi) {
  switch (nr) {
    case 1: return launch_k<P, 1, EPI_CHEB>(ctx, a, c, g64, g32, soa, v, epi, nullptr, nullptr);
    case 2: return launch_k<P, 2, EPI_CHEB>(ctx, a, c, g64, g32, soa, v, epi, nullptr, nullptr);
    case 4: return launch_k<P, 4, EPI_CHEB>(ctx, a, c, g64, g32, soa, v, epi, nullptr, nullptr);
    default: return launch_k<P, 8, EPI_CHEB>(ctx, a, c, g64, g32, soa, v, epi, nullptr, nullptr);
  }
}

template <bool E, bool B, int BLK>
void launch_tb(ggem_ctx* ctx, const PairArgs& a) {
  hipLaunchKernelGGL((k_tensors<E, B, BLK>), dim3(blocks_for(a.n, BLK)), dim3(BLK), 0, ctx->stream, a,
                     ctx->coef64);
}

template <bool E, bool B>
hipError_t launch_t(ggem_ctx* ctx, PairArgs a, int c) {
  if (c < 0 || c >= kNTensor) return hipErrorInvalidValue;
  const Shape sh = kTensorShapes[c];
  a.xcd = sh.xcd;
  switch (sh.block) {
    case 64: launch_tb<E, B, 64>(ctx, a); break;
    case 128: launch_tb<E, B, 128>(ctx, a); break;
    case 512: launch_tb<E, B, 512>(ctx, a); break;
    default: launch_tb<E, B, 256>(ctx, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_t_any(ggem_ctx* ctx, const PairArgs& a, int c) {
  const bool ex = explicit_minimg(ctx), bt = ctx->nsys > 1;
  if (bt) return ex ? launch_t<true, true>(ctx, a, c) : launch_t<false, true>(ctx, a, c);
  return ex ? launch_t<true, false>(ctx, a, c) : launch_t<false, false>(ctx, a, c);
}

}  // namespace

hipError_t launch_pair_tensors(ggem_ctx* ctx, int precision) {
  if (precision != GGEM_FP64) return hipErrorInvalidValue;  // FP32C evaluates its table in the pair sums
  if (ctx->n == 0) return hipSuccess;
  const PairArgs a = make_args(ctx);
  auto run = [&](int c) { return launch_t_any(ctx, a, c); };
  const int c = tune_pick(ctx, tune_key("tensors", precision, ctx->nsys > 1, ctx->n), kNTensor, run);
  ProfScope ps(ctx, "tensors_fp64");
  return run(c);
}

bool box_supported(const ggem_ctx* ctx) { return ctx->n > 0 && !explicit_minimg(ctx); }

hipError_t launch_box_count(ggem_ctx* ctx) {
  ProfScope ps(ctx, "box_count");
  const BoxArgs b = box_args(ctx);
  hipLaunchKernelGGL(k_box_count, dim3(n_bricks(ctx)), dim3(kBoxThreads), 0, ctx->stream, make_args(ctx), b);
  hipLaunchKernelGGL(k_box_max, dim3(1), dim3(kMaxThreads), 0, ctx->stream, b, int64_t(n_bricks(ctx)));
  return hipGetLastError();
}

hipError_t launch_box_fill(ggem_ctx* ctx, int32_t max_records, int32_t max_targets) {
  ProfScope ps(ctx, "box_fill");
  BoxArgs b = box_args(ctx);
  // LDS tables sized for the largest box and the most crowded brick
  auto cells = [](const int32_t* r) {
    return ((1 << kBrickBits) + 2 * r[0]) * ((1 << kBrickBits) + 2 * r[1]) * ((1 << kBrickBits) + 2 * r[2]);
  };
  int cap = cells(ctx->reach);
  for (int k = 0; k < ctx->nsys && ctx->nsys > 1; ++k) cap = std::max(cap, cells(ctx->sys_h[k].reach));
  b.cells_cap = std::min(cap, kBoxMaxCells);
  b.waves_cap = (max_targets + 63) / 64;
  b.recs_cap = max_records;
  if (b.waves_cap > kBoxMaxWaves || max_records > kBoxMaxRecs) return hipErrorInvalidValue;
  const size_t lds = sizeof(float4) * size_t(b.recs_cap) + sizeof(int32_t) * (2 * size_t(b.cells_cap) + b.waves_cap + 2);
  if (ctx->nsys > 1)
    hipLaunchKernelGGL(k_box_fill<true>, dim3(n_bricks(ctx)), dim3(kBoxThreads), lds, ctx->stream, make_args(ctx),
                       b);
  else
    hipLaunchKernelGGL(k_box_fill<false>, dim3(n_bricks(ctx)), dim3(kBoxThreads), lds, ctx->stream,
                       make_args(ctx), b);
  return hipGetLastError();
}

hipError_t launch_pair_beads(ggem_ctx* ctx, int precision, int epi, const double* u_glob, double* out) {
  if (ctx->n == 0) return hipSuccess;
  const PairArgs a = make_args(ctx);
  const SlotVec none{nullptr, nullptr, nullptr};
  const ChebEpi no_epi{0, 0.0, 0.0, 0.0, none, nullptr};
  auto run = [&](int c) {
    if (precision == GGEM_FP64)
      return epi == EPI_LOCAL
                 ? launch_k<GGEM_FP64, 1, EPI_LOCAL>(ctx, a, c, ctx->fg64, ctx->fg32, nullptr, none, no_epi, u_glob, out)
                 : launch_k<GGEM_FP64, 1, EPI_MOBILITY>(ctx, a, c, ctx->fg64, ctx->fg32, nullptr, none, no_epi, u_glob,
                                                        out);
    return epi == EPI_LOCAL
               ? launch_k<GGEM_FP32C, 1, EPI_LOCAL>(ctx, a, c, ctx->fg64, ctx->fg32, nullptr, none, no_epi, u_glob, out)
               : launch_k<GGEM_FP32C, 1, EPI_MOBILITY>(ctx, a, c, ctx->fg64, ctx->fg32, nullptr, none, no_epi, u_glob,
                                                       out);
  };
  // both epilogues write out[] only: idempotent, so the shape can be tuned here
  const int c = tune_pick(ctx, tune_key("pair_beads", precision, epi, ctx->n), apply_candidates(ctx, precision, 1),
                          run, apply_default(ctx, precision, 1));
  ProfScope ps(ctx, precision == GGEM_FP64 ? "pair_beads_fp64" : "pair_beads_fp32c");
  return run(c);
}

hipError_t launch_pair_points(ggem_ctx* ctx, int precision, int64_t m, const double* pts,
                              const int32_t* elem, double* u) {
  if (m == 0) return hipSuccess;
  const PairArgs a = make_args(ctx);
  ProfScope ps(ctx, precision == GGEM_FP64 ? "pair_points_fp64" : "pair_points_fp32c");
  if (precision == GGEM_FP64)
    hipLaunchKernelGGL(k_pair_points<GGEM_FP64>, dim3(blocks_for(m)), dim3(kBlock), 0, ctx->stream, a,
                       ctx->rows_d, ctx->n_rows, ctx->cell_start, ctx->fg32, ctx->fg64, m, pts, elem, u);
  else
    hipLaunchKernelGGL(k_pair_points<GGEM_FP32C>, dim3(blocks_for(m)), dim3(kBlock), 0, ctx->stream, a,
                       ctx->rows_d, ctx->n_rows, ctx->cell_start, ctx->fg32, ctx->fg64, m, pts, elem, u);
  return hipGetLastError();
}

hipError_t launch_apply(ggem_ctx* ctx, int precision, int nrhs, const SlotVec& x, const SlotVec& v,
                        const ChebEpi& epi) {
  if (ctx->n == 0) return hipSuccess;
  const PairArgs a = make_args(ctx);
  const int64_t n = ctx->n, stride = 3 * n;
  struct Chunk {
    int nr;
    ChebEpi e;
    SlotVec vv;
    const double4* g64;
    const float4* g32;
    const double* soa;
  };
  std::vector<Chunk> chunks;
  for (int done = 0; done < nrhs;) {
    const int left = nrhs - done;
    const int nr = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
    ChebEpi e = epi;
    if (e.xm1.soa) e.xm1.soa += done * stride;
    if (e.y) e.y += done * stride;
    SlotVec vv{v.soa + done * stride, v.g32 ? v.g32 + done * n : nullptr, v.g64 ? v.g64 + done * n : nullptr};
    chunks.push_back({nr, e, vv, x.g64 ? x.g64 + done * n : nullptr, x.g32 ? x.g32 + done * n : nullptr,
                      x.soa + done * stride});
    done += nr;
  }
  auto run = [&](const Chunk& k, int c) {
    return precision == GGEM_FP64 ? launch_cheb_nr<GGEM_FP64>(ctx, a, c, k.nr, k.g64, k.g32, k.soa, k.vv, k.e)
                                  : launch_cheb_nr<GGEM_FP32C>(ctx, a, c, k.nr, k.g64, k.g32, k.soa, k.vv, k.e);
  };
  // shapes: tuned on launches that only write v (kinds 0 and 1 with v not
  // aliasing x); the recurrence step (kind 2 updates y and overwrites
  // x_{k-1}) reuses the pick of its operand count
  std::vector<int> pick(chunks.size());
  for (size_t i = 0; i < chunks.size(); ++i) {
    const std::string key = tune_key("apply", precision, chunks[i].nr, n);
    const int dflt = apply_default(ctx, precision, chunks[i].nr);
    if (epi.kind != 2 && v.soa != x.soa)
      pick[i] = tune_pick(ctx, key, apply_candidates(ctx, precision, chunks[i].nr),
                          [&](int c) { return run(chunks[i], c); }, dflt);
    else
      pick[i] = tune_lookup(ctx, key, dflt);
  }
  ProfScope ps(ctx, precision == GGEM_FP64 ? "apply_fp64" : "apply_fp32c");
  for (size_t i = 0; i < chunks.size(); ++i) {
    const hipError_t e = run(chunks[i], pick[i]);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace ggem
