// kernels_vec.hip -- vector algebra of the Chebyshev recurrence on
// slot-ordered SoA vectors: layout permutes, deterministic dot products
// (fixed two-stage reduction, no atomics), axpby and fill.
//
// Reference: the PETSc Vec calls of BrownianSystem::chebyshev_polynomial_approximation
// (brownian_system.C:1030-1131): VecAXPBY(y,a,b,x) = a x + b y, VecAXPY,
// VecDot, VecNorm.
#include "ggem_device.h"

namespace ggem {

namespace {

constexpr int kBlock = 256;
constexpr int kDotBlocks = 512;

inline unsigned blocks_for(int64_t n, int b = kBlock) { return unsigned((n + b - 1) / b); }

// interleaved particle order [r][3p + c]  ->  slot SoA [r][c][s]
__global__ void k_permute_in(int64_t n, int nrhs, const int32_t* __restrict__ perm,
                             const double* __restrict__ in, double* __restrict__ out) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t p = perm[s];
  for (int r = 0; r < nrhs; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) out[r * 3 * n + c * n + s] = in[r * 3 * n + 3 * p + c];
}

__global__ void k_permute_out(int64_t n, int nrhs, const int32_t* __restrict__ perm,
                              const double* __restrict__ in, double* __restrict__ out) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t p = perm[s];
  for (int r = 0; r < nrhs; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) out[r * 3 * n + 3 * p + c] = in[r * 3 * n + c * n + s];
}

// per-slot gather records of slot SoA vectors
__global__ void k_make_gather(int64_t n, int nrhs, const double* __restrict__ soa, float4* __restrict__ g32,
                              double4* __restrict__ g64) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  for (int r = 0; r < nrhs; ++r) {
    const double x = soa[r * 3 * n + s], y = soa[r * 3 * n + n + s], z = soa[r * 3 * n + 2 * n + s];
    if (g32) g32[r * n + s] = make_float4(float(x), float(y), float(z), 0.f);
    if (g64) g64[r * n + s] = make_double4(x, y, z, 0.0);
  }
}

// y = c_0 x_0 + c_1 x_1 (brownian_system.C:1055-1058), coefficients from
// the host or per system from device parameters [da, db, c_0 ..]
__global__ void k_cheb_init_y(int64_t n, int nrhs, const double* __restrict__ x0,
                              const double* __restrict__ x1, double* __restrict__ y, double c0,
                              double c1, const double* __restrict__ dpar, int pstride,
                              const int32_t* __restrict__ sys) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (dpar) {
    const double* pp = dpar + int64_t(sys ? sys[s] : 0) * pstride;
    c0 = pp[2];
    c1 = pp[3];
  }
  for (int r = 0; r < nrhs; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t i = int64_t(r) * 3 * n + int64_t(c) * n + s;
      y[i] = c0 * x0[i] + c1 * x1[i];
    }
}

struct ParamChunk {
  double v[64];
};

__global__ void k_set_params(double* __restrict__ dst, ParamChunk p, int cnt) {
  if (int(threadIdx.x) < cnt) dst[threadIdx.x] = p.v[threadIdx.x];
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < int(blockDim.x >> 6); ++w) t += sh[w];
  __syncthreads();
  return t;
}

// stage 1: partials[r][b] over a fixed grid-stride partition
__global__ void k_dot_partial(int64_t len, int nrhs, const double* __restrict__ a,
                              const double* __restrict__ b, double* __restrict__ partials) {
  __shared__ double sh[kBlock / 64];
  for (int r = 0; r < nrhs; ++r) {
    double s = 0.0;
    const double* ar = a + r * len;
    const double* br = b + r * len;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len;
         i += int64_t(gridDim.x) * blockDim.x)
      s += ar[i] * br[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partials[r * gridDim.x + blockIdx.x] = t;
  }
}

// stage 2: one block per right-hand side
__global__ void k_dot_final(int nparts, const double* __restrict__ partials, double* __restrict__ out) {
  __shared__ double sh[kBlock / 64];
  const int r = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += partials[r * nparts + i];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) out[r] = t;
}

// Per-system dot products of slot SoA vectors ([3][n]): block k sums the
// slot range of system k (its cells' slots are contiguous).
__global__ void k_seg_dot(const SysGrid* __restrict__ sys, const int32_t* __restrict__ cell_start, int64_t n,
                          const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
  __shared__ double sh[kBlock / 64];
  const SysGrid& S = sys[blockIdx.x];
  const int64_t s0 = cell_start[S.cell_off], s1 = cell_start[S.cell_off + S.g.n_cells];
  double acc = 0.0;
  for (int c = 0; c < 3; ++c)
    for (int64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) acc += a[c * n + s] * b[c * n + s];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// w = w - ca[k] x - cb[k] y per slot of system k
__global__ void k_seg_axpy2(int64_t n, const int32_t* __restrict__ sys_slot, const double* __restrict__ ca,
                            const double* __restrict__ x, const double* __restrict__ cb,
                            const double* __restrict__ y, double* __restrict__ w) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int k = sys_slot[s];
  const double a = ca[k], b = cb[k];
#pragma unroll
  for (int c = 0; c < 3; ++c) w[c * n + s] = w[c * n + s] - a * x[c * n + s] - b * y[c * n + s];
}

// out = sc[k] x per slot of system k
__global__ void k_seg_scale(int64_t n, const int32_t* __restrict__ sys_slot, const double* __restrict__ sc,
                            const double* __restrict__ x, double* __restrict__ out) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double a = sc[sys_slot[s]];
#pragma unroll
  for (int c = 0; c < 3; ++c) out[c * n + s] = a * x[c * n + s];
}

__global__ void k_axpby(int64_t len, double a, const double* __restrict__ x, double b,
                        double* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len) y[i] = a * x[i] + b * y[i];
}

__global__ void k_fill(int64_t len, double v, double* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len) y[i] = v;
}

}  // namespace

hipError_t launch_permute_in(ggem_ctx* ctx, int nrhs, const double* in, double* out) {
  if (ctx->n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_permute_in, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, nrhs,
                     ctx->perm, in, out);
  return hipGetLastError();
}

hipError_t launch_permute_out(ggem_ctx* ctx, int nrhs, const double* in, double* out) {
  if (ctx->n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_permute_out, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n,
                     nrhs, ctx->perm, in, out);
  return hipGetLastError();
}

hipError_t launch_set_params(ggem_ctx* ctx, double* dpar, int64_t off, const double* vals, int cnt) {
  for (int base = 0; base < cnt; base += 64) {
    ParamChunk p;
    const int m = cnt - base < 64 ? cnt - base : 64;
    for (int i = 0; i < 64; ++i) p.v[i] = i < m ? vals[base + i] : 0.0;
    hipLaunchKernelGGL(k_set_params, dim3(1), dim3(64), 0, ctx->stream, dpar + off + base, p, m);
  }
  return hipGetLastError();
}

hipError_t launch_cheb_init_y(ggem_ctx* ctx, int nrhs, const double* x0, const double* x1, double* y,
                              double c0, double c1, const double* dpar, int pstride, const int32_t* sys) {
  if (ctx->n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cheb_init_y, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, nrhs, x0,
                     x1, y, c0, c1, dpar, pstride, sys);
  return hipGetLastError();
}

hipError_t launch_make_gather(ggem_ctx* ctx, int nrhs, const double* soa, float4* g32, double4* g64) {
  if (ctx->n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_make_gather, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, nrhs,
                     soa, g32, g64);
  return hipGetLastError();
}

namespace {

// Two dot products a1 . b1 and a2 . b2 in one pass: each one's partials are
// k_dot_partial's (same partition, same order within a thread).
__global__ void k_dot2_partial(int64_t len, const double* __restrict__ a1, const double* __restrict__ b1,
                               const double* __restrict__ a2, const double* __restrict__ b2,
                               double* __restrict__ partials) {
  __shared__ double sh[kBlock / 64];
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += int64_t(gridDim.x) * blockDim.x) {
    s1 += a1[i] * b1[i];
    s2 += a2[i] * b2[i];
  }
  const double t1 = block_sum(s1, sh);
  const double t2 = block_sum(s2, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = t1;
    partials[kDotBlocks + blockIdx.x] = t2;
  }
}

// k_dot_final's sum of kDotBlocks partials (one block of kBlock threads;
// the total on thread 0)
__device__ __forceinline__ double final_sum(const double* __restrict__ partials, double* sh) {
  double s = 0.0;
  for (int i = threadIdx.x; i < kDotBlocks; i += blockDim.x) s += partials[i];
  return block_sum(s, sh);
}

// One term of the reference mode's estimators and stopping rule
// (brownian_system.C:1030-1155) on the device, with the host's formulas
// and FP64 order.  st = [aux2 = (M x0).x0, y.y before the term, -, -,
// err1, err2, terms, converged]; px: partials of x_{k+1} . x_{k+1}, pmy:
// of (M y).(M y) and of y.y after the term (kept in st[1] for the next).
// Once the rule fires *stop is set; the later terms change nothing here.
__global__ void __launch_bounds__(kBlock) k_cheb_ref_term(double* st, int32_t* stop, const double* __restrict__ px,
                                                           const double* __restrict__ pmy, double xabs, double tol,
                                                           int term, int check) {
  __shared__ double sh[kBlock / 64];
  const double xx = final_sum(px, sh);
  const double mm = final_sum(pmy, sh);
  const double yy = final_sum(pmy + kDotBlocks, sh);
  if (threadIdx.x != 0) return;
  if (!*stop) {
    const double aux2 = st[0], aux4 = sqrt(st[1]), aux3 = sqrt(xx) * xabs, aux1 = mm;
    double err1 = st[4], err2 = st[5];
    if (aux2 != 0.0) err1 = sqrt(fabs(aux1 - aux2) / aux2);
    if (aux4 != 0.0) err2 = aux3 / aux4;
    st[4] = err1;
    st[5] = err2;
    st[6] = double(term);
    if (check) {
      if (err1 <= tol) {
        st[7] = 1.0;
        *stop = 1;
      } else if (err2 <= 1E-9) {
        *stop = 1;
      }
    }
  }
  st[1] = yy;
}

__global__ void k_cheb_ref_init(double* st, int32_t* stop) {
  if (threadIdx.x < 8) st[threadIdx.x] = 0.0;
  if (threadIdx.x == 0) *stop = 0;
}

}  // namespace

hipError_t launch_cheb_ref_init(ggem_ctx* ctx, double* st, int32_t* stop) {
  hipLaunchKernelGGL(k_cheb_ref_init, dim3(1), dim3(64), 0, ctx->stream, st, stop);
  return hipGetLastError();
}

hipError_t launch_cheb_ref_term(ggem_ctx* ctx, double* st, int32_t* stop, const double* px, const double* pmy,
                                double xabs, double tol, int term, int check) {
  hipLaunchKernelGGL(k_cheb_ref_term, dim3(1), dim3(kBlock), 0, ctx->stream, st, stop, px, pmy, xabs, tol, term,
                     check);
  return hipGetLastError();
}

// partials of one dot product (kDotBlocks doubles) and of two (2 kDotBlocks)
hipError_t launch_dot_partial(ggem_ctx* ctx, const double* a, const double* b, double* partials) {
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(kBlock), 0, ctx->stream, int64_t(3) * ctx->n, 1, a, b,
                     partials);
  return hipGetLastError();
}

hipError_t launch_dot2_partial(ggem_ctx* ctx, const double* a1, const double* b1, const double* a2, const double* b2,
                               double* partials) {
  hipLaunchKernelGGL(k_dot2_partial, dim3(kDotBlocks), dim3(kBlock), 0, ctx->stream, int64_t(3) * ctx->n, a1, b1,
                     a2, b2, partials);
  return hipGetLastError();
}

// partials must hold nrhs * kDotBlocks doubles
hipError_t launch_dot(ggem_ctx* ctx, int nrhs, const double* a, const double* b, double* out_d,
                      double* partials) {
  const int64_t len = 3 * ctx->n;
  ProfScope ps(ctx, "dot");
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(kBlock), 0, ctx->stream, len, nrhs, a, b,
                     partials);
  hipLaunchKernelGGL(k_dot_final, dim3(nrhs), dim3(kBlock), 0, ctx->stream, kDotBlocks, partials, out_d);
  return hipGetLastError();
}

hipError_t launch_seg_dot(ggem_ctx* ctx, const double* a, const double* b, double* out_d) {
  ProfScope ps(ctx, "seg_dot");
  hipLaunchKernelGGL(k_seg_dot, dim3(unsigned(ctx->nsys)), dim3(kBlock), 0, ctx->stream, ctx->sys_d,
                     ctx->cell_start, ctx->n, a, b, out_d);
  return hipGetLastError();
}

hipError_t launch_seg_axpy2(ggem_ctx* ctx, const double* ca, const double* x, const double* cb, const double* y,
                            double* w) {
  if (ctx->n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_seg_axpy2, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->sys_slot,
                     ca, x, cb, y, w);
  return hipGetLastError();
}

hipError_t launch_seg_scale(ggem_ctx* ctx, const double* sc, const double* x, double* out) {
  if (ctx->n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_seg_scale, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->sys_slot,
                     sc, x, out);
  return hipGetLastError();
}

hipError_t launch_axpby(ggem_ctx* ctx, int64_t len, double a, const double* x, double b, double* y) {
  if (len == 0) return hipSuccess;
  ProfScope ps(ctx, "axpby");
  hipLaunchKernelGGL(k_axpby, dim3(blocks_for(len)), dim3(kBlock), 0, ctx->stream, len, a, x, b, y);
  return hipGetLastError();
}

hipError_t launch_fill(ggem_ctx* ctx, int64_t len, double v, double* y) {
  if (len == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(len)), dim3(kBlock), 0, ctx->stream, len, v, y);
  return hipGetLastError();
}

namespace {
__global__ void k_peek(PeekArgs p) {
  const int i = int(threadIdx.x);
  if (i >= p.n) return;
  p.dst[i] = p.bytes[i] == 8 ? *static_cast<const int64_t*>(p.src[i]) : int64_t(*static_cast<const int32_t*>(p.src[i]));
}
}  // namespace

hipError_t launch_peek(ggem_ctx* ctx, const PeekArgs& p) {
  hipLaunchKernelGGL(k_peek, dim3(1), dim3(64), 0, ctx->stream, p);
  return hipGetLastError();
}

}  // namespace ggem
