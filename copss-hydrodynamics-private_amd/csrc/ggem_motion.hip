// ggem_motion.hip -- the Brownian integrator's state on the device: the
// noise (the reference's host engine kept in the context, or a
// counter-based device generator), the saved positions R0, the position
// updates of the Fixman mid-point and free-draining steps with the wall
// and periodic corrections, and the unwrapped positions.
//
// Reference: Copss::fixman_integrate (copss.C:1565-1745),
// Copss::langevin_integrate (copss.C:1784-1925), FixPoint::check_walls
// (fix/fix_point.C:27-90), RandomGenerator (random_generator.C:43-99).
#include <cmath>
#include <vector>

#include "ggem_host.h"

namespace ggem {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(int64_t n, int b = kBlock) { return unsigned((n + b - 1) / b); }

// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC 2011): ten rounds of two 32x32 -> 64-bit
// multiplies with the Weyl-sequence key schedule.
struct U4 {
  uint32_t v[4];
};
__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(M0) * c.v[0], p1 = uint64_t(M1) * c.v[2];
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    c = U4{{hi1 ^ c.v[1] ^ k0, lo1, hi0 ^ c.v[3] ^ k1, lo0}};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Values 2j and 2j + 1 of vector r: counter (lo, hi of `counter`, r, j),
// key = seed; two 53-bit uniforms u1 in (0, 1], u2 in [0, 1) from the four
// words, then Box-Muller, scaled by sqrt(2) sqrt(dt).
__global__ void k_philox_noise(int64_t len, int nrhs, uint64_t seed, uint64_t counter, double scale,
                               double* __restrict__ out) {
  const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t npair = (len + 1) / 2;
  if (j >= npair) return;
  for (int r = 0; r < nrhs; ++r) {
    const U4 w = philox4x32_10(U4{{uint32_t(counter), uint32_t(counter >> 32), uint32_t(r), uint32_t(j)}},
                               uint32_t(seed), uint32_t(seed >> 32));
    const uint64_t a = (uint64_t(w.v[0]) << 21) ^ (uint64_t(w.v[1]) >> 11);  // 53 bits
    const uint64_t b = (uint64_t(w.v[2]) << 21) ^ (uint64_t(w.v[3]) >> 11);
    const double u1 = (double(a) + 1.0) * 0x1p-53;  // (0, 1]
    const double u2 = double(b) * 0x1p-53;          // [0, 1)
    const double rad = sqrt(-2.0 * log(u1)) * scale;
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    double* o = out + int64_t(r) * len;
    o[2 * j] = rad * cs;
    if (2 * j + 1 < len) o[2 * j + 1] = rad * sn;
  }
}

struct MoveArgs {
  double cv, cd, cf, cn;
  double bmin[3], bmax[3], L[3];
  int32_t periodic[3];
  int32_t wall_type;
  double radius;
  int32_t track;
};

// x = base + cv U + cd D + cf f + cn dw, then FixPoint::check_walls.
__global__ void k_move(int64_t n, MoveArgs m, const double* __restrict__ base, const double* __restrict__ U,
                       const double* __restrict__ D, const double* __restrict__ F, const double* __restrict__ W,
                       double* __restrict__ x, double* __restrict__ rout, int32_t* __restrict__ images) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double v[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int64_t i = 3 * p + d;
    double inc = 0.0;
    if (U) inc += m.cv * U[i];
    if (D) inc += m.cd * D[i];
    if (F) inc += m.cf * F[i];
    if (W) inc += m.cn * W[i];
    v[d] = base[i] + inc;
    if (m.track) rout[i] += inc;
  }
  if (m.wall_type == 0) {
    // slit: a periodic axis wraps once each way (fix_point.C:41-45), a wall
    // axis reflects (:46-61)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (m.periodic[d]) {
        if (v[d] < m.bmin[d]) {
          v[d] += m.L[d];
          images[3 * p + d] -= 1;
        }
        if (v[d] >= m.bmax[d]) {
          v[d] -= m.L[d];
          images[3 * p + d] += 1;
        }
      } else {
        if (v[d] < m.bmin[d]) v[d] = 2.0 * m.bmin[d] - v[d];
        if (v[d] >= m.bmax[d]) v[d] = 2.0 * m.bmax[d] - v[d];
      }
    }
  } else {
    // spherical cavity about the origin (fix_point.C:67-86)
    const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (r > m.radius) {
      const double s = 2.0 * m.radius / r - 1.0;  // (2R - |x|) / |x|
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] *= s;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) x[3 * p + d] = v[d];
}

// Velocity extrema, stage 1: per block the largest and smallest |U_i|^2
// (PointMesh::set_bead_velocity keeps norm_sq and takes sqrt of the extremes).
constexpr int kExtBlocks = 240;
__global__ void k_vext_partial(int64_t n, const double* __restrict__ u, double* __restrict__ part) {
  __shared__ double smax[kBlock], smin[kBlock];
  double mx = -1.0, mn = INFINITY;
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += int64_t(gridDim.x) * blockDim.x) {
    const double a = u[3 * p], b = u[3 * p + 1], c = u[3 * p + 2];
    const double q = a * a + b * b + c * c;
    mx = fmax(mx, q);
    mn = fmin(mn, q);
  }
  smax[threadIdx.x] = mx;
  smin[threadIdx.x] = mn;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 + 2 * blockIdx.x] = smax[0];
    part[3 + 2 * blockIdx.x] = smin[0];
  }
}

// stage 2: one block; out[0] = sqrt(max), out[1] = sqrt(min)
__global__ void k_vext_final(int nparts, double* __restrict__ part) {
  __shared__ double smax[kBlock], smin[kBlock];
  double mx = -1.0, mn = INFINITY;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    mx = fmax(mx, part[2 + 2 * i]);
    mn = fmin(mn, part[3 + 2 * i]);
  }
  smax[threadIdx.x] = mx;
  smin[threadIdx.x] = mn;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[0] = sqrt(smax[0]);
    part[1] = sqrt(smin[0]);
  }
}

struct ChainGeom {
  double bmin[3], bmax[3], L[3];
  int32_t periodic[3];
};

// PMPeriodicBoundary::point_vector (pm_periodic_boundary.C:350-374) from
// bead 1 to bead 2
__device__ __forceinline__ void bond_vector(const ChainGeom& g, const double* x, int b0, int b1, double* d) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double v = x[3 * int64_t(b1) + k] - x[3 * int64_t(b0) + k];
    if (g.periodic[k]) {
      if (v <= -g.L[k] / 2.0) v += g.L[k];
      if (v > g.L[k] / 2.0) v -= g.L[k];
    }
    d[k] = v;
  }
}

// any bond at least Ls long sets the flag (the positions are not changed)
__global__ void k_chain_detect(int64_t nb, ChainGeom g, const int2* __restrict__ bonds, const double* __restrict__ x,
                               double Ls, int32_t* __restrict__ flag) {
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int2 bd = bonds[b];
  double d[3];
  bond_vector(g, x, bd.x, bd.y, d);
  if (sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) >= Ls) flag[0] = 1;
}

// PolymerChain::check_chain's loop, bond by bond in one thread, run only
// when the detection found a long bond: each correction changes bead 2,
// which a later bond may read.
__global__ void k_chain_fix(int64_t nb, ChainGeom g, const int2* __restrict__ bonds, double* __restrict__ x,
                            double Ls, const int32_t* __restrict__ flag) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || flag[0] == 0) return;
  for (int64_t b = 0; b < nb; ++b) {
    const int2 bd = bonds[b];
    double d[3];
    bond_vector(g, x, bd.x, bd.y, d);
    const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len < Ls) continue;
    const double s = Ls * 0.9 / len;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double v = x[3 * int64_t(bd.x) + k] + d[k] * s;
      if (g.periodic[k]) {  // correct_position
        if (v < g.bmin[k]) v += g.L[k];
        if (v >= g.bmax[k]) v -= g.L[k];
      }
      x[3 * int64_t(bd.y) + k] = v;
    }
  }
}

int ensure_motion(ggem_ctx* ctx) {
  if (ctx->motion_cap >= ctx->n && ctx->x_saved) return GGEM_OK;
  size_t c1 = 0, c2 = 0, c3 = 0;
  for (void* q : {static_cast<void*>(ctx->x_saved), static_cast<void*>(ctx->rout), static_cast<void*>(ctx->images)})
    if (q) (void)hipFree(q);
  ctx->x_saved = ctx->rout = nullptr;
  ctx->images = nullptr;
  ctx->motion_cap = 0;
  ctx->saved_valid = ctx->rout_valid = false;
  const size_t n3 = size_t(3) * size_t(std::max<int64_t>(ctx->n, 1));
  int rc = ensure(ctx, &ctx->x_saved, &c1, n3);
  if (!rc) rc = ensure(ctx, &ctx->rout, &c2, n3);
  if (!rc) rc = ensure(ctx, &ctx->images, &c3, n3);
  if (!rc) ctx->motion_cap = ctx->n;
  return rc;
}

// ROUT = the current positions, zero image counts
int init_unwrapped(ggem_ctx* ctx) {
  const size_t bytes = sizeof(double) * 3 * size_t(ctx->n);
  if (bytes) {
    GGEM_HIP(ctx, hipMemcpyAsync(ctx->rout, ctx->x_in, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    GGEM_HIP(ctx, hipMemsetAsync(ctx->images, 0, sizeof(int32_t) * 3 * size_t(ctx->n), ctx->stream));
  }
  ctx->rout_valid = true;
  return GGEM_OK;
}

int ensure_noise(ggem_ctx* ctx, int nrhs) {
  const int64_t len = int64_t(nrhs) * 3 * ctx->n;
  if (len > ctx->noise_cap) {
    if (ctx->noise) (void)hipFree(ctx->noise);
    ctx->noise = nullptr;
    ctx->noise_cap = 0;
    GGEM_HIP(ctx, hipMalloc(&ctx->noise, sizeof(double) * std::max<int64_t>(len, 1)));
    ctx->noise_cap = len;
  }
  ctx->noise_nrhs = nrhs;
  return GGEM_OK;
}

}  // namespace

}  // namespace ggem

using namespace ggem;

extern "C" {

int ggem_noise_seed(ggem_ctx* ctx, uint64_t seed) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  ctx->noise_gen.seed(static_cast<std::default_random_engine::result_type>(seed));
  return GGEM_OK;
}

// n_vectors calls of std_random_vector(0.0, 1.0, "gaussian") whose results
// are dropped (copss.C:1550-1553); each call builds a fresh distribution
int ggem_noise_skip(ggem_ctx* ctx, int64_t n_vectors) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (n_vectors < 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "negative vector count");
  const int64_t len = 3 * ctx->n;
  for (int64_t v = 0; v < n_vectors; ++v) {
    std::normal_distribution<double> d(0.0, 1.0);
    d.reset();
    for (int64_t i = 0; i < len; ++i) (void)d(ctx->noise_gen);
  }
  return GGEM_OK;
}

int ggem_noise_draw(ggem_ctx* ctx, double dt, double* dw) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!(dt > 0)) return fail(ctx, GGEM_ERR_INVALID_ARG, "dt must be positive");
  const int64_t len = 3 * ctx->n;
  ctx->noise_h.resize(size_t(len));
  // random_vector_normal (random_generator.C:43-67) then VecScale(dw, sqrt 2)
  std::normal_distribution<double> dist(0.0, std::sqrt(dt));
  dist.reset();
  for (int64_t i = 0; i < len; ++i) ctx->noise_h[size_t(i)] = dist(ctx->noise_gen);
  const double s2 = std::sqrt(2.0);
  for (auto& v : ctx->noise_h) v *= s2;
  if (dw && len) std::copy(ctx->noise_h.begin(), ctx->noise_h.end(), dw);
  if (len == 0) return GGEM_OK;
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure_noise(ctx, 1);
  if (rc) return rc;
  if ((rc = h2d(ctx, ctx->noise, ctx->noise_h.data(), sizeof(double) * size_t(len)))) return rc;
  return sync(ctx);  // the host staging is reused by the next draw
}

int ggem_noise_philox(ggem_ctx* ctx, uint64_t seed, uint64_t counter, int32_t nrhs, double dt) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (nrhs < 1 || !(dt > 0)) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid noise arguments");
  const int64_t len = 3 * ctx->n;
  if (len == 0) return GGEM_OK;
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure_noise(ctx, nrhs);
  if (rc) return rc;
  ProfScope ps(ctx, "philox_noise");
  hipLaunchKernelGGL(k_philox_noise, dim3(blocks_for((len + 1) / 2)), dim3(kBlock), 0, ctx->stream, len, nrhs,
                     seed, counter, std::sqrt(2.0 * dt), ctx->noise);
  GGEM_HIP(ctx, hipGetLastError());
  return GGEM_OK;
}

int ggem_get_noise(ggem_ctx* ctx, double* dw) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!dw) return fail(ctx, GGEM_ERR_INVALID_ARG, "null output");
  if (!ctx->noise || ctx->noise_nrhs < 1) return fail(ctx, GGEM_ERR_NOT_READY, "no device noise");
  const size_t len = size_t(ctx->noise_nrhs) * 3 * size_t(ctx->n);
  return len ? d2h(ctx, dw, ctx->noise, sizeof(double) * len) : GGEM_OK;
}

int ggem_save_positions(ggem_ctx* ctx) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure_motion(ctx);
  if (rc) return rc;
  if (!ctx->rout_valid && (rc = init_unwrapped(ctx))) return rc;
  if (ctx->n)
    GGEM_HIP(ctx, hipMemcpyAsync(ctx->x_saved, ctx->x_in, sizeof(double) * 3 * size_t(ctx->n),
                                 hipMemcpyDeviceToDevice, ctx->stream));
  ctx->saved_valid = true;
  return GGEM_OK;
}

int ggem_move_particles(ggem_ctx* ctx, const ggem_move* mv) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!mv || (mv->wall_type != 0 && mv->wall_type != 1) || (mv->wall_type == 1 && !(mv->cavity_radius > 0)))
    return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid move arguments");
  if (ctx->nsys > 1) return fail(ctx, GGEM_ERR_INVALID_ARG, "ggem_move_particles needs a single-system context");
  if (mv->from_saved && !ctx->saved_valid) return fail(ctx, GGEM_ERR_NOT_READY, "no saved positions");
  if (mv->c_velocity != 0.0 && !ctx->u_out_valid)
    return fail(ctx, GGEM_ERR_NOT_READY, "no velocity: call ggem_combine_velocity first");
  if (mv->c_disp != 0.0 && (!ctx->keep_d_valid || ctx->keep_nrhs < 1))
    return fail(ctx, GGEM_ERR_NOT_READY, "no kept displacement (GGEM_CHEB_KEEP)");
  if (mv->c_noise != 0.0 && (!ctx->noise || ctx->noise_nrhs < 1))
    return fail(ctx, GGEM_ERR_NOT_READY, "no device noise");
  if (ctx->n == 0) return GGEM_OK;
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure_motion(ctx);
  if (rc) return rc;
  if (!ctx->rout_valid && (rc = init_unwrapped(ctx))) return rc;
  MoveArgs m{};
  m.cv = mv->c_velocity;
  m.cd = mv->c_disp;
  m.cf = mv->c_force;
  m.cn = mv->c_noise;
  for (int d = 0; d < 3; ++d) {
    m.bmin[d] = ctx->p.box_min[d];
    m.bmax[d] = ctx->p.box_max[d];
    m.L[d] = ctx->grid.L[d];
    m.periodic[d] = ctx->grid.periodic[d];
  }
  m.wall_type = mv->wall_type;
  m.radius = mv->cavity_radius;
  m.track = mv->track != 0;
  {
    ProfScope ps(ctx, "move");
    hipLaunchKernelGGL(k_move, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, m,
                       mv->from_saved ? ctx->x_saved : ctx->x_in, m.cv != 0.0 ? ctx->u_out : nullptr,
                       m.cd != 0.0 ? ctx->keep_d : nullptr, m.cf != 0.0 ? ctx->f_in : nullptr,
                       m.cn != 0.0 ? ctx->noise : nullptr, ctx->x_in, ctx->rout, ctx->images);
    GGEM_HIP(ctx, hipGetLastError());
  }
  // the positions moved: the slot copies, tensors and velocities describe
  // the old ones (the lists stay for ggem_update_positions)
  ctx->slots_current = false;
  ctx->u_out_valid = false;
  return GGEM_OK;
}

int ggem_get_positions(ggem_ctx* ctx, double* xyz, double* unwrapped, int32_t* images) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  const size_t n3 = 3 * size_t(ctx->n);
  if (n3 == 0) return GGEM_OK;
  int rc = GGEM_OK;
  if (xyz && (rc = d2h(ctx, xyz, ctx->x_in, sizeof(double) * n3))) return rc;
  if (unwrapped || images) {
    if ((rc = ensure_motion(ctx))) return rc;
    if (!ctx->rout_valid && (rc = init_unwrapped(ctx))) return rc;
    if (unwrapped && (rc = d2h(ctx, unwrapped, ctx->rout, sizeof(double) * n3))) return rc;
    if (images && (rc = d2h(ctx, images, ctx->images, sizeof(int32_t) * n3))) return rc;
  }
  return rc;
}

int ggem_velocity_extrema(ggem_ctx* ctx, int32_t which, double* vmax, double* vmin) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (which != 0 && which != 1) return fail(ctx, GGEM_ERR_INVALID_ARG, "which must be 0 (velocity) or 1 (forces)");
  if (which == 0 && !ctx->u_out_valid)
    return fail(ctx, GGEM_ERR_NOT_READY, "no velocity: call ggem_combine_velocity first");
  const double* v = which == 0 ? ctx->u_out : ctx->f_in;
  double out[2] = {0.0, 0.0};
  if (ctx->n > 0) {
    GGEM_HIP(ctx, hipSetDevice(ctx->device));
    if (!ctx->vext) GGEM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->vext), sizeof(double) * (2 + 2 * kExtBlocks)));
    const int nb = int(std::min<int64_t>(kExtBlocks, (ctx->n + kBlock - 1) / kBlock));
    ProfScope ps(ctx, "velocity_extrema");
    hipLaunchKernelGGL(k_vext_partial, dim3(nb), dim3(kBlock), 0, ctx->stream, ctx->n, v, ctx->vext);
    hipLaunchKernelGGL(k_vext_final, dim3(1), dim3(kBlock), 0, ctx->stream, nb, ctx->vext);
    GGEM_HIP(ctx, hipGetLastError());
    int rc = d2h(ctx, out, ctx->vext, sizeof(out));
    if (rc) return rc;
  }
  if (vmax) *vmax = out[0];
  if (vmin) *vmin = out[1];
  return GGEM_OK;
}

int ggem_check_chain(ggem_ctx* ctx, double Ls, int32_t* broken) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (ctx->nsys > 1) return fail(ctx, GGEM_ERR_INVALID_ARG, "ggem_check_chain needs a single-system context");
  if (!(Ls > 0)) return fail(ctx, GGEM_ERR_INVALID_ARG, "Ls must be positive");
  if (ctx->bond_n != ctx->n) return fail(ctx, GGEM_ERR_NOT_READY, "no bonds for these particles (ggem_set_bonds)");
  if (broken) *broken = 0;
  if (ctx->n_bonds == 0) return GGEM_OK;
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  if (!ctx->chain_flag) GGEM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->chain_flag), sizeof(int32_t)));
  ChainGeom g{};
  for (int d = 0; d < 3; ++d) {
    g.bmin[d] = ctx->p.box_min[d];
    g.bmax[d] = ctx->p.box_max[d];
    g.L[d] = ctx->grid.L[d];
    g.periodic[d] = ctx->grid.periodic[d];
  }
  {
    ProfScope ps(ctx, "check_chain");
    GGEM_HIP(ctx, hipMemsetAsync(ctx->chain_flag, 0, sizeof(int32_t), ctx->stream));
    hipLaunchKernelGGL(k_chain_detect, dim3(blocks_for(ctx->n_bonds)), dim3(kBlock), 0, ctx->stream, ctx->n_bonds, g,
                       ctx->bond_list, ctx->x_in, Ls, ctx->chain_flag);
    hipLaunchKernelGGL(k_chain_fix, dim3(1), dim3(64), 0, ctx->stream, ctx->n_bonds, g, ctx->bond_list, ctx->x_in, Ls,
                       ctx->chain_flag);
    GGEM_HIP(ctx, hipGetLastError());
  }
  // the positions may have changed: as after a move
  ctx->slots_current = false;
  ctx->u_out_valid = false;
  if (broken) return d2h(ctx, broken, ctx->chain_flag, sizeof(int32_t));
  return GGEM_OK;
}

int ggem_set_unwrapped(ggem_ctx* ctx, const double* rout) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure_motion(ctx);
  if (rc) return rc;
  if ((rc = init_unwrapped(ctx))) return rc;
  if (rout && ctx->n) rc = h2d(ctx, ctx->rout, rout, sizeof(double) * 3 * size_t(ctx->n));
  return rc ? rc : sync(ctx);
}

}  // extern "C"
