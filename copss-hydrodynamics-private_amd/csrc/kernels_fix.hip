// kernels_fix.hip -- the force-field fixes of point particles on the device:
// constant forces, LJ and Gaussian pair forces over the particle neighbour
// lists, worm-like and FENE springs over the bonds, LJ and empirical walls
// of a slit or a spherical cavity, and LJ forces of discretized wall
// particles.  One launch evaluates the whole list of fixes; their sum
// replaces the context's forces (PointMesh::reinit zeroes the forces and
// each fix's compute() adds its part, point_mesh.C:249-251).
//
// Reference: FixFactory::buildFix (fix/fix_factory.C:4-84) and the
// FixPoint* classes it builds; the force laws of FixBase (fix_base.C:
// spring_force_wls :45-62, spring_force_fene :66-84, gaussian_force
// :148-160, lj_force :163-182, polymer_wall_empirical_force :202-227).
//
// One thread per slot when the lists are built (the neighbour lists are
// read coalesced by wave), else one per particle; each thread adds the fixes
// in the order given, each fix's part summed on its own first, as every
// compute() sums into a local pforce before add_particle_force.
#include <vector>

#include "ggem_host.h"

namespace ggem {

namespace {

constexpr int kFixBlock = 256;
constexpr int kMaxFixes = 8;

struct FixArgs {
  Grid g;
  int64_t n;
  int32_t nfix;
  int32_t by_slot;
  ggem_fix fix[kMaxFixes];
  const double* x;          // positions, particle order
  const double4* x4;        // positions by slot (neighbour fixes)
  const int32_t* perm;
  const int32_t* cnt;
  const int64_t* slice_off;
  const uint32_t* ell;
  const int64_t* bond_off;  // [n + 1] incidences of each particle, bond order
  const int4* bond_inc;     // (bead 1, bead 2, 1 when the particle is bead 2, 0)
  const double* cforce;     // [3n] constant forces
  const double* wall_xyz;   // [3m] discretized wall particles
  int64_t n_wall;
  double* f;                // [3n] out, particle order
};

// PMPeriodicBoundary::point_vector (pm_periodic_boundary.C:350-374)
__device__ __forceinline__ void pvec(const Grid& g, const double* a, const double* b, double* d) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double v = b[k] - a[k];
    if (g.periodic[k]) {
      if (v <= -g.half_L[k]) v += g.L[k];
      if (v > g.half_L[k]) v -= g.L[k];
    }
    d[k] = v;
  }
}

__device__ __forceinline__ double norm3(const double* r) { return sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]); }

// FixBase::lj_force: -24 eps (2 (s/r)^12 - (s/r)^6) / r^2 r_ij
__device__ __forceinline__ void lj_add(const double* r, double eps, double sigma, double* out) {
  const double rn = norm3(r);
  const double r2 = rn * rn;
  const double sr2 = sigma * sigma / r2;
  const double sr6 = sr2 * sr2 * sr2;
  const double sr12 = sr6 * sr6;
  const double c = -24. * eps * (2. * sr12 - sr6) / r2;
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k] += c * r[k];
}

// FixBase::polymer_wall_empirical_force: -c0 (1 - |r|/d0)^2 r_hat within |d0|
__device__ __forceinline__ void empirical_add(const double* r, double c0, double d0, double* out) {
  const double rn = norm3(r);
  if (fabs(rn) < fabs(d0)) {
    const double q = 1 - rn / d0;
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] += -c0 * q * q * (r[k] / rn);
  }
}

// libMesh Point::unit() times (R - |x|): the particle -> sphere wall vector
__device__ __forceinline__ void to_sphere(const double* x, double radius, double* r) {
  const double xn = norm3(x);
#pragma unroll
  for (int k = 0; k < 3; ++k) r[k] = (x[k] / xn) * (radius - xn);
}

__global__ void __launch_bounds__(kFixBlock) k_apply_fixes(FixArgs a) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  const int64_t s = t;
  const int64_t p = a.by_slot ? int64_t(a.perm[t]) : t;
  const double xi[3] = {a.x[3 * p], a.x[3 * p + 1], a.x[3 * p + 2]};
  double f[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < a.nfix; ++q) {
    const ggem_fix& fx = a.fix[q];
    double pf[3] = {0.0, 0.0, 0.0};
    switch (fx.type) {
      case GGEM_FIX_P_CONSTANT:  // FixPointConstant::compute (fix_point_constant.C)
#pragma unroll
        for (int k = 0; k < 3; ++k) pf[k] = a.cforce[3 * p + k];
        break;
      case GGEM_FIX_LJ_CUT:     // FixPointLJCut::compute (fix_point_lj_cut.C:38-59)
      case GGEM_FIX_GAUSSIAN: {  // FixPointGaussian::compute (fix_point_gaussian.C)
        const double4 xs = a.x4[s];
        const double xo[3] = {xs.x, xs.y, xs.z};
        const int cnt = a.cnt[s];
        const int64_t base = a.slice_off[s >> 6] + (s & 63);
        for (int k = 0; k < cnt; ++k) {
          const double4 xt = a.x4[a.ell[base + int64_t(k) * kSlice] & kSlotMask];
          const double xn[3] = {xt.x, xt.y, xt.z};
          double r[3];
          pvec(a.g, xo, xn, r);  // the neighbour vector r_ij = x_j - x_i
          if (fx.type == GGEM_FIX_LJ_CUT) {
            if (norm3(r) <= fx.p[2]) lj_add(r, fx.p[0], fx.p[1], pf);
          } else {
            const double rn = norm3(r);
            const double c = -fx.p[0] * fx.p[1] * exp(-fx.p[1] * rn * rn);
#pragma unroll
            for (int d = 0; d < 3; ++d) pf[d] += c * r[d];
          }
        }
        break;
      }
      case GGEM_FIX_WORM_LIKE_SPRING:  // FixPointWLS::compute (fix_point_wls.C:66-96)
      case GGEM_FIX_FENE: {
        for (int64_t i = a.bond_off[p], e = a.bond_off[p + 1]; i < e; ++i) {
          const int4 b = a.bond_inc[i];
          double R[3];
          pvec(a.g, a.x + 3 * int64_t(b.x), a.x + 3 * int64_t(b.y), R);
          const double Rn = norm3(R);
          const double ratio = Rn / fx.p[1];
          double c;
          if (fx.type == GGEM_FIX_WORM_LIKE_SPRING) {  // c1 (1/(1 - r)^2 - 1 + 4 r)
            double v0 = 1.0 - ratio;
            if (fabs(v0) < 1e-3) v0 = v0 / fabs(v0) * 1e-3;
            const double tmp = 1. / (v0 * v0);
            c = fx.p[0] * (tmp - 1. + 4. * ratio);
          } else {  // (k Ls) r / (1 - r^2)
            double v0 = 1.0 - ratio * ratio;
            if (fabs(v0) < 1e-3) v0 = v0 / fabs(v0) * 1e-3 * 1e-3;
            c = fx.p[0] * fx.p[1] * ratio / v0;
          }
          // bead 1 takes F_ij, bead 2 F_ji = -F_ij, bond by bond
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double F = c * R[d] / Rn;
            f[d] += b.z ? -F : F;
          }
        }
        break;
      }
      case GGEM_FIX_SLIT_WALL_LJ_CUT:     // FixPointSlitWallLJCut::compute
      case GGEM_FIX_SLIT_WALL_EMPIRICAL_DNA:  // FixPointSlitWallEmpiricalDNA::compute
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (a.g.periodic[d]) continue;
          double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
          lo[d] = a.g.bmin[d] - xi[d];
          hi[d] = a.g.bmin[d] + a.g.L[d] - xi[d];
          double w[3] = {0.0, 0.0, 0.0};
          if (fx.type == GGEM_FIX_SLIT_WALL_LJ_CUT) {
            if (norm3(lo) <= fx.p[2]) lj_add(lo, fx.p[0], fx.p[1], w);
            pf[d] += w[d];
            w[0] = w[1] = w[2] = 0.0;
            if (norm3(hi) <= fx.p[2]) lj_add(hi, fx.p[0], fx.p[1], w);
            pf[d] += w[d];
          } else {
            empirical_add(lo, fx.p[0], fx.p[1], w);
            pf[d] += w[d];
            w[0] = w[1] = w[2] = 0.0;
            empirical_add(hi, fx.p[0], fx.p[1], w);
            pf[d] += w[d];
          }
        }
        break;
      case GGEM_FIX_SPHERE_WALL_LJ_CUT: {  // FixPointSphereWallLJCut::compute
        double r[3];
        to_sphere(xi, fx.p[3], r);
        if (norm3(r) <= fx.p[2]) lj_add(r, fx.p[0], fx.p[1], pf);
        break;
      }
      case GGEM_FIX_SPHERE_WALL_EMPIRICAL_DNA: {  // FixPointSphereWallEmpiricalDNA::compute
        double r[3];
        to_sphere(xi, fx.p[2], r);
        empirical_add(r, fx.p[0], fx.p[1], pf);
        break;
      }
      case GGEM_FIX_DISCRETIZED_WALL_LJ_CUT:  // FixPointDiscretizedWallLJCut::compute
        for (int64_t j = 0; j < a.n_wall; ++j) {
          double r[3];
          pvec(a.g, xi, a.wall_xyz + 3 * j, r);
          if (norm3(r) <= fx.p[2]) lj_add(r, fx.p[0], fx.p[1], pf);
        }
        break;
      default:
        break;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) f[d] += pf[d];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) a.f[3 * p + d] = f[d];
}

bool uses_lists(int type) { return type == GGEM_FIX_LJ_CUT || type == GGEM_FIX_GAUSSIAN; }
bool uses_bonds(int type) { return type == GGEM_FIX_WORM_LIKE_SPRING || type == GGEM_FIX_FENE; }

}  // namespace

}  // namespace ggem

using namespace ggem;

extern "C" {

int ggem_set_bonds(ggem_ctx* ctx, int64_t n_bonds, const int64_t* bonds) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (ctx->nsys > 1) return fail(ctx, GGEM_ERR_INVALID_ARG, "ggem_set_bonds needs a single-system context");
  if (n_bonds < 0 || (n_bonds > 0 && !bonds)) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid bond array");
  const int64_t n = ctx->n;
  for (int64_t b = 0; b < 2 * n_bonds; ++b)
    if (bonds[b] < 0 || bonds[b] >= n)
      return fail(ctx, GGEM_ERR_INVALID_ARG, "bond " + std::to_string(b / 2) + " names bead " +
                                                 std::to_string(bonds[b]) + " of " + std::to_string(n));
  // the incidences of each bead in bond order (a bead's forces add bond by bond)
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t b = 0; b < n_bonds; ++b) {
    ++off[bonds[2 * b] + 1];
    ++off[bonds[2 * b + 1] + 1];
  }
  for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
  std::vector<int4> inc(std::max<int64_t>(2 * n_bonds, 1));
  {
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    for (int64_t b = 0; b < n_bonds; ++b) {
      const int32_t b1 = int32_t(bonds[2 * b]), b2 = int32_t(bonds[2 * b + 1]);
      inc[pos[b1]++] = make_int4(b1, b2, 0, 0);
      inc[pos[b2]++] = make_int4(b1, b2, 1, 0);
    }
  }
  std::vector<int2> list(std::max<int64_t>(n_bonds, 1));
  for (int64_t b = 0; b < n_bonds; ++b) list[b] = make_int2(int32_t(bonds[2 * b]), int32_t(bonds[2 * b + 1]));
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure(ctx, &ctx->bond_off, &ctx->bond_off_cap, size_t(n + 1));
  if (!rc) rc = ensure(ctx, &ctx->bond_inc, &ctx->bond_inc_cap, inc.size());
  if (!rc) rc = ensure(ctx, &ctx->bond_list, &ctx->bond_list_cap, list.size());
  if (rc) return rc;
  if ((rc = h2d(ctx, ctx->bond_off, off.data(), sizeof(int64_t) * (n + 1)))) return rc;
  if ((rc = h2d(ctx, ctx->bond_inc, inc.data(), sizeof(int4) * inc.size()))) return rc;
  if ((rc = h2d(ctx, ctx->bond_list, list.data(), sizeof(int2) * list.size()))) return rc;
  ctx->bond_n = n;
  ctx->n_bonds = n_bonds;
  return sync(ctx);
}

int ggem_set_constant_forces(ggem_ctx* ctx, const double* f) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!f && ctx->n > 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "null forces");
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  const size_t len = size_t(3) * size_t(ctx->n);
  int rc = ensure(ctx, &ctx->cforce, &ctx->cforce_cap, std::max<size_t>(len, 1));
  if (rc) return rc;
  if (len && (rc = h2d(ctx, ctx->cforce, f, sizeof(double) * len))) return rc;
  ctx->cforce_n = ctx->n;
  return sync(ctx);
}

int ggem_set_wall_particles(ggem_ctx* ctx, int64_t m, const double* xyz) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (m < 0 || (m > 0 && !xyz)) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid wall particles");
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  int rc = ensure(ctx, &ctx->wall_xyz, &ctx->wall_cap, size_t(std::max<int64_t>(3 * m, 1)));
  if (rc) return rc;
  if (m && (rc = h2d(ctx, ctx->wall_xyz, xyz, sizeof(double) * 3 * size_t(m)))) return rc;
  ctx->n_wall = m;
  return sync(ctx);
}

int ggem_apply_fixes(ggem_ctx* ctx, int32_t nfix, const ggem_fix* fixes, double* f) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (ctx->nsys > 1) return fail(ctx, GGEM_ERR_INVALID_ARG, "ggem_apply_fixes needs a single-system context");
  if (nfix < 0 || nfix > kMaxFixes || (nfix > 0 && !fixes))
    return fail(ctx, GGEM_ERR_INVALID_ARG, "between 0 and 8 fixes");
  bool lists = false;
  for (int q = 0; q < nfix; ++q) {
    const ggem_fix& x = fixes[q];
    const std::string which = "fix " + std::to_string(q) + ": ";
    if (x.type < GGEM_FIX_P_CONSTANT || x.type > GGEM_FIX_DISCRETIZED_WALL_LJ_CUT)
      return fail(ctx, GGEM_ERR_INVALID_ARG, which + "unknown type " + std::to_string(x.type));
    if (uses_lists(x.type)) lists = true;
    if (uses_bonds(x.type) && (ctx->bond_n != ctx->n || !(x.p[1] > 0)))
      return fail(ctx, GGEM_ERR_NOT_READY, which + "springs need ggem_set_bonds for these particles and Ls > 0");
    if (x.type == GGEM_FIX_P_CONSTANT && ctx->cforce_n != ctx->n)
      return fail(ctx, GGEM_ERR_NOT_READY, which + "no constant forces for these particles (ggem_set_constant_forces)");
    if ((x.type == GGEM_FIX_LJ_CUT || x.type == GGEM_FIX_SLIT_WALL_LJ_CUT || x.type == GGEM_FIX_SPHERE_WALL_LJ_CUT ||
         x.type == GGEM_FIX_DISCRETIZED_WALL_LJ_CUT) && !(x.p[1] > 0 && x.p[2] >= 0))
      return fail(ctx, GGEM_ERR_INVALID_ARG, which + "lj_cut needs sigma > 0 and rcut >= 0");
  }
  int rc;
  if (lists) {
    if ((rc = check_ready(ctx))) return rc;
    if ((rc = ensure_ell(ctx))) return rc;
  }
  if (ctx->n == 0) return GGEM_OK;
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  FixArgs a{};
  a.g = ctx->grid;
  a.n = ctx->n;
  a.nfix = nfix;
  a.by_slot = ctx->built && ctx->slots_current ? 1 : 0;
  for (int q = 0; q < nfix; ++q) a.fix[q] = fixes[q];
  a.x = ctx->x_in;
  a.x4 = ctx->x4;
  a.perm = ctx->perm;
  a.cnt = ctx->nbr_cnt;
  a.slice_off = ctx->slice_off;
  a.ell = ctx->ell;
  a.bond_off = ctx->bond_off;
  a.bond_inc = ctx->bond_inc;
  a.cforce = ctx->cforce;
  a.wall_xyz = ctx->wall_xyz;
  a.n_wall = ctx->n_wall;
  a.f = ctx->f_in;
  {
    ProfScope ps(ctx, "apply_fixes");
    hipLaunchKernelGGL(k_apply_fixes, dim3(unsigned((ctx->n + kFixBlock - 1) / kFixBlock)), dim3(kFixBlock), 0,
                       ctx->stream, a);
    GGEM_HIP(ctx, hipGetLastError());
  }
  if (ctx->built) {
    const hipError_t e = launch_gather_forces(ctx);
    if (e != hipSuccess) return hip_fail(ctx, e, "gather_forces");
  }
  ctx->u_out_valid = false;
  if (f) return d2h(ctx, f, ctx->f_in, sizeof(double) * 3 * size_t(ctx->n));
  return GGEM_OK;
}

}  // extern "C"
