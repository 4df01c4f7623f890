// kernels_fe.hip -- the FEM side of the particle path on the device: the
// HEX20 interpolant of the nodal FEM velocity at each bead, in the element
// the bead was binned to, and the element force vectors of the spreading
// summed into one force per node.
//
// Reference for the nodal force: AssembleStokes::assemble_global_F adds each
// element's Fe into the global RHS by its dof indices
// (rhs->add_vector(Fe, _dof_indices[elem_id]), assemble_stokes.C:508-513);
// with one velocity dof per node and component, the point-force part is
// F[node_l][d] += Fe[e][d][l].
//
// Reference: PMSystemStokes::compute_point_velocity takes
// Uglobal(d) = point_value(var_d, pt, *elem) (pm_system_stokes.C:441-443),
// libMesh's System::point_value = sum_l phi_l(xi) u[dof_l] over the
// element's 20 velocity dofs in node order, with xi the point mapped to the
// reference cube [-1, 1]^3.  The generated meshes are axis-aligned boxes of
// size h (build_cube, copss.C:995-1004), so the map is xi = 2 (x - c) / h
// with c the element centre, and the second-order shape functions are the
// serendipity ones of HEX20 (8 vertices, then the 12 edge mid-nodes
// 0-1, 1-2, 2-3, 3-0, 0-4, 1-5, 2-6, 3-7, 4-5, 5-6, 6-7, 7-4).
//
// One thread per slot: the beads of one element sit in consecutive slots,
// so a wave reads each element's 20 node numbers and 60 values once
// through its L1, and the only scattered access is the 24-byte result in
// particle order.  Blocks are renumbered so that each XCD takes one
// contiguous range of slots (whole bricks of cells): the nodes that
// neighbouring elements share are then fetched into one L2, not several.
#include "ggem_device.h"

namespace ggem {

namespace {

constexpr int kFeBlock = 256;
constexpr int kFeGroup = 10;  // nodal values loaded per group

// phi_l(a, b, c) of the 20 HEX20 serendipity functions in libMesh's order
__device__ __forceinline__ void hex20_phi(double a, double b, double c, double* phi) {
  const double am = 1.0 - a, ap = 1.0 + a, bm = 1.0 - b, bp = 1.0 + b, cm = 1.0 - c, cp = 1.0 + c;
  // vertices: 1/8 (1 + a a_l)(1 + b b_l)(1 + c c_l)(a a_l + b b_l + c c_l - 2)
  phi[0] = 0.125 * am * bm * cm * (-a - b - c - 2.0);
  phi[1] = 0.125 * ap * bm * cm * (a - b - c - 2.0);
  phi[2] = 0.125 * ap * bp * cm * (a + b - c - 2.0);
  phi[3] = 0.125 * am * bp * cm * (-a + b - c - 2.0);
  phi[4] = 0.125 * am * bm * cp * (-a - b + c - 2.0);
  phi[5] = 0.125 * ap * bm * cp * (a - b + c - 2.0);
  phi[6] = 0.125 * ap * bp * cp * (a + b + c - 2.0);
  phi[7] = 0.125 * am * bp * cp * (-a + b + c - 2.0);
  // edge mid-nodes: 1/4 (1 - t^2) along the edge's axis times the two
  // linear factors of its position on the other axes
  const double qa = am * ap, qb = bm * bp, qc = cm * cp;
  phi[8] = 0.25 * qa * bm * cm;   // 0-1
  phi[9] = 0.25 * ap * qb * cm;   // 1-2
  phi[10] = 0.25 * qa * bp * cm;  // 2-3
  phi[11] = 0.25 * am * qb * cm;  // 3-0
  phi[12] = 0.25 * am * bm * qc;  // 0-4
  phi[13] = 0.25 * ap * bm * qc;  // 1-5
  phi[14] = 0.25 * ap * bp * qc;  // 2-6
  phi[15] = 0.25 * am * bp * qc;  // 3-7
  phi[16] = 0.25 * qa * bm * cp;  // 4-5
  phi[17] = 0.25 * ap * qb * cp;  // 5-6
  phi[18] = 0.25 * qa * bp * cp;  // 6-7
  phi[19] = 0.25 * am * qb * cp;  // 7-4
}

__global__ __launch_bounds__(kFeBlock) void k_interp_hex20(Grid g, int64_t n, const int32_t* __restrict__ perm,
                                                           const int32_t* __restrict__ slot_cell,
                                                           const double4* __restrict__ x4,
                                                           const int32_t* __restrict__ conn,
                                                           const double* __restrict__ node_u,
                                                           double* __restrict__ out) {
  const int64_t s = xcd_block(blockIdx.x, gridDim.x) * kFeBlock + threadIdx.x;
  if (s >= n) return;
  int ci, cj, ck;
  cell_ijk(g, slot_cell[s], ci, cj, ck);
  const int64_t e = ci + int64_t(g.n[0]) * (cj + int64_t(g.n[1]) * ck);
  const double4 p = x4[s];
  const double a = 2.0 * (p.x - (g.bmin[0] + (double(ci) + 0.5) * g.h[0])) / g.h[0];
  const double b = 2.0 * (p.y - (g.bmin[1] + (double(cj) + 0.5) * g.h[1])) / g.h[1];
  const double c = 2.0 * (p.z - (g.bmin[2] + (double(ck) + 0.5) * g.h[2])) / g.h[2];
  double phi[20];
  hex20_phi(a, b, c, phi);
  const int32_t* cn = conn + 20 * e;
  int nds[20];
#pragma unroll
  for (int l = 0; l < 20; ++l) nds[l] = cn[l];
  // the nodal values in two groups of ten, each group's 30 loads issued
  // together before its sums (one load at a time left the kernel waiting
  // on memory latency; all 20 at once cost occupancy); sums in node order
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
#pragma unroll
  for (int l0 = 0; l0 < 20; l0 += kFeGroup) {
    double v[3 * kFeGroup];
#pragma unroll
    for (int j = 0; j < kFeGroup; ++j) {
      const int64_t nd = nds[l0 + j];
      v[3 * j] = node_u[3 * nd];
      v[3 * j + 1] = node_u[3 * nd + 1];
      v[3 * j + 2] = node_u[3 * nd + 2];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < kFeGroup; ++j) {
      u0 += phi[l0 + j] * v[3 * j];
      u1 += phi[l0 + j] * v[3 * j + 1];
      u2 += phi[l0 + j] * v[3 * j + 2];
    }
  }
  const int64_t q = perm[s];
  out[3 * q] = u0;
  out[3 * q + 1] = u1;
  out[3 * q + 2] = u2;
}

// Nodal force: each node sums the element vectors Fe[e][l][d] (the
// spreading's node-major layout: one 24-byte read per incidence) of its
// incidences (e, l) in ascending order.  A gather, so no atomics and the
// same sums on every run.  Threads take the nodes brick by brick of their
// first element (order), which keeps a brick's element vectors in L2 while
// its nodes are summed; in plain node order each element was fetched once
// per plane of nodes touching it.
__global__ __launch_bounds__(kFeBlock) void k_gather_nodes(int64_t nn, const int32_t* __restrict__ order,
                                                           const int64_t* __restrict__ off,
                                                           const int32_t* __restrict__ inc,
                                                           const double* __restrict__ Fe,
                                                           double* __restrict__ out) {
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * kFeBlock + threadIdx.x;
  if (t >= nn) return;
  const int64_t v = order[t];
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  const int64_t k1 = off[v + 1];
  for (int64_t k = off[v]; k < k1; ++k) {
    const int64_t el = inc[k];
    const int64_t e = el / 20, l = el - 20 * e;
    const double* fe = Fe + 60 * e + 3 * l;
    f0 += fe[0];
    f1 += fe[1];
    f2 += fe[2];
  }
  out[3 * v] = f0;
  out[3 * v + 1] = f1;
  out[3 * v + 2] = f2;
}

}  // namespace

hipError_t launch_gather_nodes(ggem_ctx* ctx, const double* Fe, double* out, bool condensed) {
  const int64_t nn = ctx->fe_nodes;
  if (nn <= 0) return hipSuccess;
  // without a periodic axis the condensed incidences are the plain ones
  const bool c = condensed && ctx->fe_cinc;
  ProfScope ps(ctx, "gather_nodes");
  hipLaunchKernelGGL(k_gather_nodes, dim3(unsigned((nn + kFeBlock - 1) / kFeBlock)), dim3(kFeBlock), 0,
                     ctx->stream, nn, ctx->fe_node_order, c ? ctx->fe_cinc_off : ctx->fe_inc_off,
                     c ? ctx->fe_cinc : ctx->fe_inc, Fe, out);
  return hipGetLastError();
}

hipError_t launch_interp_hex20(ggem_ctx* ctx, const int32_t* conn, const double* node_u, double* out) {
  const int64_t n = ctx->n;
  if (n == 0) return hipSuccess;
  ProfScope ps(ctx, "interp_hex20");
  hipLaunchKernelGGL(k_interp_hex20, dim3(unsigned((n + kFeBlock - 1) / kFeBlock)), dim3(kFeBlock), 0,
                     ctx->stream, ctx->grid, n, ctx->perm, ctx->slot_cell, ctx->x4, conn, node_u, out);
  return hipGetLastError();
}

}  // namespace ggem

// ---- force fields: FENE bead-spring chains and LJ slit walls -------------
//
// Reference: FixBase::spring_force_fene (fix_base.C:78-97) between
// consecutive beads of a chain, F = c1 (r/Ls) / (1 - (r/Ls)^2) r_hat with
// c1 = k Ls, Ls = b_max, the bond vector by the minimum image on periodic
// axes; and FixPointSlitWallLJCut's wall force, FixBase::lj_force
// (fix_base.C:168-186) of the particle -> wall vector within rcut, for the
// two walls of every non-periodic axis.  One thread per particle in
// particle order: f = F(bond to the next bead) - F(bond from the previous
// bead), then the walls axis by axis, low wall first.
namespace ggem {

namespace {

__device__ __forceinline__ void fene_bond(const Grid& g, const double* x, int64_t i, double k, double b,
                                          double* fb) {
  double r[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double v = x[3 * (i + 1) + d] - x[3 * i + d];
    if (g.periodic[d]) {
      if (v <= -g.half_L[d]) v += g.L[d];
      if (v > g.half_L[d]) v -= g.L[d];
    }
    r[d] = v;
  }
  const double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  const double ratio = R / b;
  double val0 = 1.0 - ratio * ratio;
  constexpr double kTol = 1e-3;  // FixBase's singularity guard (fix_base.h:50)
  if (fabs(val0) < kTol) val0 = val0 / fabs(val0) * kTol * kTol;
  const double c = (k * b) * ratio / val0;
#pragma unroll
  for (int d = 0; d < 3; ++d) fb[d] = c * r[d] / R;
}

struct WallPos {
  double w[3][2];  // the box faces as given (box_min, box_max)
};

__global__ void k_fix_forces(Grid g, WallPos wp, int64_t n, const double* __restrict__ x, int32_t chain,
                             double k, double b, int32_t walls, double eps, double sigma, double rcut,
                             double* __restrict__ f) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double fi[3] = {0.0, 0.0, 0.0};
  if (chain > 1) {
    const int64_t j = i % chain;
    double fb[3];
    if (j + 1 < chain) {
      fene_bond(g, x, i, k, b, fb);
#pragma unroll
      for (int d = 0; d < 3; ++d) fi[d] += fb[d];
    }
    if (j > 0) {
      fene_bond(g, x, i - 1, k, b, fb);
#pragma unroll
      for (int d = 0; d < 3; ++d) fi[d] -= fb[d];
    }
  }
  if (walls) {
    const double s2 = sigma * sigma;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (g.periodic[d]) continue;
      const double xd = x[3 * i + d];
#pragma unroll
      for (int w = 0; w < 2; ++w) {
        const double rw = wp.w[d][w] - xd;
        if (fabs(rw) <= rcut) {
          const double r2 = rw * rw;
          const double q = s2 / r2;
          const double sr6 = q * q * q;
          fi[d] += -24.0 * eps * (2.0 * sr6 * sr6 - sr6) / r2 * rw;
        }
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) f[3 * i + d] = fi[d];
}

}  // namespace

hipError_t launch_fix_forces(ggem_ctx* ctx, int32_t chain, double k, double b, int32_t walls, double eps,
                             double sigma, double rcut) {
  if (ctx->n == 0) return hipSuccess;
  WallPos wp;
  for (int d = 0; d < 3; ++d) {
    wp.w[d][0] = ctx->p.box_min[d];
    wp.w[d][1] = ctx->p.box_max[d];
  }
  ProfScope ps(ctx, "fix_forces");
  hipLaunchKernelGGL(k_fix_forces, dim3(unsigned((ctx->n + kFeBlock - 1) / kFeBlock)), dim3(kFeBlock), 0,
                     ctx->stream, ctx->grid, wp, ctx->n, ctx->x_in, chain, k, b, walls, eps, sigma, rcut,
                     ctx->f_in);
  return hipGetLastError();
}

}  // namespace ggem
