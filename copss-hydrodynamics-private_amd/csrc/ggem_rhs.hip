// ggem_rhs.hip -- the right-hand side of the global Stokes system on the
// device: element vectors after the spreading (or the inlet traction) and
// the penalty rows of the no-slip walls, and their sum into one force per
// velocity node with the periodic images condensed.
//
// Reference: AssembleStokes::assemble_global_F (assemble_stokes.C:473-516):
// per element compute_element_rhs (point-force spreading, :643-672, or the
// inlet traction of the "undisturbed" option, :687-727), apply_bc_by_penalty
// (:889-996) on the sides that select_boundary_side keeps (:839-885: no
// neighbour, not on a periodic or inlet face), constrain_element_vector
// (:508) and rhs->add_vector (:513).  The penalty adds penalty * uvw to
// Fe(d n_nodes + l) per side node (penalize_elem_matrix_vector,
// assemble_system.C:70-107) with uvw the side's undisturbed velocity
// (shear, :929-948) minus the local velocity local_velocity_fluid(node,
// elem) (pm_system_stokes.C:572-593 -> GGEMStokes::local_solution_field),
// or the unbounded solution minus it in the validation case (:949-976,
// AnalyticalSolutionStokes::exact_solution_infinite_domain,
// analytical_solution_stokes.C:66-102).
//
// The sides on walls and inlets depend on the grid alone: the host lists
// one item per (element, local node) on such a side, with bit masks of the
// element's wall and inlet sides holding that node (libMesh's hex side
// order: 0 z-, 1 y-, 2 x+, 3 y+, 4 x-, 5 z+).  The local velocity of all
// items is one field-point pair sum; then each item thread adds its sides'
// traction and penalty terms, sides ascending, to the element vector entry
// that it alone owns.  The whole right-hand side is a few launches over the
// boundary items plus the spreading, which dominates.
#include <cmath>
#include <vector>

#include "ggem_device.h"
#include "ggem_host.h"

namespace ggem {

namespace {

constexpr int kRhsBlock = 256;

// HEX20 nodes on the reference cube (libMesh order: the vertices, then the
// mid-nodes of edges 0-1, 1-2, 2-3, 3-0, 0-4, 1-5, 2-6, 3-7, 4-5, 5-6, 6-7,
// 7-4), hex side s: its axis, its face (-1 / +1) and its first, second and
// fourth vertex (libMesh's side node order, which fixes the side's
// reference axes and so the order of its face quadrature points)
constexpr int kVert[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                             {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
constexpr int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                              {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
constexpr int kSideAxis[6] = {2, 1, 0, 1, 0, 2};
constexpr int kSideFace[6] = {-1, -1, 1, 1, -1, 1};
constexpr int kSideVert[6][3] = {{0, 3, 1}, {0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}, {4, 5, 7}};

void node_ref(int l, int* xi) {
  for (int d = 0; d < 3; ++d)
    xi[d] = l < 8 ? kVert[l][d] : (kVert[kEdge[l - 8][0]][d] + kVert[kEdge[l - 8][1]][d]) / 2;
}

// bit s set when node l lies on hex side s
int node_sides(int l) {
  int xi[3];
  node_ref(l, xi);
  int m = 0;
  for (int s = 0; s < 6; ++s)
    if (xi[kSideAxis[s]] == kSideFace[s]) m |= 1 << s;
  return m;
}

// the 20 HEX20 shape functions at a reference point (host)
void hex20_phi_host(const double* X, double* phi) {
  for (int l = 0; l < 20; ++l) {
    int xi[3];
    node_ref(l, xi);
    const double fa = 1.0 + X[0] * xi[0], fb = 1.0 + X[1] * xi[1], fc = 1.0 + X[2] * xi[2];
    if (l < 8)
      phi[l] = 0.125 * fa * fb * fc * (X[0] * xi[0] + X[1] * xi[1] + X[2] * xi[2] - 2.0);
    else if (xi[0] == 0)
      phi[l] = 0.25 * (1.0 - X[0] * X[0]) * fb * fc;
    else if (xi[1] == 0)
      phi[l] = 0.25 * fa * (1.0 - X[1] * X[1]) * fc;
    else
      phi[l] = 0.25 * fa * fb * (1.0 - X[2] * X[2]);
  }
}

// JxW_face[qp] phi_l(qp) of every side (compute_element_rhs, :687-727 with
// QGauss(2, SECOND) on the side: 2 x 2 Gauss points, the side's first
// reference axis fastest; JxW = h_u h_v / 4 on the axis-aligned box)
void traction_table(const double* h, double* tab /* [6][4][20] */) {
  const double gp[2] = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)};
  for (int s = 0; s < 6; ++s) {
    const int* v0 = kVert[kSideVert[s][0]];
    const int* v1 = kVert[kSideVert[s][1]];
    const int* v3 = kVert[kSideVert[s][2]];
    double jxw = 0.25;
    for (int d = 0; d < 3; ++d)
      if (d != kSideAxis[s]) jxw *= h[d];
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        double X[3], phi[20];
        for (int d = 0; d < 3; ++d)
          X[d] = v0[d] + 0.5 * (gp[i] + 1.0) * (v1[d] - v0[d]) + 0.5 * (gp[j] + 1.0) * (v3[d] - v0[d]);
        hex20_phi_host(X, phi);
        for (int l = 0; l < 20; ++l) tab[(s * 4 + i + 2 * j) * 20 + l] = jxw * phi[l];
      }
  }
}

int rhs_key(const ggem_rhs_opts* o) {
  return (o->inlet[0] ? 1 : 0) | (o->inlet[1] ? 2 : 0) | (o->inlet[2] ? 4 : 0) |
         (o->option == GGEM_RHS_UNDISTURBED ? 8 : 0);
}

// wall / inlet items of every boundary element, elements and nodes ascending
void build_items(const int32_t* n, const int32_t* periodic, const ggem_rhs_opts* o, std::vector<int2>& items) {
  items.clear();
  const bool undist = o->option == GGEM_RHS_UNDISTURBED;
  int nm[20];
  for (int l = 0; l < 20; ++l) nm[l] = node_sides(l);
  const int n0 = n[0], n1 = n[1], n2 = n[2];
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < n1; ++j) {
      const bool jk_inner = j > 0 && j < n1 - 1 && k > 0 && k < n2 - 1;
      for (int i = 0; i < n0; ++i) {
        if (jk_inner && i > 0 && i < n0 - 1) {
          i = n0 - 2;  // skip to the last element of the row
          continue;
        }
        const int idx[3] = {i, j, k};
        int wm = 0, tm = 0;
        for (int s = 0; s < 6; ++s) {
          const int a = kSideAxis[s];
          const bool on = kSideFace[s] < 0 ? idx[a] == 0 : idx[a] == n[a] - 1;
          if (!on) continue;
          if (!periodic[a] && !o->inlet[a]) wm |= 1 << s;
          if (undist && o->inlet[a]) tm |= 1 << s;
        }
        if (!(wm | tm)) continue;
        const int64_t e = i + int64_t(n0) * (j + int64_t(n1) * k);
        for (int l = 0; l < 20; ++l)
          if (nm[l] & (wm | tm)) items.push_back(make_int2(int(e), l | ((nm[l] & wm) << 8) | ((nm[l] & tm) << 16)));
      }
    }
}

// Node point of item q in the coordinates of libMesh's build_cube (a unit
// lattice i / (2 n) scaled to the box: x = lo + (hi - lo) i / (2 n)) and
// the item's element.
__global__ void k_item_points(Grid g, int64_t m, const int2* __restrict__ items, double* __restrict__ pts,
                              int32_t* __restrict__ elem) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const int2 it = items[q];
  int ijk[3];
  elem_ijk(g, it.x, ijk[0], ijk[1], ijk[2]);
  const int l = it.y & 0xFF;
  // lattice offset of node l within its element, in {0, 1, 2}^3
  const int v = l < 8 ? l : 0;
  const int a = l < 8 ? l : (l < 12 ? l - 8 : (l < 16 ? l - 12 : l - 16));
  int off[3];
  if (l < 8) {
    off[0] = (v == 1 || v == 2 || v == 5 || v == 6) ? 2 : 0;
    off[1] = (v == 2 || v == 3 || v == 6 || v == 7) ? 2 : 0;
    off[2] = v >= 4 ? 2 : 0;
  } else if (l < 12 || l >= 16) {
    // edges of the bottom (8-11) and top (16-19) faces: 0-1, 1-2, 2-3, 3-0
    const int z = l >= 16 ? 2 : 0;
    const int ox[4] = {1, 2, 1, 0}, oy[4] = {0, 1, 2, 1};
    off[0] = ox[a];
    off[1] = oy[a];
    off[2] = z;
  } else {
    // vertical edges 0-4, 1-5, 2-6, 3-7
    const int ox[4] = {0, 2, 2, 0}, oy[4] = {0, 0, 2, 2};
    off[0] = ox[a];
    off[1] = oy[a];
    off[2] = 1;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double t = double(2 * ijk[d] + off[d]) / double(2 * g.n[d]);
    pts[3 * q + d] = t * g.L[d] + g.bmin[d];
  }
  elem[q] = it.x;
}

// AnalyticalSolutionStokes::exact_solution_infinite_domain: the unbounded
// smoothed tensor with xi (GGEMStokes::green_tensor_unbounded_smoothed,
// ggem_stokes.C:201-251) at pt - x_i, no periodic image, summed over all
// particles in particle order.
__global__ void k_unbounded(GGEMConst k, int64_t n, const double* __restrict__ x, const double* __restrict__ f,
                            int64_t m, const double* __restrict__ pts, double* __restrict__ u) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const double p0 = pts[3 * q], p1 = pts[3 * q + 1], p2 = pts[3 * q + 2];
  const double a = k.ksi;
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double d[3] = {p0 - x[3 * i], p1 - x[3 * i + 1], p2 - x[3 * i + 2]};
    const double fi[3] = {f[3 * i], f[3 * i + 1], f[3 * i + 2]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double G[3][3];
    if (r < 1e-6) {  // GGEMSystem::r_eps (ggem_system.h:116)
      const double c0 = k.eps * 4.0 * a / 1.77245385090551602729;
#pragma unroll
      for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int t = 0; t < 3; ++t) G[s][t] = s == t ? c0 : 0.0;
    } else {
      const double r2 = r * r;
      const double c1 = erf(a * r) / r;
      const double c2 = 2.0 * a / 1.77245385090551602729 * exp(-(a * a) * r2);
#pragma unroll
      for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          const double dij = s == t ? 1.0 : 0.0, xx = d[s] * d[t] / r2;
          G[s][t] = ((dij + xx) * c1 + (dij - xx) * c2) * k.eps;
        }
    }
    u0 += G[0][0] * fi[0];
    u0 += G[0][1] * fi[1];
    u0 += G[0][2] * fi[2];
    u1 += G[1][0] * fi[0];
    u1 += G[1][1] * fi[1];
    u1 += G[1][2] * fi[2];
    u2 += G[2][0] * fi[0];
    u2 += G[2][1] * fi[1];
    u2 += G[2][2] * fi[2];
  }
  u[3 * q] = u0;
  u[3 * q + 1] = u1;
  u[3 * q + 2] = u2;
}

struct RhsArgs {
  int64_t m;
  const int2* items;
  int32_t undist, validation, node_major;
  int32_t shear[3], sdir[3];
  double rate[3], pin[3], lo[3], hi[3];
  double penalty;
  const double* tab;   // [6][4][20] (inlet traction)
  const double* uloc;  // [3m]
  const double* uex;   // [3m] (validation)
  double* Fe;          // [n_elem][3][20], node_major: [n_elem][20][3]
};

// The traction terms, then the penalty terms, of the item's sides in side
// order, each added on its own as the reference's loops do (products
// rounded before the sum: no contraction into fused multiply-adds).
__global__ void k_rhs_items(RhsArgs a) {
#pragma clang fp contract(off)
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.m) return;
  const int2 it = a.items[q];
  const int l = it.y & 0xFF, wm = (it.y >> 8) & 63, tm = (it.y >> 16) & 63;
  constexpr int axis[6] = {2, 1, 0, 1, 0, 2};
  constexpr int face[6] = {-1, -1, 1, 1, -1, 1};
  double* fe = a.Fe + 60 * int64_t(it.x);
  const int stride = a.node_major ? 1 : 20;
  const int base = a.node_major ? 3 * l : l;
  double v[3] = {fe[base], fe[base + stride], fe[base + 2 * stride]};
  for (int s = 0; s < 6; ++s) {
    if (!((tm >> s) & 1)) continue;
    const int ax = axis[s];
    double w = ax == 0 ? v[0] : (ax == 1 ? v[1] : v[2]);
    const double p = a.pin[ax];
    for (int qp = 0; qp < 4; ++qp) w = w + a.tab[(s * 4 + qp) * 20 + l] * p;
    if (ax == 0) v[0] = w;
    else if (ax == 1) v[1] = w;
    else v[2] = w;
  }
  if (wm) {
    double ul[3] = {0.0, 0.0, 0.0}, ue[3] = {0.0, 0.0, 0.0};
    if (!a.undist) {
#pragma unroll
      for (int d = 0; d < 3; ++d) ul[d] = a.uloc[3 * q + d];
      if (a.validation) {
#pragma unroll
        for (int d = 0; d < 3; ++d) ue[d] = a.uex[3 * q + d];
      }
    }
    for (int s = 0; s < 6; ++s) {
      if (!((wm >> s) & 1)) continue;
      const int ax = axis[s];
      double uvw[3] = {0.0, 0.0, 0.0};
      if (a.shear[ax]) {
        const double val = a.rate[ax] * (face[s] < 0 ? a.lo[ax] : a.hi[ax]);
        const int c = a.sdir[ax];
        uvw[0] = c == 0 ? val : uvw[0];
        uvw[1] = c == 1 ? val : uvw[1];
        uvw[2] = c == 2 ? val : uvw[2];
      }
      if (!a.undist) {
#pragma unroll
        for (int d = 0; d < 3; ++d) uvw[d] = (a.validation ? ue[d] : uvw[d]) - ul[d];
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = v[d] + a.penalty * uvw[d];
    }
  }
  fe[base] = v[0];
  fe[base + stride] = v[1];
  fe[base + 2 * stride] = v[2];
}

unsigned rhs_blocks(int64_t m) { return unsigned((m + kRhsBlock - 1) / kRhsBlock); }

int check_opts(ggem_ctx* ctx, const ggem_rhs_opts* o) {
  if (!o) return fail(ctx, GGEM_ERR_INVALID_ARG, "null right-hand side options");
  if (o->option != GGEM_RHS_DISTURBED && o->option != GGEM_RHS_UNDISTURBED)
    return fail(ctx, GGEM_ERR_INVALID_ARG, "option must be GGEM_RHS_DISTURBED or GGEM_RHS_UNDISTURBED");
  for (int a = 0; a < 3; ++a)
    if (o->shear[a] && (o->shear_direction[a] < 0 || o->shear_direction[a] > 2))
      return fail(ctx, GGEM_ERR_INVALID_ARG, "shear_direction must be 0, 1 or 2");
  if (!std::isfinite(o->penalty)) return fail(ctx, GGEM_ERR_INVALID_ARG, "the penalty must be finite");
  return GGEM_OK;
}

int ensure_items(ggem_ctx* ctx, const ggem_rhs_opts* o) {
  const int key = rhs_key(o);
  if (ctx->rhs_key == key) return GGEM_OK;
  std::vector<int2> items;
  build_items(ctx->grid.n, ctx->grid.periodic, o, items);
  int rc = ensure(ctx, &ctx->rhs_items, &ctx->rhs_items_cap, std::max<size_t>(items.size(), 1));
  if (rc) return rc;
  if (!items.empty() && (rc = h2d(ctx, ctx->rhs_items, items.data(), sizeof(int2) * items.size()))) return rc;
  if ((rc = sync(ctx))) return rc;
  ctx->rhs_m = int64_t(items.size());
  ctx->rhs_key = key;
  return GGEM_OK;
}

// Element vectors of the options into the pool's Fe ([e][d][l], or
// [e][l][d] when node_major).
int rhs_fill(ggem_ctx* ctx, int32_t nqp_1d, const double* IF, const ggem_rhs_opts* o, bool node_major) {
  int rc = check_opts(ctx, o);
  if (rc) return rc;
  if ((rc = single_only(ctx, "the right-hand side assembly"))) return rc;
  const bool undist = o->option == GGEM_RHS_UNDISTURBED;
  if (!undist) {
    if ((rc = check_ready(ctx))) return rc;
    if (nqp_1d != 2 && nqp_1d != 3) return fail(ctx, GGEM_ERR_INVALID_ARG, "nqp_1d must be 2 or 3");
    if (!IF) return fail(ctx, GGEM_ERR_INVALID_ARG, "null IF table");
    if ((rc = ensure_spread(ctx, nqp_1d))) return rc;
  }
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  const int64_t ne = ctx->grid.n_elem;
  Pools& P = pools(ctx);
  if ((rc = ensure(ctx, &P.fe, &P.fe_cap, size_t(ne) * 60))) return rc;
  hipError_t e;
  if (!undist) {
    const int nqp = nqp_1d * nqp_1d * nqp_1d;
    if ((rc = ensure(ctx, &P.ifd, &P.ifd_cap, size_t(20 * nqp)))) return rc;
    if ((rc = h2d(ctx, P.ifd, IF, sizeof(double) * 20 * nqp))) return rc;
    e = launch_spread(ctx, ctx->p.precision, nqp_1d, 20, P.ifd, nullptr, P.fe, node_major);
    if (e != hipSuccess) return hip_fail(ctx, e, "spread");
  } else if ((e = launch_fill(ctx, ne * 60, 0.0, P.fe)) != hipSuccess) {
    return hip_fail(ctx, e, "fill");
  }
  if ((rc = ensure_items(ctx, o))) return rc;
  const int64_t m = ctx->rhs_m;
  if (m == 0) return GGEM_OK;
  RhsArgs a{};
  a.m = m;
  a.items = ctx->rhs_items;
  a.undist = undist ? 1 : 0;
  a.validation = (!undist && o->validation) ? 1 : 0;
  a.node_major = node_major ? 1 : 0;
  for (int d = 0; d < 3; ++d) {
    a.shear[d] = o->shear[d] ? 1 : 0;
    a.sdir[d] = o->shear_direction[d];
    a.rate[d] = o->shear_rate[d];
    a.pin[d] = o->inlet_pressure[d];
    a.lo[d] = ctx->p.box_min[d];
    a.hi[d] = ctx->p.box_max[d];
  }
  a.penalty = o->penalty;
  a.Fe = P.fe;
  if (undist && (o->inlet[0] || o->inlet[1] || o->inlet[2])) {
    if (!ctx->rhs_tab) GGEM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->rhs_tab), sizeof(double) * 480));
    std::vector<double> tab(480);
    traction_table(ctx->grid.h, tab.data());
    if ((rc = h2d(ctx, ctx->rhs_tab, tab.data(), sizeof(double) * 480))) return rc;
    a.tab = ctx->rhs_tab;
  }
  if (!undist) {
    if ((rc = ensure(ctx, &ctx->rhs_buf, &ctx->rhs_buf_cap, size_t(9 * m)))) return rc;
    if ((rc = ensure(ctx, &ctx->rhs_el, &ctx->rhs_el_cap, size_t(m)))) return rc;
    double* pts = ctx->rhs_buf;
    double* uloc = pts + 3 * m;
    double* uex = pts + 6 * m;
    {
      ProfScope ps(ctx, "rhs_item_points");
      hipLaunchKernelGGL(k_item_points, dim3(rhs_blocks(m)), dim3(kRhsBlock), 0, ctx->stream, ctx->grid, m,
                         ctx->rhs_items, pts, ctx->rhs_el);
      GGEM_HIP(ctx, hipGetLastError());
    }
    if ((e = launch_pair_points(ctx, ctx->p.precision, m, pts, ctx->rhs_el, uloc)) != hipSuccess)
      return hip_fail(ctx, e, "pair_points");
    if (a.validation) {
      ProfScope ps(ctx, "rhs_unbounded");
      hipLaunchKernelGGL(k_unbounded, dim3(rhs_blocks(m)), dim3(kRhsBlock), 0, ctx->stream, ctx->gc, ctx->n,
                         ctx->x_in, ctx->f_in, m, pts, uex);
      GGEM_HIP(ctx, hipGetLastError());
    }
    a.uloc = uloc;
    a.uex = uex;
  }
  ProfScope ps(ctx, "rhs_penalty");
  hipLaunchKernelGGL(k_rhs_items, dim3(rhs_blocks(m)), dim3(kRhsBlock), 0, ctx->stream, a);
  GGEM_HIP(ctx, hipGetLastError());
  return GGEM_OK;
}

}  // namespace
}  // namespace ggem

using namespace ggem;

extern "C" {

int ggem_util_rhs_wall_items(const int32_t* n_elem, const int32_t* periodic, const ggem_rhs_opts* o, int64_t* m,
                             int32_t* items) {
  if (!n_elem || !periodic || !m) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null argument");
  for (int d = 0; d < 3; ++d)
    if (n_elem[d] < 1) return fail(nullptr, GGEM_ERR_INVALID_ARG, "n_elem must be positive");
  int rc = check_opts(nullptr, o);
  if (rc) return rc;
  std::vector<int2> it;
  build_items(n_elem, periodic, o, it);
  *m = int64_t(it.size());
  if (items)
    for (size_t q = 0; q < it.size(); ++q) {
      items[2 * q] = it[q].x;
      items[2 * q + 1] = it[q].y;
    }
  return GGEM_OK;
}

int ggem_util_hex20_traction_table(const double* h, double* tab) {
  if (!h || !tab) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null argument");
  traction_table(h, tab);
  return GGEM_OK;
}

int ggem_assemble_elem_rhs(ggem_ctx* ctx, int32_t nqp_1d, const double* IF, const ggem_rhs_opts* o, double* Fe) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!Fe) return fail(ctx, GGEM_ERR_INVALID_ARG, "null output");
  int rc = rhs_fill(ctx, nqp_1d, IF, o, false);
  if (rc) return rc;
  return d2h(ctx, Fe, pools(ctx).fe, sizeof(double) * 60 * size_t(ctx->grid.n_elem));
}

int ggem_assemble_rhs(ggem_ctx* ctx, int32_t nqp_1d, const double* IF, const ggem_rhs_opts* o, int32_t condense,
                      double* F_node) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (ctx->fe_nodes < 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "no FE mesh (call ggem_set_fe_mesh)");
  int rc = rhs_fill(ctx, nqp_1d, IF, o, true);
  if (rc) return rc;
  const hipError_t e = launch_gather_nodes(ctx, pools(ctx).fe, ctx->fe_u, condense != 0);
  if (e != hipSuccess) return hip_fail(ctx, e, "gather_nodes");
  if (!F_node) return GGEM_OK;  // stays in the context for the caller's solve
  return d2h(ctx, F_node, ctx->fe_u, sizeof(double) * 3 * size_t(ctx->fe_nodes));
}

}  // extern "C"
