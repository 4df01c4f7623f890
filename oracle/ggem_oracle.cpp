// ggem_oracle.cpp -- FP64 CPU restatement of the reference's GGEM particle path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
// CPU baseline timed by bench.py.  It is never linked into the product.
//
// Arithmetic follows the reference statement by statement (same operation
// order, libm erf/erfc/exp/sqrt, no FMA contraction: built with
// -ffp-contract=off).  The only departures, all deliberate:
//  * particles are addressed by index (the reference copies the whole
//    particle-pointer vector per neighbour, point_mesh.h:241-243);
//  * element ownership of a particle uses the canonical uniform-grid rule
//    documented at locate() below instead of libMesh's PointLocatorTree;
//  * loops over beads / elements / points run on OpenMP threads; every output
//    element is produced by one thread in the reference's order, so results
//    do not depend on the thread count;
//  * Chebyshev dots and norms are plain sequential sums (PETSc uses BLAS
//    ddot/dnrm2; the difference is at the ulp level).
#include "ggem_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

const double PI = 3.14159265358979323846;  // libMesh::pi
const double SQRT_PI = std::sqrt(PI);       // ggem_system.h:110
const double PI_23 = std::pow(PI, 3. / 2.); // ggem_system.h:113
const double R_EPS = 1E-6;                  // ggem_system.h:116

int g_threads = 0;

inline int nthreads() { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

// GGEMStokes::set_mu / set_alpha / set_ksi (ggem_stokes.C:39-114) and the
// search radii of Copss::create_domain_mesh (copss.C:986-992, :1074-1075).
int derive(const orc_params* p, orc_consts* c) {
  std::memset(c, 0, sizeof(*c));
  for (int d = 0; d < 3; ++d) {
    if (p->n_elem[d] <= 0 || !(p->box_max[d] > p->box_min[d])) return -1;
    c->L[d] = p->box_max[d] - p->box_min[d];
    c->h[d] = c->L[d] / double(p->n_elem[d]);
  }
  if (!(p->mu > 0) || !(p->alpha > 0)) return -1;
  c->one_eight_pi_mu = 1. / (8. * PI * p->mu);
  c->alpha = p->alpha;
  c->alpha2 = p->alpha * p->alpha;
  const double alpha3 = c->alpha2 * p->alpha;
  c->alpha3_pi_23 = alpha3 / PI_23;
  c->two_alpha_sqrt_pi = 2. * p->alpha / SQRT_PI;
  c->coeff_1 = c->one_eight_pi_mu * 4.0 * p->alpha / SQRT_PI;
  double ksi = p->ksi;
  if (!(ksi > 0)) {
    if (p->point_type == ORC_LAGRANGIAN_POINT) ksi = 1.0 / (p->ibm_beta * p->hmin);
    else ksi = SQRT_PI / (3. * p->br0);
  }
  c->ksi = ksi;
  c->ksi2 = ksi * ksi;
  c->two_ksi_sqrt_pi = 2. * ksi / SQRT_PI;
  c->coeff_2 = c->one_eight_pi_mu * 4.0 * (ksi - p->alpha) / SQRT_PI;
  double hmax = std::max(std::max(c->h[0], c->h[1]), c->h[2]);
  c->search_radius_p = 4. / p->alpha;
  c->search_radius_e = p->search_radius_e > 0 ? p->search_radius_e : 0.5 * hmax + 4. / p->alpha;
  return 0;
}

// PMPeriodicBoundary::point_vector (pm_periodic_boundary.C:350-374)
inline void point_vector(const orc_params* p, const orc_consts* c, const double* p0,
                         const double* p1, double* d) {
  for (int i = 0; i < 3; ++i) {
    d[i] = p1[i] - p0[i];
    if (p->periodic[i]) {
      const double xsize = c->L[i];
      if (d[i] <= -xsize / 2.0) d[i] += xsize;
      if (d[i] > xsize / 2.0) d[i] -= xsize;
    }
  }
}

// GGEMStokes::green_tensor_local_regularized (ggem_stokes.C:304-357)
inline void green_local(const orc_consts* c, const double* x, double* G) {
  const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (r < R_EPS) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) G[3 * i + j] = (i == j) ? c->coeff_2 : 0.0 * c->coeff_2;
    return;
  }
  const double r2 = r * r;
  const double kr = c->ksi * r, kr2 = kr * kr;
  const double ar = c->alpha * r, ar2 = ar * ar;
  const double c1 = (std::erf(kr) - std::erf(ar)) / r;
  const double c2 = c->two_ksi_sqrt_pi * std::exp(-kr2) - c->two_alpha_sqrt_pi * std::exp(-ar2);
  for (int i = 0; i < 3; ++i) {
    for (int j = i; j < 3; ++j) {
      const double dij = (i == j) ? 1.0 : 0.0;
      const double xixj_r2 = x[i] * x[j] / r2;
      const double tmp1 = (dij + xixj_r2) * c1;
      const double tmp2 = (dij - xixj_r2) * c2;
      G[3 * i + j] = tmp1 + tmp2;
    }
    for (int j = i - 1; j >= 0; --j) G[3 * i + j] = G[3 * j + i];
  }
  for (int k = 0; k < 9; ++k) G[k] *= c->one_eight_pi_mu;
}

// GGEMStokes::green_tensor_unbounded_smoothed (ggem_stokes.C:201-251)
inline void green_unbounded(const orc_consts* c, const double* x, double a, double* G) {
  const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (r < R_EPS) {
    const double c0 = c->one_eight_pi_mu * 4.0 * a / SQRT_PI;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) G[3 * i + j] = (i == j) ? c0 : 0.0;
    return;
  }
  const double r2 = r * r, a2 = a * a;
  const double c1 = std::erf(a * r) / r;
  const double c2 = 2. * a / SQRT_PI * std::exp(-a2 * r2);
  for (int i = 0; i < 3; ++i) {
    for (int j = i; j < 3; ++j) {
      const double dij = (i == j) ? 1.0 : 0.0;
      const double xixj_r2 = x[i] * x[j] / r2;
      const double tmp1 = (dij + xixj_r2) * c1;
      const double tmp2 = (dij - xixj_r2) * c2;
      G[3 * i + j] = tmp1 + tmp2;
    }
    for (int j = i - 1; j >= 0; --j) G[3 * i + j] = G[3 * j + i];
  }
  for (int k = 0; k < 9; ++k) G[k] *= c->one_eight_pi_mu;
}

// GGEMStokes::smoothed_force_exp (ggem_stokes.C:117-128)
inline double smoothed_force_exp(const orc_consts* c, const double* r) {
  const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  const double a2r2 = c->alpha2 * r2;
  return c->alpha3_pi_23 * std::exp(-a2r2) * (2.5 - a2r2);
}

// u(i) += GT(i,j) * f(j), i outer, j inner (ggem_stokes.C:418-420, :469-471)
inline void gemv_acc(const double* G, const double* f, double* u) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) u[i] += G[3 * i + j] * f[j];
}

}  // namespace

struct orc_mesh {
  orc_params p;
  orc_consts c;
  int64_t n_elem = 0;
  // elem-elem CSR (point_mesh.C:293-443)
  std::vector<int64_t> ee_off;
  std::vector<int32_t> ee_ids;
  bool ee_built = false;
  // particles
  int64_t n = 0;
  std::vector<double> x, f;        // interleaved 3n
  std::vector<int32_t> elem;       // PointParticle::_elem_id (-1 = unset)
  // elem-point containing and neighbour lists (point_mesh.C:573-799)
  std::vector<int64_t> ec_off, en_off;
  std::vector<int32_t> ec_ids, en_ids;
  // particle lists (point_mesh.C:803-848), kept per particle for the
  // point_particle.h:189-194 quirk
  std::vector<std::vector<int32_t>> plist;
};

namespace {

inline int64_t eid(const orc_mesh* m, int64_t i, int64_t j, int64_t k) {
  return i + m->p.n_elem[0] * (j + int64_t(m->p.n_elem[1]) * k);
}

// libMesh Elem::centroid(): average of the 8 vertices of the HEX20 element
// built by MeshTools::Generation::build_cube, whose node coordinates are
// (ii / (2 n)) * (max - min) + min.  Vertex order 0..7 of a hex.
void centroid(const orc_mesh* m, int64_t e, double* cp) {
  const int nx = m->p.n_elem[0], ny = m->p.n_elem[1];
  const int64_t i = e % nx, j = (e / nx) % ny, k = e / (int64_t(nx) * ny);
  double lo[3], hi[3];
  const int64_t ijk[3] = {i, j, k};
  for (int d = 0; d < 3; ++d) {
    const double n2 = 2.0 * double(m->p.n_elem[d]);
    lo[d] = (double(2 * ijk[d]) / n2) * (m->p.box_max[d] - m->p.box_min[d]) + m->p.box_min[d];
    hi[d] = (double(2 * ijk[d] + 2) / n2) * (m->p.box_max[d] - m->p.box_min[d]) + m->p.box_min[d];
  }
  // vertices: (lo,lo,lo) (hi,lo,lo) (hi,hi,lo) (lo,hi,lo) then z = hi
  const int vx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
  const int vy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
  const int vz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  double s[3] = {0, 0, 0};
  for (int v = 0; v < 8; ++v) {
    s[0] += vx[v] ? hi[0] : lo[0];
    s[1] += vy[v] ? hi[1] : lo[1];
    s[2] += vz[v] ? hi[2] : lo[2];
  }
  for (int d = 0; d < 3; ++d) cp[d] = s[d] / 8.0;
}

// build_elem_elem_neighbor_list, literal form: every other element whose
// centroid lies strictly inside R_e (minimum image), ascending id
// (point_mesh.C:333-355).  O(n_e^2): small grids only.
void ee_literal(orc_mesh* m) {
  const int64_t ne = m->n_elem;
  std::vector<double> cen(3 * ne);
  for (int64_t e = 0; e < ne; ++e) centroid(m, e, &cen[3 * e]);
  std::vector<std::vector<int32_t>> lists(ne);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads())
  for (int64_t e = 0; e < ne; ++e) {
    for (int64_t i = 0; i < ne; ++i) {
      if (i == e) continue;
      double d[3];
      point_vector(&m->p, &m->c, &cen[3 * e], &cen[3 * i], d);
      const double nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (nrm < m->c.search_radius_e) lists[e].push_back(int32_t(i));
    }
  }
  m->ee_off.assign(ne + 1, 0);
  for (int64_t e = 0; e < ne; ++e) m->ee_off[e + 1] = m->ee_off[e] + lists[e].size();
  m->ee_ids.resize(m->ee_off[ne]);
  for (int64_t e = 0; e < ne; ++e)
    std::copy(lists[e].begin(), lists[e].end(), m->ee_ids.begin() + m->ee_off[e]);
}

// The same list by a window search: candidate cells within ceil(R_e/h) cells
// per axis (distinct minimum-image cells on periodic axes), tested with the
// same vertex-average centroids and the same strict comparison, then sorted.
void ee_window(orc_mesh* m) {
  const int64_t ne = m->n_elem;
  int w[3];
  for (int d = 0; d < 3; ++d) w[d] = int(std::ceil(m->c.search_radius_e / m->c.h[d])) + 1;
  std::vector<std::vector<int32_t>> lists(ne);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads())
  for (int64_t e = 0; e < ne; ++e) {
    const int nx = m->p.n_elem[0], ny = m->p.n_elem[1];
    const int64_t ijk[3] = {e % nx, (e / nx) % ny, e / (int64_t(nx) * ny)};
    double c0[3];
    centroid(m, e, c0);
    std::vector<int64_t> cand[3];
    for (int d = 0; d < 3; ++d) {
      const int n = m->p.n_elem[d];
      for (int o = -w[d]; o <= w[d]; ++o) {
        int64_t q = ijk[d] + o;
        if (m->p.periodic[d]) {
          q = ((q % n) + n) % n;
        } else if (q < 0 || q >= n) {
          continue;
        }
        cand[d].push_back(q);
      }
      std::sort(cand[d].begin(), cand[d].end());
      cand[d].erase(std::unique(cand[d].begin(), cand[d].end()), cand[d].end());
    }
    auto& L = lists[e];
    for (int64_t kk : cand[2])
      for (int64_t jj : cand[1])
        for (int64_t ii : cand[0]) {
          const int64_t o = eid(m, ii, jj, kk);
          if (o == e) continue;
          double c1[3], d[3];
          centroid(m, o, c1);
          point_vector(&m->p, &m->c, c0, c1, d);
          const double nrm = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          if (nrm < m->c.search_radius_e) L.push_back(int32_t(o));
        }
    std::sort(L.begin(), L.end());
  }
  m->ee_off.assign(ne + 1, 0);
  for (int64_t e = 0; e < ne; ++e) m->ee_off[e + 1] = m->ee_off[e] + lists[e].size();
  m->ee_ids.resize(m->ee_off[ne]);
  for (int64_t e = 0; e < ne; ++e)
    std::copy(lists[e].begin(), lists[e].end(), m->ee_ids.begin() + m->ee_off[e]);
}

// Canonical element ownership (replaces libMesh's point locator,
// point_mesh.C:594-617): per axis c = floor((x - min) / h), clamped to
// [0, n-1]; a coordinate outside [min, max] is out of the domain (error, as
// the reference's failed locator at point_mesh.C:602-612).
int32_t locate(const orc_mesh* m, const double* x) {
  int64_t c[3];
  for (int d = 0; d < 3; ++d) {
    if (!(x[d] >= m->p.box_min[d]) || !(x[d] <= m->p.box_max[d])) return -1;
    int64_t q = int64_t(std::floor((x[d] - m->p.box_min[d]) / m->c.h[d]));
    if (q < 0) q = 0;
    if (q > m->p.n_elem[d] - 1) q = m->p.n_elem[d] - 1;
    c[d] = q;
  }
  return int32_t(eid(m, c[0], c[1], c[2]));
}

// Elem::contains_point with libMesh's TOLERANCE (1e-6) on the reference
// element of an axis-aligned hex: |2 (x - centre) / h| <= 1 + 1e-6 per axis
// (point_mesh.C:589-596 keeps the old element when this holds).
bool contains(const orc_mesh* m, int32_t e, const double* x) {
  const int nx = m->p.n_elem[0], ny = m->p.n_elem[1];
  const int64_t ijk[3] = {e % nx, (e / nx) % ny, e / (int64_t(nx) * ny)};
  for (int d = 0; d < 3; ++d) {
    const double cen = m->p.box_min[d] + (double(ijk[d]) + 0.5) * m->c.h[d];
    const double xi = 2.0 * (x[d] - cen) / m->c.h[d];
    if (!(std::fabs(xi) <= 1.0 + 1e-6)) return false;
  }
  return true;
}

}  // namespace

extern "C" {

int orc_consts_get(const orc_params* p, orc_consts* c) { return derive(p, c); }

void orc_point_vector(const orc_params* p, const double* p0, const double* p1, double* d) {
  orc_consts c;
  derive(p, &c);
  point_vector(p, &c, p0, p1, d);
}

// PMPeriodicBoundary::correct_position (pm_periodic_boundary.C:305-328)
void orc_correct_position(const orc_params* p, double* x) {
  for (int i = 0; i < 3; ++i) {
    if (p->periodic[i]) {
      const double L = p->box_max[i] - p->box_min[i];
      if (x[i] < p->box_min[i]) x[i] += L;
      if (x[i] >= p->box_max[i]) x[i] -= L;
    }
  }
}

// FixPoint::check_walls (fix/fix_point.C:27-90) for one point: slit walls
// (wall_type 0: a periodic axis wraps once each way and counts the image,
// a wall axis reflects) or a spherical cavity of radius R about the origin
// (wall_type 1: a point outside is reflected to pt_unit 2R - pt).
void orc_check_walls(const orc_params* p, int wall_type, double radius, double* x, int32_t* count) {
  if (wall_type == 0) {
    for (int i = 0; i < 3; ++i) {
      const double L = p->box_max[i] - p->box_min[i];
      if (p->periodic[i]) {
        if (x[i] < p->box_min[i]) {
          x[i] += L;
          count[i] -= 1;
        }
        if (x[i] >= p->box_max[i]) {
          x[i] -= L;
          count[i] += 1;
        }
      } else {
        if (x[i] < p->box_min[i]) x[i] = 2. * p->box_min[i] - x[i];
        if (x[i] >= p->box_max[i]) x[i] = 2. * p->box_max[i] - x[i];
      }
    }
  } else {
    const double nrm = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (nrm > radius) {
      double u[3];
      for (int i = 0; i < 3; ++i) u[i] = x[i] / nrm;
      for (int i = 0; i < 3; ++i) x[i] = u[i] * (2. * radius) - x[i];
    }
  }
}

void orc_green_local_regularized(const orc_params* p, const double* x, double* G9) {
  orc_consts c;
  derive(p, &c);
  green_local(&c, x, G9);
}

void orc_green_unbounded_smoothed(const orc_params* p, const double* x, double a, double* G9) {
  orc_consts c;
  derive(p, &c);
  green_unbounded(&c, x, a, G9);
}

double orc_smoothed_force_exp(const orc_params* p, const double* r) {
  orc_consts c;
  derive(p, &c);
  return smoothed_force_exp(&c, r);
}

orc_mesh* orc_mesh_create(const orc_params* p) {
  orc_mesh* m = new orc_mesh();
  m->p = *p;
  if (derive(p, &m->c) != 0) {
    delete m;
    return nullptr;
  }
  m->n_elem = int64_t(p->n_elem[0]) * p->n_elem[1] * p->n_elem[2];
  return m;
}

void orc_mesh_destroy(orc_mesh* m) { delete m; }

int orc_mesh_build_elem_elem(orc_mesh* m, int method) {
  if (method == 0) ee_literal(m);
  else ee_window(m);
  m->ee_built = true;
  return 0;
}

int64_t orc_mesh_elem_elem_len(orc_mesh* m, int64_t e) { return m->ee_off[e + 1] - m->ee_off[e]; }
const int32_t* orc_mesh_elem_elem(orc_mesh* m, int64_t e) { return m->ee_ids.data() + m->ee_off[e]; }

// PointMesh::reinit (point_mesh.C:245-289) with a neighbour-list update:
// build_elem_point_containing_list -> build_elem_point_neighbor_list ->
// build_point_point_neighbor_list.  Forces are set from f (the reference
// zeroes them at :249-251 and the fixes refill them).
// PointMesh::reinit(false) (point_mesh.C:245-289 without the list
// update): new positions and forces (zero when null), the lists and element
// ids kept; every r_ij is formed from the new positions when it is used
// (reinit_neighbor_vector, :220-241, point_vector).
int orc_mesh_update_positions(orc_mesh* m, const double* xyz, const double* f) {
  if (m->n == 0 || m->en_off.empty()) return -1;
  m->x.assign(xyz, xyz + 3 * m->n);
  if (f) m->f.assign(f, f + 3 * m->n);
  else m->f.assign(3 * m->n, 0.0);
  return 0;
}

int orc_mesh_reinit(orc_mesh* m, int64_t n, const double* xyz, const double* f) {
  if (!m->ee_built) orc_mesh_build_elem_elem(m, m->n_elem <= 20000 ? 0 : 1);
  if (m->n != n) {
    m->n = n;
    m->elem.assign(n, -1);
    m->plist.assign(n, std::vector<int32_t>());
  }
  m->x.assign(xyz, xyz + 3 * n);
  if (f) m->f.assign(f, f + 3 * n);
  else m->f.assign(3 * n, 0.0);
  // build_elem_point_containing_list (point_mesh.C:573-648)
  for (int64_t i = 0; i < n; ++i) {
    bool use_locator = m->elem[i] == -1 || !contains(m, m->elem[i], &m->x[3 * i]);
    if (use_locator) {
      const int32_t e = locate(m, &m->x[3 * i]);
      if (e < 0) return -2;
      m->elem[i] = e;
    }
  }
  const int64_t ne = m->n_elem;
  m->ec_off.assign(ne + 1, 0);
  for (int64_t i = 0; i < n; ++i) m->ec_off[m->elem[i] + 1]++;
  for (int64_t e = 0; e < ne; ++e) m->ec_off[e + 1] += m->ec_off[e];
  m->ec_ids.resize(n);
  {
    std::vector<int64_t> cur(m->ec_off.begin(), m->ec_off.end() - 1);
    for (int64_t i = 0; i < n; ++i) m->ec_ids[cur[m->elem[i]]++] = int32_t(i);  // ascending id
  }
  // build_elem_point_neighbor_list (point_mesh.C:652-799): own element first,
  // then the elem-elem list in its (ascending) order
  m->en_off.assign(ne + 1, 0);
  for (int64_t e = 0; e < ne; ++e) {
    int64_t cnt = m->ec_off[e + 1] - m->ec_off[e];
    for (int64_t q = m->ee_off[e]; q < m->ee_off[e + 1]; ++q) {
      const int32_t o = m->ee_ids[q];
      cnt += m->ec_off[o + 1] - m->ec_off[o];
    }
    m->en_off[e + 1] = m->en_off[e] + cnt;
  }
  m->en_ids.resize(m->en_off[ne]);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads())
  for (int64_t e = 0; e < ne; ++e) {
    int64_t w = m->en_off[e];
    for (int64_t q = m->ec_off[e]; q < m->ec_off[e + 1]; ++q) m->en_ids[w++] = m->ec_ids[q];
    for (int64_t q = m->ee_off[e]; q < m->ee_off[e + 1]; ++q) {
      const int32_t o = m->ee_ids[q];
      for (int64_t s = m->ec_off[o]; s < m->ec_off[o + 1]; ++s) m->en_ids[w++] = m->ec_ids[s];
    }
  }
  // build_point_point_neighbor_list (point_mesh.C:803-848) +
  // PointParticle::set_neighbor_list (point_particle.h:189-194)
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads())
  for (int64_t i = 0; i < n; ++i) {
    const int32_t e = m->elem[i];
    std::vector<int32_t> L(m->en_ids.begin() + m->en_off[e], m->en_ids.begin() + m->en_off[e + 1]);
    L.erase(std::remove(L.begin(), L.end(), int32_t(i)), L.end());
    if (L.size() > 0 || !m->p.keep_stale_lists) m->plist[i].swap(L);
  }
  return 0;
}

int orc_mesh_set_forces(orc_mesh* m, const double* f) {
  m->f.assign(f, f + 3 * m->n);
  return 0;
}

int32_t orc_mesh_elem_of(orc_mesh* m, int64_t i) { return m->elem[i]; }
int64_t orc_mesh_elem_point_len(orc_mesh* m, int64_t e) { return m->en_off[e + 1] - m->en_off[e]; }
const int32_t* orc_mesh_elem_point(orc_mesh* m, int64_t e) { return m->en_ids.data() + m->en_off[e]; }
int64_t orc_mesh_point_len(orc_mesh* m, int64_t i) { return int64_t(m->plist[i].size()); }
const int32_t* orc_mesh_point(orc_mesh* m, int64_t i) { return m->plist[i].data(); }

int64_t orc_mesh_total_pairs(orc_mesh* m) {
  int64_t s = 0;
  for (auto& L : m->plist) s += int64_t(L.size());
  return s;
}

int orc_mesh_export_csr(orc_mesh* m, int64_t* offsets, int32_t* ids) {
  offsets[0] = 0;
  for (int64_t i = 0; i < m->n; ++i) {
    const auto& L = m->plist[i];
    if (ids) std::copy(L.begin(), L.end(), ids + offsets[i]);
    offsets[i + 1] = offsets[i] + int64_t(L.size());
  }
  return 0;
}

// GGEMStokes::local_solution_bead (ggem_stokes.C:427-503); the neighbour
// vector r_ij = point_vector(x_i, x_j) is PointMesh::reinit_neighbor_vector
// (point_mesh.C:220-241), recomputed from the current positions.
int orc_local_velocity_beads(orc_mesh* m, double* u) {
  const int64_t n = m->n;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads())
  for (int64_t i = 0; i < n; ++i) {
    double ui[3] = {0, 0, 0};
    const double* xi = &m->x[3 * i];
    for (int32_t j : m->plist[i]) {
      double d[3], G[9];
      point_vector(&m->p, &m->c, xi, &m->x[3 * j], d);
      green_local(&m->c, d, G);
      gemv_acc(G, &m->f[3 * j], ui);
    }
    if (m->p.point_type == ORC_LAGRANGIAN_POINT) {
      const double z[3] = {0, 0, 0};
      double G[9];
      green_local(&m->c, z, G);
      gemv_acc(G, &m->f[3 * i], ui);
    }
    u[3 * i] = ui[0];
    u[3 * i + 1] = ui[1];
    u[3 * i + 2] = ui[2];
  }
  return 0;
}

// GGEMStokes::global_self_exclusion (ggem_stokes.C:506-523)
int orc_global_self_exclusion(orc_mesh* m, double* uex) {
  for (int64_t k = 0; k < 3 * m->n; ++k) uex[k] = m->c.coeff_1 * m->f[k];
  return 0;
}

// PMSystemStokes::compute_point_velocity("disturbed") (pm_system_stokes.C:441-457):
// POLYMER_BEAD: U = U_glob + U_loc - U_ex + f; LAGRANGIAN_POINT: U_glob + U_loc;
// POINT_PARTICLE falls through both branches and keeps U = 0.
int orc_combine_velocity(orc_mesh* m, const double* u_glob, double* U) {
  const int64_t n = m->n;
  std::vector<double> ul(3 * n);
  orc_local_velocity_beads(m, ul.data());
  for (int64_t k = 0; k < 3 * n; ++k) {
    const double ug = u_glob ? u_glob[k] : 0.0;
    if (m->p.point_type == ORC_POLYMER_BEAD) {
      const double uex = m->c.coeff_1 * m->f[k];
      U[k] = ug + ul[k] - uex + m->f[k];
    } else if (m->p.point_type == ORC_LAGRANGIAN_POINT) {
      U[k] = ug + ul[k];
    } else {
      U[k] = 0.0;
    }
  }
  return 0;
}

int32_t orc_locate_point(orc_mesh* m, const double* x) { return locate(m, x); }

// GGEMStokes::local_solution_field (ggem_stokes.C:360-424): sum over the
// element's particle neighbour list, own element included, no exclusion.
int orc_local_velocity_points(orc_mesh* m, int64_t np, const double* pts, const int32_t* elem,
                              double* u) {
  for (int64_t q = 0; q < np; ++q)
    if (elem[q] < 0 || elem[q] >= m->n_elem) return -3;  // ggem_stokes.C:369-380
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads())
  for (int64_t q = 0; q < np; ++q) {
    double uq[3] = {0, 0, 0};
    const int32_t e = elem[q];
    for (int64_t s = m->en_off[e]; s < m->en_off[e + 1]; ++s) {
      const int32_t j = m->en_ids[s];
      double d[3], G[9];
      point_vector(&m->p, &m->c, &pts[3 * q], &m->x[3 * j], d);
      green_local(&m->c, d, G);
      gemv_acc(G, &m->f[3 * j], uq);
    }
    u[3 * q] = uq[0];
    u[3 * q + 1] = uq[1];
    u[3 * q + 2] = uq[2];
  }
  return 0;
}

// AnalyticalSolutionStokes::exact_solution_infinite_domain
// (analytical_solution_stokes.C:66-102): sum over all particles of
// G_xi(pt0 - x_i) f_i, no periodic image.
int orc_exact_solution_infinite(orc_mesh* m, int64_t np, const double* pts, double* u) {
  for (int64_t q = 0; q < np; ++q) {
    double uq[3] = {0, 0, 0};
    for (int64_t i = 0; i < m->n; ++i) {
      double d[3], G[9];
      for (int k = 0; k < 3; ++k) d[k] = pts[3 * q + k] - m->x[3 * i + k];
      green_unbounded(&m->c, d, m->c.ksi, G);
      gemv_acc(G, &m->f[3 * i], uq);
    }
    for (int k = 0; k < 3; ++k) u[3 * q + k] = uq[k];
  }
  return 0;
}

// Gauss-Legendre quadrature points of QGauss(3, order) on a uniform
// axis-aligned hex: x fastest (qp = a + n (b + n c)), 1D points ascending;
// physical point = centre + xi * h / 2.
int orc_qp_coords(orc_mesh* m, int nqp_1d, int64_t e, double* q) {
  double g[3];
  if (nqp_1d == 2) {
    g[0] = -std::sqrt(1. / 3.);
    g[1] = std::sqrt(1. / 3.);
  } else if (nqp_1d == 3) {
    g[0] = -std::sqrt(3. / 5.);
    g[1] = 0.0;
    g[2] = std::sqrt(3. / 5.);
  } else {
    return -1;
  }
  const int nx = m->p.n_elem[0], ny = m->p.n_elem[1];
  const int64_t ijk[3] = {e % nx, (e / nx) % ny, e / (int64_t(nx) * ny)};
  double cen[3], hh[3];
  for (int d = 0; d < 3; ++d) {
    cen[d] = m->p.box_min[d] + (double(ijk[d]) + 0.5) * m->c.h[d];
    hh[d] = 0.5 * m->c.h[d];
  }
  int qp = 0;
  for (int c = 0; c < nqp_1d; ++c)
    for (int b = 0; b < nqp_1d; ++b)
      for (int a = 0; a < nqp_1d; ++a) {
        q[3 * qp + 0] = cen[0] + g[a] * hh[0];
        q[3 * qp + 1] = cen[1] + g[b] * hh[1];
        q[3 * qp + 2] = cen[2] + g[c] * hh[2];
        ++qp;
      }
  return 0;
}

// AssembleStokes::compute_element_rhs, point-force branch
// (assemble_stokes.C:643-672) for every element of assemble_global_F
// (:473-516): rho(qp) = sum_p g(point_vector(q, x_p)) f_p, then
// Fe[j*n_u + l] += IF[l*nqp + qp] * rho_j(qp).  Elements with an empty list
// keep Fe = 0 (pf_flag false, :489-494).
int orc_spread(orc_mesh* m, int nqp_1d, int n_u, const double* IF, double* rho, double* Fe) {
  const int nqp = nqp_1d * nqp_1d * nqp_1d;
  if (nqp_1d != 2 && nqp_1d != 3) return -1;
  const int64_t ne = m->n_elem;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads())
  for (int64_t e = 0; e < ne; ++e) {
    double q[27 * 3];
    orc_qp_coords(m, nqp_1d, e, q);
    std::vector<double> fe(3 * n_u, 0.0);
    const int64_t b = m->en_off[e], t = m->en_off[e + 1];
    for (int qp = 0; qp < nqp; ++qp) {
      double rg[3] = {0, 0, 0};
      if (t > b) {
        for (int64_t s = b; s < t; ++s) {
          const int32_t pid = m->en_ids[s];
          double d[3];
          point_vector(&m->p, &m->c, &q[3 * qp], &m->x[3 * pid], d);
          const double g = smoothed_force_exp(&m->c, d);
          for (int j = 0; j < 3; ++j) rg[j] += g * m->f[3 * pid + j];
        }
        for (int j = 0; j < 3; ++j)
          for (int l = 0; l < n_u; ++l) fe[j * n_u + l] += IF[l * nqp + qp] * rg[j];
      }
      if (rho)
        for (int j = 0; j < 3; ++j) rho[(e * nqp + qp) * 3 + j] = rg[j];
    }
    if (Fe) std::copy(fe.begin(), fe.end(), Fe + e * 3 * n_u);
  }
  return 0;
}

// Chebyshev::chebyshev_coefficients with the "Gauss" nodes
// (chebyshev.C:44-107, :136-151)
void orc_chebyshev_coefficients(int N, double da, double db, double* coef) {
  const double RN = double(N);
  std::vector<double> x(N + 1), w(N + 1);
  for (int i = 0; i <= N; ++i) {
    const double j = double(i);
    x[i] = std::cos((2. * j + 1.) * PI / (2. * RN + 2.));
    w[i] = PI / (RN + 1.);
  }
  for (int k = 0; k <= N; ++k) {
    coef[k] = 0.0;
    for (int j = 0; j <= N; ++j) {
      const double tmpt = PI * double(2. * j + 1.) * double(k) / (2. * RN + 2.);
      const double fj = 1.0 / std::sqrt((x[j] - db) / da);
      coef[k] += w[j] * std::cos(tmpt) * fj;
    }
    double ck = 1.0;
    if (k == 0) ck = 2.0;
    coef[k] *= 2. / (PI * ck);
  }
}

// _MatMult_Stokes (brownian_system.C:44-140) with the host FEM part removed:
// forces <- f, then compute_point_velocity("disturbed") with U_glob = 0.
int orc_mobility_apply(orc_mesh* m, const double* f, double* u) {
  m->f.assign(f, f + 3 * m->n);
  return orc_combine_velocity(m, nullptr, u);
}

int orc_mobility_dense(orc_mesh* m, double* Md) {
  const int64_t n3 = 3 * m->n;
  std::vector<double> e(n3, 0.0), col(n3);
  for (int64_t c = 0; c < n3; ++c) {
    e[c] = 1.0;
    orc_mobility_apply(m, e.data(), col.data());
    for (int64_t r = 0; r < n3; ++r) Md[r * n3 + c] = col[r];
    e[c] = 0.0;
  }
  return 0;
}

namespace {
double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
double nrm(const std::vector<double>& a) { return std::sqrt(dot(a, a)); }
}  // namespace

// BrownianSystem::chebyshev_polynomial_approximation (brownian_system.C:967-1178).
// PETSc semantics: VecAXPBY(y, a, b, x): y = a x + b y; VecAXPY(y, a, x): y += a x.
int orc_cheb_brownian(orc_mesh* m, int N, double eig_min, double eig_max, double tol, int mode,
                      const double* dw_in, double* y_out, double* disp, orc_cheb_info* info) {
  const int64_t n3 = 3 * m->n;
  if (N < 1 || !(eig_max > eig_min)) return -1;
  const double d_eig = eig_max - eig_min;
  const double da = 2.0 / d_eig;
  const double db = -(eig_max + eig_min) / d_eig;
  std::vector<double> X(N + 1);
  orc_chebyshev_coefficients(N, da, db, X.data());
  std::vector<double> x0(dw_in, dw_in + n3), x1(n3), xmid(n3), dwm(n3), dw(n3, 0.0);
  int applies = 0;
  bool conv = false;
  double err1 = 0, err2 = 0;
  int iters = 1;
  orc_mobility_apply(m, x0.data(), x1.data());  // x1 = M x0
  ++applies;
  const double aux2 = dot(x1, x0);
  for (int64_t k = 0; k < n3; ++k) x1[k] = db * x0[k] + da * x1[k];
  for (int64_t k = 0; k < n3; ++k) dw[k] += X[0] * x0[k];
  double aux4 = nrm(dw);
  for (int64_t k = 0; k < n3; ++k) dw[k] += X[1] * x1[k];
  double aux3 = nrm(x1) * std::fabs(X[1]);
  if (mode == 0) {
    orc_mobility_apply(m, dw.data(), dwm.data());
    ++applies;
    const double aux1 = dot(dwm, dwm);
    if (aux2 != 0.0) err1 = std::sqrt(std::fabs(aux1 - aux2) / aux2);
    if (aux4 != 0.0) err2 = aux3 / aux4;
  }
  for (int i = 1; i < N; ++i) {
    if (mode == 0) aux4 = nrm(dw);
    orc_mobility_apply(m, x1.data(), xmid.data());
    ++applies;
    for (int64_t k = 0; k < n3; ++k) xmid[k] = 2. * db * x1[k] + 2. * da * xmid[k];
    for (int64_t k = 0; k < n3; ++k) xmid[k] += -1.0 * x0[k];
    for (int64_t k = 0; k < n3; ++k) dw[k] += X[i + 1] * xmid[k];
    iters = i + 1;
    if (mode == 0) {
      aux3 = nrm(xmid) * std::fabs(X[i + 1]);
      orc_mobility_apply(m, dw.data(), dwm.data());
      ++applies;
      const double aux1 = dot(dwm, dwm);
      if (aux2 != 0.0) err1 = std::sqrt(std::fabs(aux1 - aux2) / aux2);
      if (aux4 != 0.0) err2 = aux3 / aux4;
      if (err1 <= tol) {
        conv = true;
        break;
      }
      if (err2 <= 1E-9) break;
    }
    x0.swap(x1);
    x1.swap(xmid);
  }
  if (mode != 0) conv = true;
  std::copy(dw.begin(), dw.end(), y_out);
  if (disp) {  // hi_ewald: M y (copss.C:1686)
    orc_mobility_apply(m, dw.data(), disp);
    ++applies;
  }
  if (info) {
    info->converged = conv ? 1 : 0;
    info->iterations = iters;
    info->applies = applies;
    info->err1 = err1;
    info->err2 = err2;
  }
  return 0;
}

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) { return nthreads(); }

}  // extern "C"
