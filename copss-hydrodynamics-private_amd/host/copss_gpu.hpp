// copss_gpu.hpp -- C++ host classes over the copss_ggem C-ABI that keep the
// reference's GGEM host interface: the same class and method names, the
// same argument meaning, and errors raised where the reference calls
// libmesh_error() (here: std::runtime_error carrying ggem_last_error()).
//
//   PointMesh     <- PointMesh<3> (point_mesh.h): particles, reinit(bool&),
//                    element-point neighbour lists
//   GGEMStokes    <- GGEMStokes (ggem_stokes.h): local_solution_bead,
//                    local_solution_field, global_self_exclusion,
//                    smoothed force spreading (AssembleStokes::compute_element_rhs)
//   BrownianSystem <- BrownianSystem (brownian_system.h):
//                    chebyshev_polynomial_approximation, hi_ewald,
//                    compute_eigenvalues
//
// The reference computes one bead or one element per call; here a call
// computes every bead / element once on the GPU and later calls for other
// beads read the cached result until the particles or forces change.
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "copss_ggem.h"

namespace copss_amd {

using Point = std::array<double, 3>;

inline void check(int rc, const ggem_ctx* ctx) {
  if (rc != GGEM_OK) throw std::runtime_error(std::string("copss_ggem: ") + ggem_last_error(ctx));
}

class PointMesh {
 public:
  explicit PointMesh(const ggem_params& p) : params_(p) {
    ggem_ctx* c = nullptr;
    check(ggem_create(&p, &c), nullptr);
    ctx_ = c;
  }
  ~PointMesh() { ggem_destroy(ctx_); }
  PointMesh(const PointMesh&) = delete;
  PointMesh& operator=(const PointMesh&) = delete;

  ggem_ctx* ctx() const { return ctx_; }

  // Particle store (the reference's PointParticle objects): positions and
  // forces, interleaved [3*i + j].  They reach the device at the next
  // reinit, which decides whether the lists are rebuilt.
  void set_particles(const std::vector<double>& xyz, const std::vector<double>& f) {
    if (xyz.size() % 3 != 0 || (!f.empty() && f.size() != xyz.size()))
      throw std::runtime_error("copss_ggem: particle arrays must hold 3 values per particle");
    if (xyz.size() / 3 != n_) lists_ = false;  // a new particle set has no lists to keep
    xyz_ = xyz;
    n_ = xyz.size() / 3;
    force_ = f.empty() ? std::vector<double>(xyz.size(), 0.0) : f;
    pending_ = true;
    ++version_;
  }
  void set_forces(const std::vector<double>& f) {
    if (f.size() != 3 * n_) throw std::runtime_error("copss_ggem: force array size");
    force_ = f;
    if (!pending_) check(ggem_set_forces(ctx_, force_.data()), ctx_);
    ++version_;
  }

  // PointMesh<3>::reinit(bool& neighbor_list_update_flag) (point_mesh.C:245-289).
  // With the flag set the lists are rebuilt and the flag cleared, as the
  // reference does; without it the lists, element ids and element lists of
  // the last build are kept and only the positions (and forces) change
  // (ggem_update_positions).  The first reinit of a particle set builds.
  void reinit(bool& neighbor_list_update_flag) {
    if (neighbor_list_update_flag || !lists_) {
      if (pending_) check(ggem_set_particles(ctx_, int64_t(n_), xyz_.data(), force_.data()), ctx_);
      check(ggem_build_neighbors(ctx_), ctx_);
      neighbor_list_update_flag = false;
      lists_ = true;
    } else {
      check(ggem_update_positions(ctx_, xyz_.data(), force_.data()), ctx_);
    }
    pending_ = false;
    ++version_;
  }

  std::size_t num_particles() const { return n_; }
  const std::vector<double>& positions() const { return xyz_; }
  const std::vector<double>& forces() const { return force_; }
  uint64_t version() const { return version_; }
  int64_t n_elem() const {
    return int64_t(params_.n_elem[0]) * params_.n_elem[1] * params_.n_elem[2];
  }

  // PointParticle::elem_id of every particle
  std::vector<int32_t> elem_ids() const {
    std::vector<int32_t> e(n_);
    check(ggem_get_elem_ids(ctx_, e.data()), ctx_);
    return e;
  }

  // PointParticle::neighbor_list of particle i (point_mesh.C:803-848)
  std::vector<int32_t> neighbor_list(std::size_t i) const {
    std::vector<int64_t> off(n_ + 1);
    check(ggem_get_neighbor_csr(ctx_, off.data(), nullptr), ctx_);
    std::vector<int32_t> ids(std::max<int64_t>(off[n_], 1));
    check(ggem_get_neighbor_csr(ctx_, off.data(), ids.data()), ctx_);
    return std::vector<int32_t>(ids.begin() + off[i], ids.begin() + off[i + 1]);
  }

  // PointMesh::get_elem_point_neighbor_list (point_mesh.h:537-541)
  std::vector<int32_t> get_elem_point_neighbor_list(int64_t elem_id) const {
    int64_t off[2];
    check(ggem_get_elem_point_csr(ctx_, elem_id, 1, off, nullptr), ctx_);
    std::vector<int32_t> ids(std::max<int64_t>(off[1], 1));
    check(ggem_get_elem_point_csr(ctx_, elem_id, 1, off, ids.data()), ctx_);
    ids.resize(off[1]);
    return ids;
  }

 private:
  ggem_params params_;
  ggem_ctx* ctx_ = nullptr;
  std::size_t n_ = 0;
  std::vector<double> xyz_, force_;
  uint64_t version_ = 0;
  bool pending_ = false;  // positions / forces not yet on the device
  bool lists_ = false;    // the device holds lists of this particle set
};

class GGEMStokes {
 public:
  // GGEMStokes::local_solution_bead (ggem_stokes.h:187-189)
  Point local_solution_bead(PointMesh* pm, const std::size_t& bead_id,
                            const std::string& force_type) const {
    require_regularized(force_type);
    if (bead_id >= pm->num_particles()) throw std::runtime_error("copss_ggem: bead id out of range");
    if (cache_version_ != pm->version() || cache_mesh_ != pm) {
      u_loc_.assign(3 * pm->num_particles(), 0.0);
      check(ggem_local_velocity_beads(pm->ctx(), u_loc_.data()), pm->ctx());
      cache_version_ = pm->version();
      cache_mesh_ = pm;
    }
    return {u_loc_[3 * bead_id], u_loc_[3 * bead_id + 1], u_loc_[3 * bead_id + 2]};
  }

  // GGEMStokes::local_solution_field (ggem_stokes.h:176-179); batched form
  // for many points with their element ids.
  Point local_solution_field(PointMesh* pm, const Point& ptx, const std::string& force_type,
                             const int32_t& ptx_elem_id) const {
    require_regularized(force_type);
    Point u;
    check(ggem_local_velocity_points(pm->ctx(), 1, ptx.data(), &ptx_elem_id, u.data()), pm->ctx());
    return u;
  }
  std::vector<double> local_solution_field(PointMesh* pm, const std::vector<double>& pts,
                                           const std::vector<int32_t>& elem) const {
    std::vector<double> u(pts.size());
    check(ggem_local_velocity_points(pm->ctx(), int64_t(pts.size() / 3), pts.data(),
                                     elem.empty() ? nullptr : elem.data(), u.data()),
          pm->ctx());
    return u;
  }

  // GGEMStokes::global_self_exclusion (ggem_stokes.h:195-197)
  Point global_self_exclusion(PointMesh* pm, const std::size_t& pid0) const {
    ggem_consts c;
    check(ggem_get_consts(pm->ctx(), &c), pm->ctx());
    const auto& f = pm->forces();
    return {c.coeff_1 * f[3 * pid0], c.coeff_1 * f[3 * pid0 + 1], c.coeff_1 * f[3 * pid0 + 2]};
  }

  // PMSystemStokes::compute_point_velocity("disturbed") (pm_system_stokes.C:386-531)
  // with the FEM velocity at the beads supplied by the caller.
  void compute_point_velocity(PointMesh* pm, const std::vector<double>& u_global,
                              std::vector<double>& pv) const {
    pv.assign(3 * pm->num_particles(), 0.0);
    check(ggem_combine_velocity(pm->ctx(), u_global.empty() ? nullptr : u_global.data(), pv.data()),
          pm->ctx());
  }

  // The velocity mesh once (20 node numbers per element in HEX20 order),
  // then compute_point_velocity with point_value (pm_system_stokes.C:441-443)
  // evaluated on the device from the nodal velocity node_u[3 * node + d].
  void set_velocity_mesh(PointMesh* pm, int64_t n_nodes, const std::vector<int32_t>& elem_nodes) const {
    if (int64_t(elem_nodes.size()) != 20 * pm->n_elem())
      throw std::runtime_error("set_velocity_mesh: elem_nodes must hold 20 node numbers per element");
    check(ggem_set_fe_mesh(pm->ctx(), n_nodes, elem_nodes.data()), pm->ctx());
  }
  void compute_point_velocity_fe(PointMesh* pm, const std::vector<double>& node_u,
                                 std::vector<double>& pv) const {
    pv.assign(3 * pm->num_particles(), 0.0);
    check(ggem_combine_velocity_fe(pm->ctx(), node_u.data(), pv.data()), pm->ctx());
  }

  // AssembleStokes::assemble_global_F spreading part: Fe of every element
  // (compute_element_rhs, assemble_stokes.C:643-672), IF table as
  // assemble_int_force builds it.
  void element_force_vectors(PointMesh* pm, int nqp_1d, const std::vector<double>& IF,
                             std::vector<double>& Fe, std::vector<double>* rho = nullptr) const {
    const int nqp = nqp_1d * nqp_1d * nqp_1d;
    const int n_u = int(IF.size()) / nqp;
    const int64_t ne = pm->n_elem();
    Fe.assign(size_t(ne) * 3 * n_u, 0.0);
    if (rho) rho->assign(size_t(ne) * nqp * 3, 0.0);
    check(ggem_spread(pm->ctx(), nqp_1d, n_u, IF.data(), rho ? rho->data() : nullptr, Fe.data()),
          pm->ctx());
  }

 private:
  static void require_regularized(const std::string& force_type) {
    if (force_type != "regularized")  // ggem_stokes.C:403-409, :458-464
      throw std::runtime_error("copss_ggem: invalid force_type: " + force_type);
  }
  mutable std::vector<double> u_loc_;
  mutable uint64_t cache_version_ = ~uint64_t(0);
  mutable const PointMesh* cache_mesh_ = nullptr;
};

class BrownianSystem {
 public:
  explicit BrownianSystem(PointMesh* pm) : pm_(pm) {}

  // BrownianSystem::chebyshev_polynomial_approximation (brownian_system.h:239-243):
  // dw is input and output (M^{-1/2} dw), returns convergence.
  bool chebyshev_polynomial_approximation(const std::size_t N, const double eigen_min,
                                          const double eigen_max, const double tol_cheb,
                                          std::vector<double>* dw) {
    std::vector<double> y(dw->size());
    ggem_cheb_info info;
    check(ggem_cheb_brownian(pm_->ctx(), op_ptr(), int32_t(N), eigen_min, eigen_max, tol_cheb,
                             GGEM_CHEB_REFERENCE, 1, dw->data(), y.data(), nullptr, &info),
          pm_->ctx());
    *dw = y;
    last_ = info;
    return info.converged != 0;
  }

  // BrownianSystem::hi_ewald (brownian_system.h:249-252): u = M f
  void hi_ewald(const std::vector<double>& f, std::vector<double>& u) {
    u.assign(f.size(), 0.0);
    if (op_.apply) {
      if (op_.apply(op_.user, f.data(), u.data()) != 0) throw std::runtime_error("copss_ggem: operator failed");
      return;
    }
    // _MatMult_Stokes assigns f to the particle forces (brownian_system.C:76)
    // and leaves it there
    pm_->set_forces(f);
    check(ggem_combine_velocity(pm_->ctx(), nullptr, u.data()), pm_->ctx());
  }

  // BrownianSystem::compute_eigenvalues (brownian_system.C:803-843), Lanczos
  // in place of SLEPc Krylov-Schur; raw extreme Ritz values.
  void compute_eigenvalues(double& eig_min, double& eig_max, int iters = 40, uint64_t seed = 1) {
    check(ggem_lanczos_bounds(pm_->ctx(), op_ptr(), iters, seed, &eig_min, &eig_max), pm_->ctx());
  }

  // FEM-inclusive mobility: route every application through a host callback.
  void set_operator(const ggem_op& op) { op_ = op; }
  const ggem_cheb_info& last_info() const { return last_; }

 private:
  const ggem_op* op_ptr() const { return op_.apply ? &op_ : nullptr; }
  PointMesh* pm_;
  ggem_op op_{nullptr, nullptr};
  ggem_cheb_info last_{};
};

}  // namespace copss_amd
