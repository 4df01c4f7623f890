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
//   AssembleStokes <- AssembleStokes (assemble_stokes.h): assemble_global_F
//                    (spreading or inlet traction, penalty rows, periodic
//                    constraints) into one force per velocity node
//   FixFactory    <- FixFactory::buildFix (fix/fix_factory.C:4-84) with each
//                    fix's initParams; PointMesh::apply_fixes runs the fixes'
//                    compute() as reinit_system does
//
// The reference computes one bead or one element per call; here a call
// computes every bead / element once on the GPU and later calls for other
// beads read the cached result until the particles or forces change.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
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

  // The fixes' compute() of PMSystemStokes::reinit_system
  // (pm_system_stokes.C:178-185) after a reinit: their summed forces become
  // the particle forces (the reinit zeroed them).  The bonds of the springs
  // (bead 1, bead 2 pairs, 0-based) are set once the particles are on the
  // device, i.e. after the first reinit.
  void set_bonds(const std::vector<int64_t>& bonds) {
    if (pending_) throw std::runtime_error("copss_ggem: set_bonds needs a reinit after set_particles");
    check(ggem_set_bonds(ctx_, int64_t(bonds.size() / 2), bonds.data()), ctx_);
  }
  void apply_fixes(const std::vector<ggem_fix>& fixes) {
    if (pending_) throw std::runtime_error("copss_ggem: apply_fixes needs a reinit after set_particles");
    force_.assign(3 * n_, 0.0);
    check(ggem_apply_fixes(ctx_, int32_t(fixes.size()), fixes.data(), force_.data()), ctx_);
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

// The settings of the right-hand side that the reference reads from its
// equation_systems parameters (shear, shear_direction, shear_rate, inlet,
// inlet_pressure, simulation_name) and its quadrature (the IF table of
// assemble_int_force, assemble_stokes.C:733-757).
struct StokesRhsParams {
  int nqp_1d = 3;
  std::vector<double> IF;  // 20 * nqp_1d^3
  std::array<bool, 3> shear{{false, false, false}}, inlet{{false, false, false}};
  std::array<int, 3> shear_direction{{0, 0, 0}};
  std::array<double, 3> shear_rate{{0, 0, 0}}, inlet_pressure{{0, 0, 0}};
  std::string simulation_name = "general";
};

class AssembleStokes {
 public:
  explicit AssembleStokes(const StokesRhsParams& p) : p_(p) {}

  // AssembleStokes::assemble_global_F (assemble_stokes.C:473-516) for
  // option "disturbed" or "undisturbed": the element vectors (spreading or
  // inlet traction, then apply_bc_by_penalty), constrain_element_vector
  // with the periodic constraints, and rhs->add_vector, returned as one
  // force per node of the velocity mesh set by
  // GGEMStokes::set_velocity_mesh (F_node[3 node + d]).
  void assemble_global_F(PointMesh* pm, const std::string& option, std::vector<double>& F_node) const {
    if (n_nodes_ <= 0) throw std::runtime_error("assemble_global_F: set_node_count of the velocity mesh first");
    ggem_rhs_opts o = opts(option);
    F_node.assign(3 * size_t(n_nodes_), 0.0);
    check(ggem_assemble_rhs(pm->ctx(), p_.nqp_1d, p_.IF.empty() ? nullptr : p_.IF.data(), &o, 1, F_node.data()),
          pm->ctx());
  }
  // The element vectors before the constraints, Fe[e][d][l] (the Fe of the
  // element loop after apply_bc_by_penalty).
  void element_vectors(PointMesh* pm, const std::string& option, std::vector<double>& Fe) const {
    ggem_rhs_opts o = opts(option);
    Fe.assign(size_t(pm->n_elem()) * 60, 0.0);
    check(ggem_assemble_elem_rhs(pm->ctx(), p_.nqp_1d, p_.IF.empty() ? nullptr : p_.IF.data(), &o, Fe.data()),
          pm->ctx());
  }
  // the node count of the mesh given to GGEMStokes::set_velocity_mesh
  void set_node_count(int64_t n_nodes) { n_nodes_ = n_nodes; }

 private:
  ggem_rhs_opts opts(const std::string& option) const {
    if (option != "disturbed" && option != "undisturbed")
      throw std::runtime_error("assemble_global_F: invalid option " + option);
    ggem_rhs_opts o{};
    o.option = option == "disturbed" ? GGEM_RHS_DISTURBED : GGEM_RHS_UNDISTURBED;
    o.validation = p_.simulation_name == "ggem_validation" ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
      o.shear[a] = p_.shear[a];
      o.shear_direction[a] = p_.shear_direction[a];
      o.shear_rate[a] = p_.shear_rate[a];
      o.inlet[a] = p_.inlet[a];
      o.inlet_pressure[a] = p_.inlet_pressure[a];
    }
    o.penalty = 1E6;  // assemble_stokes.C:910
    return o;
  }
  StokesRhsParams p_;
  int64_t n_nodes_ = 0;
};

// What FixFactory::buildFix's fixes read from the equation_systems
// parameters in their initParams.
struct FixParams {
  std::string point_particle_model = "polymer_chain";
  std::string wall_type = "slit";
  std::vector<double> wall_params;  // slit: 6 box bounds; sphere: radius
  double bead_radius = 0, bk = 0, Nks = 0, Ss2 = 0;
};

class FixFactory {
 public:
  // FixFactory::buildFix (fix/fix_factory.C:4-84) for point particles, with
  // the parameters of the fix's initParams from force_params (the control
  // file's values for that fix) and the system's parameters.
  static ggem_fix buildFix(const std::string& fix_name, const std::vector<double>& force_params,
                           const FixParams& sp) {
    ggem_fix f{};
    auto need = [&](std::size_t k) {
      if (force_params.size() != k)
        throw std::runtime_error("The force type '" + fix_name + "' requires " + std::to_string(k) + " parameter");
    };
    auto polymer_only = [&]() {
      if (sp.point_particle_model != "polymer_chain")
        throw std::runtime_error("The force field '" + fix_name + "' is only for polymer_chain");
    };
    auto copy = [&](int k) {
      for (int i = 0; i < k; ++i) f.p[i] = force_params[i];
    };
    if (fix_name == "lj_cut") {  // fix_point_lj_cut.C
      need(3);
      f.type = GGEM_FIX_LJ_CUT;
      copy(3);
    } else if (fix_name == "gaussian") {  // fix_point_gaussian.C:14-30
      need(2);
      f.type = GGEM_FIX_GAUSSIAN;
      copy(2);
    } else if (fix_name == "gaussian_dna") {  // fix_point_gaussian_dna.C:31-58
      polymer_only();
      need(1);
      const double r = sp.bead_radius, ev = force_params[0] * r * r * r;
      const double c1_tmp = 3. / (4. * M_PI * sp.Ss2), c1_tmp3 = c1_tmp * c1_tmp * c1_tmp;
      f.type = GGEM_FIX_GAUSSIAN;
      f.p[0] = ev * sp.Nks * sp.Nks * std::sqrt(c1_tmp3);
      f.p[1] = 3. * r * r / (4. * sp.Ss2);
    } else if (fix_name == "worm_like_spring") {  // fix_point_wls.C:31-60
      polymer_only();
      need(0);
      f.type = GGEM_FIX_WORM_LIKE_SPRING;
      f.p[0] = sp.bead_radius / (2.0 * sp.bk);
      f.p[1] = sp.Nks * sp.bk / sp.bead_radius;
    } else if (fix_name == "wall/lj_cut") {
      need(3);
      copy(3);
      if (sp.wall_type == "slit") {
        f.type = GGEM_FIX_SLIT_WALL_LJ_CUT;
      } else if (sp.wall_type == "sphere") {
        f.type = GGEM_FIX_SPHERE_WALL_LJ_CUT;
        f.p[3] = sp.wall_params.at(0);
      } else {
        throw std::runtime_error("Error: FixFactory::buildFix: undefined wall_type for point_particles: " + fix_name);
      }
    } else if (fix_name == "discretized_wall/lj_cut") {  // wall particles via PointMesh's context
      need(4);
      f.type = GGEM_FIX_DISCRETIZED_WALL_LJ_CUT;
      copy(3);
    } else if (fix_name == "wall/empirical_dna") {  // fix_point_slitWall_empirical_dna.C:35-62
      polymer_only();
      const double c1 = sp.bead_radius / sp.bk, c2 = c1 / std::sqrt(sp.Nks);
      f.p[0] = 25.0 * c1;
      f.p[1] = 0.5 / c2;
      if (sp.wall_type == "slit") {
        f.type = GGEM_FIX_SLIT_WALL_EMPIRICAL_DNA;
      } else if (sp.wall_type == "sphere") {
        f.type = GGEM_FIX_SPHERE_WALL_EMPIRICAL_DNA;
        f.p[2] = sp.wall_params.at(0);
      } else {
        throw std::runtime_error("Error: FixFactory::buildFix: undefined wall_type for point_particles: " + fix_name);
      }
    } else if (fix_name == "p_constant") {
      f.type = GGEM_FIX_P_CONSTANT;
    } else {
      throw std::runtime_error("Error: undefined force field for point_particles: " + fix_name);
    }
    return f;
  }
};

}  // namespace copss_amd
