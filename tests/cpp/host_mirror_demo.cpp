// Drives the C++ host classes (PointMesh / GGEMStokes / BrownianSystem) the
// way the reference's point-particle app drives its own, on the GGEM
// validation case and on 100 periodic beads, and prints the results as JSON
// for tests/test_gpu_host_mirror.py to compare with the oracle.
#include <cmath>
#include <cstdio>
#include <vector>

#include "copss_gpu.hpp"

static ggem_params base_params() {
  ggem_params p{};
  p.mu = 1.0 / (6.0 * M_PI);
  p.alpha = 0.5;
  p.br0 = 1.0;
  p.ibm_beta = 0.75;
  p.point_type = GGEM_POLYMER_BEAD;
  p.precision = GGEM_FP64;
  p.compat_flags = GGEM_COMPAT_ELEM_HYSTERESIS;
  return p;
}

static void print_vec(const char* key, const std::vector<double>& v, bool last = false) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? "," : "", v[i]);
  std::printf("]%s\n", last ? "" : ",");
}

int main() {
  using namespace copss_amd;
  std::printf("{\n");
  // GGEM validation case: 3 beads, 15^3 elements on [-15,15]^3, alpha 0.35
  {
    ggem_params p = base_params();
    p.alpha = 0.35;
    for (int d = 0; d < 3; ++d) {
      p.box_min[d] = -15;
      p.box_max[d] = 15;
      p.n_elem[d] = 15;
    }
    PointMesh pm(p);
    pm.set_particles({-5, 0, 0, 0, 0, 0, 5, 0, 0}, {0.3333, 0, 0, 0.3333, 0, 0, 0.3333, 0, 0});
    bool update = true;
    pm.reinit(update);
    GGEMStokes ggem;
    std::vector<double> pts;
    for (int i = 0; i <= 200; ++i) {
      pts.push_back(-15.0 + i * 0.15);
      pts.push_back(0);
      pts.push_back(0);
    }
    print_vec("ggem3_field_x", ggem.local_solution_field(&pm, pts, {}));
    std::vector<double> ub;
    for (size_t b = 0; b < 3; ++b) {
      Point u = ggem.local_solution_bead(&pm, b, "regularized");
      Point ex = ggem.global_self_exclusion(&pm, b);
      for (int d = 0; d < 3; ++d) ub.push_back(u[d] - ex[d]);
    }
    print_vec("ggem3_bead_minus_exclusion", ub);
    // AssembleStokes::assemble_global_F's element vectors on the walled box:
    // spreading at 8 Gauss points, penalty rows with sheared z walls
    StokesRhsParams rp;
    rp.nqp_1d = 2;
    rp.IF.resize(20 * 8);
    const double h2[3] = {2, 2, 2};
    ggem_util_hex20_if_table(h2, 2, rp.IF.data());
    rp.shear = {{false, false, true}};
    rp.shear_direction = {{0, 0, 0}};
    rp.shear_rate = {{0, 0, 0.05}};
    AssembleStokes as(rp);
    std::vector<double> Fe_w;
    as.element_vectors(&pm, "disturbed", Fe_w);
    print_vec("ggem3_rhs_Fe", Fe_w);
    std::printf("\"update_flag_cleared\": %d,\n", update ? 0 : 1);
    bool threw = false;
    try {
      ggem.local_solution_bead(&pm, 0, "singular");
    } catch (const std::runtime_error&) {
      threw = true;
    }
    std::printf("\"bad_force_type_throws\": %d,\n", threw ? 1 : 0);
  }
  // 100 periodic beads: velocity combination, spreading, Brownian step
  {
    ggem_params p = base_params();
    for (int d = 0; d < 3; ++d) {
      p.box_min[d] = 0;
      p.box_max[d] = 40;
      p.n_elem[d] = 20;
      p.periodic[d] = 1;
    }
    std::vector<double> x(300), f(300), dw(300), IF(20 * 27);
    ggem_util_rsa_positions(&p, 100, 2.2, 1001, x.data());
    ggem_util_normal(300, 2001, f.data());
    ggem_util_reference_noise(300, 123456789, 1e-3, 0, dw.data());
    const double h[3] = {2, 2, 2};
    ggem_util_hex20_if_table(h, 3, IF.data());
    PointMesh pm(p);
    pm.set_particles(x, f);
    bool update = true;
    pm.reinit(update);
    GGEMStokes ggem;
    std::vector<double> ug(300, 0.25), pv, Fe;
    ggem.compute_point_velocity(&pm, ug, pv);
    print_vec("b100_point_velocity", pv);
    // the same with U_glob interpolated on the device from a uniform nodal
    // velocity of 0.25 (every element with its own 20 nodes)
    std::vector<int32_t> conn(20 * 8000);
    for (int32_t k = 0; k < 20 * 8000; ++k) conn[k] = k;
    ggem.set_velocity_mesh(&pm, 20 * 8000, conn);
    std::vector<double> node_u(3 * 20 * 8000, 0.25), pv_fe;
    ggem.compute_point_velocity_fe(&pm, node_u, pv_fe);
    print_vec("b100_point_velocity_fe", pv_fe);
    ggem.element_force_vectors(&pm, 3, IF, Fe);
    double s = 0;
    for (double v : Fe) s += v * v;
    std::printf("\"b100_Fe_norm\": %.17g,\n", std::sqrt(s));
    BrownianSystem bs(&pm);
    double lo, hi;
    bs.compute_eigenvalues(lo, hi, 60, 5);
    std::printf("\"b100_eig\": [%.17g, %.17g],\n", lo, hi);
    lo /= 1.05;
    hi *= 1.05;
    std::vector<double> y = dw;
    const bool conv = bs.chebyshev_polynomial_approximation(30, lo, hi, 0.1, &y);
    std::vector<double> disp;
    bs.hi_ewald(y, disp);
    std::printf("\"b100_cheb\": {\"converged\": %d, \"terms\": %d, \"lo\": %.17g, \"hi\": %.17g},\n", conv ? 1 : 0,
                bs.last_info().terms, lo, hi);
    print_vec("b100_dw", dw);
    print_vec("b100_y", y);
    print_vec("b100_disp", disp);
    // PointMesh::reinit(false): moved positions on the kept lists and elements
    std::vector<int32_t> list0 = pm.neighbor_list(0), elem0 = pm.elem_ids();
    std::vector<double> x2(x);
    for (size_t i = 0; i < x2.size(); ++i) {
      x2[i] += 0.37 * std::sin(1.3 * double(i));
      if (x2[i] < 0) x2[i] += 40;  // correct_position's single shift
      if (x2[i] >= 40) x2[i] -= 40;
    }
    pm.set_particles(x2, f);
    bool keep = false;
    pm.reinit(keep);
    std::vector<double> u2;
    for (size_t b = 0; b < 100; ++b) {
      Point u = ggem.local_solution_bead(&pm, b, "regularized");
      u2.insert(u2.end(), u.begin(), u.end());
    }
    print_vec("b100_x2", x2);
    print_vec("b100_reinit_false_u", u2);
    std::printf("\"b100_lists_kept\": %d,\n", (pm.neighbor_list(0) == list0 && pm.elem_ids() == elem0) ? 1 : 0);
    bool rebuild = true;
    pm.reinit(rebuild);
    std::vector<double> u3;
    for (size_t b = 0; b < 100; ++b) {
      Point u = ggem.local_solution_bead(&pm, b, "regularized");
      u3.insert(u3.end(), u.begin(), u.end());
    }
    print_vec("b100_reinit_true_u", u3);
    // the periodic box's nodal right-hand side, condensed (every element
    // with its own 20 nodes, as set above)
    StokesRhsParams rq;
    rq.nqp_1d = 3;
    rq.IF = IF;
    AssembleStokes aq(rq);
    aq.set_node_count(20 * 8000);
    std::vector<double> F_node;
    aq.assemble_global_F(&pm, "disturbed", F_node);
    print_vec("b100_rhs_F_node", F_node);
    // fixes built by FixFactory from control-file style parameters, bonds
    // along chains of 10 beads
    std::vector<int64_t> bonds;
    for (int64_t c = 0; c < 10; ++c)
      for (int64_t i = 0; i < 9; ++i) {
        bonds.push_back(10 * c + i);
        bonds.push_back(10 * c + i + 1);
      }
    pm.set_bonds(bonds);
    FixParams sp;
    sp.bead_radius = 0.077;
    sp.bk = 0.106;
    sp.Nks = 200.0;
    sp.Ss2 = sp.Nks * sp.bk * sp.bk / 6.0;
    std::vector<ggem_fix> fixes = {FixFactory::buildFix("gaussian_dna", {2.628506592}, sp),
                                   FixFactory::buildFix("worm_like_spring", {}, sp),
                                   FixFactory::buildFix("lj_cut", {1.0, 1.0, 2.5}, sp)};
    pm.apply_fixes(fixes);
    print_vec("b100_fix_forces", pm.forces());
    bool bad_fix = false;
    try {
      FixFactory::buildFix("spring", {}, sp);
    } catch (const std::runtime_error&) {
      bad_fix = true;
    }
    std::printf("\"unknown_fix_throws\": %d\n", bad_fix ? 1 : 0);
  }
  std::printf("}\n");
  return 0;
}
