/*
 * copss_ggem.h -- C-ABI of the MI355X (gfx950) GGEM particle path.
 *
 * Drop-in boundary for the reference's GGEM hot path.  Each entry point below
 * names the reference interface it replaces (file:line, relative to the
 * reference source root).  Per-bead and per-element reference calls are
 * batched into whole-array calls.
 *
 * Conventions
 *  - Vectors crossing the ABI use the reference's global layout: 3 doubles
 *    per particle, interleaved [3*i + j] in particle-id order
 *    (brownian_system.C:253-263, pm_system_stokes.C:494-497).
 *  - Host pointers unless the call says otherwise.  Inputs are copied to the
 *    device; outputs are copied back.  An output pointer may be NULL: the
 *    result then stays in the context's device buffers (see the
 *    ggem_get_* calls), which is how callers keep a pipeline on the GPU.
 *  - Every call returns 0 on success and a negative GGEM_ERR_* code on
 *    failure; ggem_last_error() describes the last failure.  Nothing throws
 *    across the ABI (the reference aborts through libmesh_error(); a C++
 *    caller turns a non-zero return into that).
 *  - A context is bound to one HIP device and one stream.  Calls on one
 *    context are not reentrant; different contexts are independent.  Calls
 *    are stream ordered and return after their outputs are on the host
 *    (or, with NULL outputs, after the work is queued).
 */
#ifndef COPSS_GGEM_H
#define COPSS_GGEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define GGEM_OK 0
#define GGEM_ERR_INVALID_ARG (-1)
#define GGEM_ERR_OUT_OF_DOMAIN (-2)  /* a particle lies outside the mesh       */
#define GGEM_ERR_INVALID_ELEM (-3)   /* ggem_stokes.C:369-380                   */
#define GGEM_ERR_HIP (-4)            /* HIP runtime failure (text in last_error) */
#define GGEM_ERR_NOT_READY (-5)      /* call order: particles / lists missing   */
#define GGEM_ERR_STENCIL_TIE (-6)    /* R_e within 1e-9 of a cell-centre distance */
#define GGEM_ERR_NO_DEVICE (-7)      /* no HIP device visible                   */
#define GGEM_ERR_CAPACITY (-8)       /* a size exceeds what this build supports */

/* ---- parameters ------------------------------------------------------- */
/* PointType, point_particle.h:36-41 */
enum { GGEM_POLYMER_BEAD = 0, GGEM_POINT_PARTICLE = 1, GGEM_LAGRANGIAN_POINT = 2 };
/* arithmetic of the pair sum and the spreading */
enum { GGEM_FP64 = 0, GGEM_FP32C = 1 };
/* compat_flags */
#define GGEM_COMPAT_ELEM_HYSTERESIS 1u /* keep a particle's element while it still
                                          contains it (point_mesh.C:589-596)       */
#define GGEM_COMPAT_STALE_LISTS 2u     /* a particle whose new neighbour list is
                                          empty keeps its previous one
                                          (PointParticle::set_neighbor_list,
                                          point_particle.h:189-194); plain
                                          contexts only                          */

/*
 * Reference sources of each field: viscosity_0 / alpha / br0 / ibm_beta /
 * minimum mesh size (AssembleStokes::init_ggem_stokes, assemble_stokes.C:1061-1085;
 * values from copss_point_particle_system.C:178-317); the slit box,
 * periodicity and n_mesh (copss.C:971-1127); the element search radius
 * R_e = 0.5 h_max + 4/alpha (copss.C:1074-1075).
 */
typedef struct ggem_params {
  double mu;              /* viscosity; 1/(6 pi) in the reference's units        */
  double alpha;           /* GGEM splitting parameter                            */
  double ksi;             /* > 0: this xi; <= 0: GGEMStokes::set_ksi's rule       */
  double br0;             /* bead radius (1)                                     */
  double ibm_beta;        /* LAGRANGIAN_POINT: xi = 1/(ibm_beta * hmin)           */
  double hmin;
  double box_min[3];
  double box_max[3];
  int32_t periodic[3];
  int32_t n_elem[3];      /* uniform hex grid; element id = i + nx*(j + ny*k)     */
  double search_radius_e; /* > 0: this R_e; <= 0: 0.5*h_max + 4/alpha            */
  int32_t point_type;     /* GGEM_POLYMER_BEAD ...                                */
  int32_t precision;      /* GGEM_FP64 or GGEM_FP32C                              */
  uint32_t compat_flags;  /* GGEM_COMPAT_*                                        */
  int32_t device;         /* HIP device ordinal                                  */
} ggem_params;

/* Derived constants, GGEMStokes::set_mu/set_alpha/set_ksi (ggem_stokes.C:39-114) */
typedef struct ggem_consts {
  double one_eight_pi_mu, alpha, alpha2, alpha3_pi_23, two_alpha_sqrt_pi, coeff_1;
  double ksi, ksi2, two_ksi_sqrt_pi, coeff_2;
  double search_radius_e, search_radius_p;
  double h[3], L[3];
  int32_t stencil_rows;   /* (dj,dk) rows of the element stencil              */
  int32_t stencil_cells;  /* elements within R_e of an interior element, self incl. */
} ggem_consts;

typedef struct ggem_ctx ggem_ctx;

/* Context: owns all device buffers and one HIP stream.  Replaces the
 * construction of PointMesh (point_mesh.C:139-156) and the GGEMStokes set-up
 * (assemble_stokes.C:1061-1085).  *out is NULL on failure. */
int ggem_create(const ggem_params* params, ggem_ctx** out);
int ggem_destroy(ggem_ctx* ctx);
const char* ggem_last_error(const ggem_ctx* ctx); /* never NULL; ctx may be NULL */
int ggem_get_consts(const ggem_ctx* ctx, ggem_consts* out);
int ggem_set_precision(ggem_ctx* ctx, int32_t precision);
int ggem_sync(ggem_ctx* ctx);

/* ---- particles -------------------------------------------------------- */
/* Positions (and forces, may be NULL = zero) of n particles, copied to the
 * device.  Replaces the PointParticle store read by PointMesh::reinit
 * (point_mesh.C:245-289; point_particle.h:51-363).  Positions must already be
 * wrapped into the box on periodic axes (PMPeriodicBoundary::correct_position,
 * pm_periodic_boundary.C:305-328). */
int ggem_set_particles(ggem_ctx* ctx, int64_t n, const double* xyz, const double* f);
/* Forces only: the "assign" branch of BrownianSystem::extract_particle_vector
 * used by _MatMult_Stokes (brownian_system.C:76). */
int ggem_set_forces(ggem_ctx* ctx, const double* f);

/* ---- (a) neighbour build ---------------------------------------------- */
/* PointMesh<3>::reinit(bool& update_flag) with update_flag true
 * (point_mesh.h:387, point_mesh.C:245-289): bins particles into elements,
 * builds the element-point lists and the particle lists. */
int ggem_build_neighbors(ggem_ctx* ctx);
/* PointMesh<3>::reinit(bool& update_flag) with update_flag false: new
 * positions (3n, particle order; NULL: the context's current ones, after
 * ggem_move_particles) and forces (NULL: zero, as the reinit zeroes them,
 * point_mesh.C:249-251); the particle lists, element ids and element lists
 * of the last build are kept and every pair displacement is recomputed from
 * the new positions with the explicit minimum image
 * (reinit_neighbor_vector, point_mesh.C:220-241).  Plain contexts only. */
int ggem_update_positions(ggem_ctx* ctx, const double* xyz, const double* f);
/* Element of each particle: PointParticle::elem_id after
 * build_elem_point_containing_list (point_mesh.C:573-648). */
int ggem_get_elem_ids(ggem_ctx* ctx, int32_t* elem /* n */);
/* Number of (ordered) neighbour pairs sum_i |nbr(i)| of the last build. */
int ggem_neighbor_pairs(ggem_ctx* ctx, int64_t* npairs);
/* Particle neighbour lists in the reference's order as CSR
 * (build_point_point_neighbor_list, point_mesh.C:803-848): offsets[n+1],
 * ids[npairs].  ids may be NULL to fetch offsets only. */
int ggem_get_neighbor_csr(ggem_ctx* ctx, int64_t* offsets, int32_t* ids);
/* Element-point neighbour list of elements [e0, e0+ne): offsets[ne+1] (from
 * 0) and ids; build_elem_point_neighbor_list (point_mesh.C:652-799). */
int ggem_get_elem_point_csr(ggem_ctx* ctx, int64_t e0, int64_t ne, int64_t* offsets, int32_t* ids);

/* ---- (b) local velocity ----------------------------------------------- */
/* GGEMStokes::local_solution_bead for every bead (ggem_stokes.h:187-189,
 * ggem_stokes.C:427-503): u = sum_nbr G_reg(r_ij) f_j (+ the self term for
 * Lagrangian points).  u: 3n or NULL. */
int ggem_local_velocity_beads(ggem_ctx* ctx, double* u);
/* GGEMStokes::global_self_exclusion for every bead (ggem_stokes.C:506-523). */
int ggem_global_self_exclusion(ggem_ctx* ctx, double* uex);
/* PMSystemStokes::compute_point_velocity("disturbed") combination
 * (pm_system_stokes.C:441-457): U = U_glob + U_loc - coeff_1 f + f for
 * polymer beads, U_glob + U_loc for Lagrangian points, 0 for point
 * particles (both reference branches skip that type).  u_glob may be NULL
 * (= 0); recomputes U_loc from the current forces. */
int ggem_combine_velocity(ggem_ctx* ctx, const double* u_glob, double* U);
/* GGEMStokes::local_solution_field at m points (ggem_stokes.h:176-179,
 * ggem_stokes.C:360-424), each with the element that contains it.  elem may
 * be NULL: the element is then located with the particle binning rule. */
int ggem_local_velocity_points(ggem_ctx* ctx, int64_t m, const double* pts, const int32_t* elem,
                               double* u);

/* ---- FEM velocity at the beads ---------------------------------------- */
/* The velocity mesh for ggem_global_velocity_beads: n_nodes nodes and, for
 * each element e = i + nx (j + ny k), its 20 node numbers in libMesh's HEX20
 * order (8 vertices, then the edge mid-nodes 0-1, 1-2, 2-3, 3-0, 0-4, 1-5,
 * 2-6, 3-7, 4-5, 5-6, 6-7, 7-4), as Elem::node_id gives them for the
 * build_cube mesh (copss.C:995-1004).  Copied to the device once; every
 * number must lie in [0, n_nodes).  Single-system contexts only. */
int ggem_set_fe_mesh(ggem_ctx* ctx, int64_t n_nodes, const int32_t* elem_nodes /* 20*n_elem */);
/* U_glob of every bead: point_value(u/v/w, pt, elem) of
 * PMSystemStokes::compute_point_velocity (pm_system_stokes.C:441-443), the
 * HEX20 interpolant of the nodal velocity node_u[3*node + d] in the bead's
 * element of the last neighbour build.  u_glob: 3n, particle order. */
int ggem_global_velocity_beads(ggem_ctx* ctx, const double* node_u, double* u_glob);
/* ggem_combine_velocity with U_glob interpolated on the device from node_u
 * (pm_system_stokes.C:441-457 as a whole); U may be NULL. */
int ggem_combine_velocity_fe(ggem_ctx* ctx, const double* node_u, double* U);
/* Spreading (as ggem_spread with n_u = 20) with the element vectors summed
 * into one force per node of the ggem_set_fe_mesh mesh:
 * F_node[3*node + d] = sum over the node's (element, local node l) of
 * Fe[e][d][l], elements ascending: the point-force part of
 * rhs->add_vector(Fe, dof_indices) in AssembleStokes::assemble_global_F
 * (assemble_stokes.C:508-513), in node order.  Periodic images that the
 * caller numbers as one node are summed into it. */
int ggem_spread_nodes(ggem_ctx* ctx, int32_t nqp_1d, const double* IF, double* F_node /* 3*n_nodes */);

/* ---- global right-hand side: walls, inlet traction, periodic images ---- */
/* AssembleStokes::assemble_global_F (assemble_stokes.C:473-516) with the
 * option of the call ("disturbed" / "undisturbed"):
 *  - compute_element_rhs: the point-force spreading (disturbed, :643-672)
 *    or the traction of the inlet faces, + JxW_face phi_l p_j on every side
 *    of an inlet axis j (undisturbed, :687-727);
 *  - apply_bc_by_penalty (:889-996): on every wall side (a box face that is
 *    neither periodic nor an inlet, select_boundary_side :839-885) and every
 *    node l of it, Fe(d 20 + l) += penalty uvw_d (penalize_elem_matrix_vector,
 *    assemble_system.C:70-107), uvw = the wall's undisturbed velocity
 *    (shear[a]: component shear_direction[a] = shear_rate[a] times the
 *    face's coordinate, :929-948) minus the local velocity of the node in
 *    that element (disturbed), or the unbounded solution minus it
 *    (validation, the "ggem_validation" case, :949-976);
 *  - constrain_element_vector (:508) with libMesh's periodic constraints:
 *    every node on the max face of a periodic axis takes the node at its
 *    image on the min faces as its constraint, so the condensed nodal force
 *    of that node moves there and its own row is zero. */
#define GGEM_RHS_DISTURBED 0
#define GGEM_RHS_UNDISTURBED 1
typedef struct ggem_rhs_opts {
  int32_t option;             /* GGEM_RHS_DISTURBED / GGEM_RHS_UNDISTURBED */
  int32_t validation;         /* wall value = exact_solution_infinite_domain */
  int32_t shear[3];           /* per axis: shear of that axis's two walls */
  int32_t shear_direction[3]; /* velocity component of the shear, 0..2 */
  int32_t inlet[3];           /* inlet axes (no penalty; traction when undisturbed) */
  int32_t pad;
  double shear_rate[3];
  double inlet_pressure[3];
  double penalty;             /* 1e6 in the reference (assemble_stokes.C:910) */
} ggem_rhs_opts;
/* Element vectors Fe[n_elem][3][20] (the order of Fe(d 20 + l)) after the
 * spreading or traction and the penalty rows, before the constraints.  The
 * disturbed option needs built lists; IF as for ggem_spread (20 nodes). */
int ggem_assemble_elem_rhs(ggem_ctx* ctx, int32_t nqp_1d, const double* IF, const ggem_rhs_opts* o,
                           double* Fe /* 60*n_elem */);
/* The same element vectors summed into one force per node of the
 * ggem_set_fe_mesh mesh, elements ascending (rhs->add_vector, :513); with
 * condense != 0 the periodic images are condensed as constrain_element_vector
 * does (each min-face node sums its images' incidences in element order, the
 * max-face rows are zero).  F_node may be NULL: the sums stay on the device
 * (queued, not synchronised). */
int ggem_assemble_rhs(ggem_ctx* ctx, int32_t nqp_1d, const double* IF, const ggem_rhs_opts* o, int32_t condense,
                      double* F_node /* 3*n_nodes */);

/* ---- (c) spreading ---------------------------------------------------- */
/* AssembleStokes::compute_element_rhs point-force branch for every element
 * (assemble_stokes.h:134-141, assemble_stokes.C:643-672 via
 * assemble_global_F :473-516).  nqp_1d = 2 or 3 Gauss points per axis
 * (qp = a + n (b + n c)); IF[n_u*nqp] with IF[l*nqp + qp] = JxW[qp]*phi_l(qp)
 * (assemble_int_force, assemble_stokes.C:733-757), identical for every
 * element of the uniform grid.  rho[n_e][nqp][3] and Fe[n_e][3][n_u] may each
 * be NULL. */
int ggem_spread(ggem_ctx* ctx, int32_t nqp_1d, int32_t n_u, const double* IF, double* rho,
                double* Fe);

/* ---- force fields ----------------------------------------------------- */
/* Bead-spring chains and slit walls, the fixes of the polymer-slit system. */
typedef struct ggem_fix_params {
  int32_t chain_beads;  /* beads per chain, consecutive particle ids; <= 1: no springs */
  int32_t walls;        /* != 0: LJ walls at box_min / box_max of each non-periodic axis */
  double fene_k;        /* FENE: F = k Ls (r/Ls) / (1 - (r/Ls)^2) r_hat, Ls = fene_bmax */
  double fene_bmax;
  double lj_eps, lj_sigma, lj_rcut;  /* wall force within |wall - x| <= lj_rcut */
} ggem_fix_params;
/* Forces of the particles from their current positions: FENE springs between
 * consecutive beads of each chain (FixBase::spring_force_fene,
 * fix_base.C:78-97; bond vectors by the minimum image) plus the LJ wall force
 * of FixPointSlitWallLJCut (FixBase::lj_force, fix_base.C:168-186).  The
 * result replaces the context's forces (as ggem_set_forces would) and is
 * copied to f (3n) unless f is NULL.  Single-system contexts only. */
int ggem_fix_forces(ggem_ctx* ctx, const ggem_fix_params* fp, double* f);

/* Force-field fixes of point particles, FixFactory::buildFix
 * (fix/fix_factory.C:4-84) and the FixPoint* classes it builds, evaluated
 * on the device.  p[] holds each fix's parameters in the reference's
 * dimensionless form:
 *  P_CONSTANT               each particle's constant force (ggem_set_constant_forces;
 *                           FixPointConstant, the data file's force columns)
 *  LJ_CUT                   {epsilon, sigma, rcut}: FixBase::lj_force of every
 *                           neighbour-list pair within rcut (FixPointLJCut,
 *                           fix_point_lj_cut.C:38-59; needs a build)
 *  GAUSSIAN                 {c1, c2}: -c1 c2 exp(-c2 r^2) r_ij over the
 *                           neighbour lists (FixPointGaussian; gaussian_dna is
 *                           this law with c1, c2 from ev, Ss2, bk, Nks,
 *                           fix_point_gaussian_dna.C; needs a build)
 *  WORM_LIKE_SPRING         {c1 = a/(2 bk), Ls = Nks bk/a}: spring_force_wls over
 *                           the bonds (FixPointWLS, fix_point_wls.C:66-96)
 *  FENE                     {k, b_max}: spring_force_fene with c1 = k b_max,
 *                           Ls = b_max over the bonds
 *  SLIT_WALL_LJ_CUT         {epsilon, sigma, rcut}: the two box faces of every
 *                           non-periodic axis (FixPointSlitWallLJCut)
 *  SPHERE_WALL_LJ_CUT       {epsilon, sigma, rcut, radius} (FixPointSphereWallLJCut)
 *  SLIT_WALL_EMPIRICAL_DNA  {c0, d0} (FixPointSlitWallEmpiricalDNA)
 *  SPHERE_WALL_EMPIRICAL_DNA {c0, d0, radius} (FixPointSphereWallEmpiricalDNA)
 *  DISCRETIZED_WALL_LJ_CUT  {epsilon, sigma, rcut}: every wall particle
 *                           (ggem_set_wall_particles) within rcut
 *                           (FixPointDiscretizedWallLJCut)
 * Pair vectors use point_vector's minimum image. */
enum {
  GGEM_FIX_P_CONSTANT = 1,
  GGEM_FIX_LJ_CUT = 2,
  GGEM_FIX_GAUSSIAN = 3,
  GGEM_FIX_WORM_LIKE_SPRING = 4,
  GGEM_FIX_FENE = 5,
  GGEM_FIX_SLIT_WALL_LJ_CUT = 6,
  GGEM_FIX_SPHERE_WALL_LJ_CUT = 7,
  GGEM_FIX_SLIT_WALL_EMPIRICAL_DNA = 8,
  GGEM_FIX_SPHERE_WALL_EMPIRICAL_DNA = 9,
  GGEM_FIX_DISCRETIZED_WALL_LJ_CUT = 10
};
typedef struct ggem_fix {
  int32_t type;
  int32_t pad;
  double p[4];
} ggem_fix;
/* The bonds of PolymerChain (polymer_chain.C:188-202): n_bonds pairs of bead
 * ids (bonds[2b], bonds[2b+1]), 0-based.  Bead 1 of a bond takes F_ij, bead 2
 * F_ji = -F_ij, bond by bond. */
int ggem_set_bonds(ggem_ctx* ctx, int64_t n_bonds, const int64_t* bonds);
/* Each particle's constant force (3n, particle order) for P_CONSTANT. */
int ggem_set_constant_forces(ggem_ctx* ctx, const double* f);
/* Positions of m discretized wall particles (3m) for DISCRETIZED_WALL_LJ_CUT. */
int ggem_set_wall_particles(ggem_ctx* ctx, int64_t m, const double* xyz);
/* The forces of nfix (<= 8) fixes at the context's current positions, added
 * in the order given (each reinit zeroes the forces and every fix adds its
 * part, point_mesh.C:249-251); the sum replaces the context's forces and is
 * copied to f (3n) unless f is NULL.  Single-system contexts only. */
int ggem_apply_fixes(ggem_ctx* ctx, int32_t nfix, const ggem_fix* fixes, double* f);

/* ---- (d) Brownian step ------------------------------------------------ */
/* Mobility operator u = M f: _MatMult_Stokes (brownian_system.h:54-59,
 * brownian_system.C:44-140) with the host FEM part U_glob left out, i.e.
 * the device part of compute_point_velocity.  f, u: 3n host arrays. */
int ggem_mobility_apply(ggem_ctx* ctx, const double* f, double* u);

/* Brownian noise kept on the device: nrhs vectors of 3n values (particle
 * order, interleaved), used by ggem_cheb_brownian when its dw is NULL.  The
 * noise vector dw of fixman_integrate (copss.C:1569-1572). */
int ggem_set_noise(ggem_ctx* ctx, int32_t nrhs, const double* dw);

/* Host-assisted operator for a FEM-inclusive mobility: u = M f on host
 * arrays (3n, particle order).  Returns 0 on success. */
typedef struct ggem_op {
  void* user;
  int (*apply)(void* user, const double* f, double* u);
} ggem_op;

typedef struct ggem_cheb_info {
  int32_t converged;  /* reference-compat mode: the stopping rule fired on err1 */
  int32_t terms;      /* highest Chebyshev index used                          */
  int32_t applies;    /* operator applications                                 */
  int32_t pad;
  double err1, err2;  /* estimators of brownian_system.C:1076-1131             */
} ggem_cheb_info;

enum { GGEM_CHEB_REFERENCE = 0, GGEM_CHEB_FIXED = 1 };
/* OR'ed into the mode of ggem_cheb_brownian: y and the displacement M y are
 * computed on the device and kept there, in particle order, for
 * ggem_move_particles, ggem_displacement_apply and ggem_get_brownian, also
 * when the y and disp pointers are NULL. */
#define GGEM_CHEB_KEEP 0x100

/* BrownianSystem::chebyshev_polynomial_approximation (brownian_system.h:239-243,
 * brownian_system.C:967-1178) followed by hi_ewald (copss.C:1686):
 *   y = M^{-1/2} dw by the three-term Chebyshev recurrence on
 *   [eig_min, eig_max] (coefficients: chebyshev.C:57-151), disp = M y.
 * mode GGEM_CHEB_REFERENCE: the reference's estimators and stopping rule
 * (err1 <= tol or err2 <= 1e-9, at most N terms); GGEM_CHEB_FIXED: exactly N
 * terms and no estimator applications.  nrhs noise vectors are processed
 * together (dw, y, disp: nrhs blocks of 3n); dw == NULL uses the noise set
 * by ggem_set_noise (nrhs must match).  op == NULL: device M; else the
 * recurrence runs on the device and op->apply is called per application.
 * y and disp may be NULL; info (nrhs entries) may be NULL. */
int ggem_cheb_brownian(ggem_ctx* ctx, const ggem_op* op, int32_t N, double eig_min,
                       double eig_max, double tol, int32_t mode, int32_t nrhs, const double* dw,
                       double* y, double* disp, ggem_cheb_info* info);

/* enable != 0: ggem_cheb_brownian in GGEM_CHEB_FIXED mode with op == NULL
 * replays one captured HIP graph of its N + 1 mobility applications and
 * vector updates instead of launching them one by one.  The capture is
 * redone when the particle set, the neighbour lists, N, nrhs, the precision
 * or whether disp is asked for changes; the eigenvalue bounds and the
 * Chebyshev coefficients are device parameters written before each replay.
 * No counterpart in the reference, whose PETSc loop issues each operation
 * itself (brownian_system.C:1030-1131); results equal the non-graph path. */
int ggem_set_graph(ggem_ctx* ctx, int32_t enable);

/* Extreme-eigenvalue estimates of M by Lanczos from a seeded start vector,
 * replacing the SLEPc Krylov-Schur bounds of BrownianSystem::compute_eigenvalues
 * (brownian_system.C:527-657, :803-843).  Raw Ritz values: the caller applies
 * eig_factor as copss.C:1607 does. */
int ggem_lanczos_bounds(ggem_ctx* ctx, const ggem_op* op, int32_t iters, uint64_t seed,
                        double* lmin, double* lmax);

/* ---- Brownian integration on the device -------------------------------- */
/* Copies of the vectors kept by the last ggem_cheb_brownian with
 * GGEM_CHEB_KEEP (or updated by ggem_displacement_apply): y = M^{-1/2} dw
 * and disp = M y, nrhs blocks of 3n in particle order; either may be NULL. */
int ggem_get_brownian(ggem_ctx* ctx, double* y, double* disp);
/* disp = M y with the kept y, at the context's current positions and
 * neighbour lists: hi_ewald(M, dw, dw_mid) at the mid-point of
 * Copss::fixman_integrate (copss.C:1730).  The result replaces the kept
 * displacement; it is copied to disp (3n x nrhs) unless disp is NULL. */
int ggem_displacement_apply(ggem_ctx* ctx, double* disp);

/* Noise stream kept in the context: the engine of RandomGenerator
 * (random_generator.C:43-67, one std::default_random_engine for the whole
 * run, seeded by set_seed :96-99; copss.C:1389).  ggem_noise_seed restarts
 * the stream; ggem_noise_skip draws and discards n_vectors vectors of 3n
 * N(0,1) values (the restart catch-up of copss.C:1544-1555);
 * ggem_noise_draw draws the next vector dw = sqrt(2) N(0, sqrt(dt)) over
 * 3n values (std_random_vector + VecScale, copss.C:1565-1572), which
 * becomes the device noise (nrhs = 1, as ggem_set_noise) and is copied to
 * dw unless dw is NULL.  The draws run on the host (the engine is
 * sequential); each costs O(n). */
int ggem_noise_seed(ggem_ctx* ctx, uint64_t seed);
int ggem_noise_skip(ggem_ctx* ctx, int64_t n_vectors);
int ggem_noise_draw(ggem_ctx* ctx, double dt, double* dw);
/* Device noise from a counter-based generator: nrhs vectors of 3n values
 * sqrt(2) N(0, sqrt(dt)) (Philox4x32-10 keyed by seed, then the Box-Muller
 * transform; values 2j and 2j + 1 of vector r come from the counter words
 * (counter low, counter high, r, j)), written
 * straight to the device noise.  No counterpart in the reference, whose
 * noise is the host engine above; for throughput runs with many noise
 * vectors. */
int ggem_noise_philox(ggem_ctx* ctx, uint64_t seed, uint64_t counter, int32_t nrhs, double dt);
/* A copy of the device noise (noise_nrhs blocks of 3n, particle order). */
int ggem_get_noise(ggem_ctx* ctx, double* dw);

/* Positions updated on the device, Copss::fixman_integrate /
 * langevin_integrate (copss.C:1676-1745, :1916-1925):
 *   x = base + c_velocity U + c_disp D + c_force f + c_noise dw
 * with base the positions saved by ggem_save_positions (R0) when from_saved
 * is set, else the current positions; U the velocity of the last
 * ggem_combine_velocity / ggem_combine_velocity_fe, D the kept displacement,
 * f the context's forces (the free-draining velocity of langevin_integrate)
 * and dw the device noise.  Then FixPoint::check_walls (fix_point.C:27-90):
 * wall_type 0 (slit): a periodic axis wraps once, x < min: x += L,
 * x >= max: x -= L (counting the image crossings), and a wall axis
 * reflects, x < min: 2 min - x, x >= max: 2 max - x; wall_type 1: a point
 * outside the spherical cavity of cavity_radius about the origin is
 * reflected to (2 R - |x|) x/|x|.  With track != 0 the increment (without
 * the wrap) is added to the unwrapped positions ROUT (copss.C:1740-1743).
 * The new positions become the context's; a neighbour build or
 * ggem_update_positions follows before the next pair sum. */
typedef struct ggem_move {
  double c_velocity, c_disp, c_force, c_noise;
  int32_t from_saved;
  int32_t wall_type;     /* 0 slit (box faces), 1 spherical cavity */
  double cavity_radius;
  int32_t track;
  int32_t pad;
} ggem_move;
int ggem_save_positions(ggem_ctx* ctx);
int ggem_move_particles(ggem_ctx* ctx, const ggem_move* mv);
/* Current positions (3n, wrapped), the unwrapped positions ROUT and the
 * periodic image counts (PointParticle::counter); any output may be NULL. */
int ggem_get_positions(ggem_ctx* ctx, double* xyz, double* unwrapped, int32_t* images);
/* ROUT and image counts restart from rout (3n; NULL: the current
 * positions, zero counts), as Copss::run reads vector_RIN (copss.C:1392-1413). */
int ggem_set_unwrapped(ggem_ctx* ctx, const double* rout);
/* PointMesh::set_bead_velocity's extrema (point_mesh.C:1042-1061): the
 * largest and smallest |v_i| over the particles, the input of
 * PMSystemStokes::get_dt's adaptive step (pm_system_stokes.C:77-134).
 * which = 0: v = the velocity of the last ggem_combine_velocity(_fe)
 * (fixman_integrate, copss.C:1476-1491); which = 1: v = the context's
 * forces (langevin_integrate takes the force as the velocity, :1812-1822).
 * Either output may be NULL. */
int ggem_velocity_extrema(ggem_ctx* ctx, int32_t which, double* vmax, double* vmin);
/* PolymerChain::check_chain (polymer_chain.C:1826-1867) on the context's
 * positions, over the bonds of ggem_set_bonds in bond order: a bond whose
 * minimum-image length is at least Ls moves its bead 2 to 0.9 Ls from bead
 * 1 along the bond, then into the box on periodic axes
 * (PMPeriodicBoundary::correct_position, pm_periodic_boundary.C:305-328);
 * later bonds see the moved bead.  broken (may be NULL, then nothing is
 * synchronised) = 1 if some bond was at least Ls. */
int ggem_check_chain(ggem_ctx* ctx, double Ls, int32_t* broken);

/* ---- batched suspensions ---------------------------------------------- */
/* Many independent suspensions in one context: the particles, cells and
 * neighbour lists of all systems share one device store, so every neighbour
 * build, pair sum and operator application covers the whole batch in one
 * launch.  No counterpart in the reference, which runs one suspension per
 * process (copss.C:1565-1745); each system follows the same arithmetic as a
 * plain context built from its own params.
 * params[nsys]: per-system box, grid, periodicity and search radius; mu,
 * alpha, xi, point type, precision, compat flags and device must agree.
 * On a batched context the whole-store calls work unchanged (build,
 * local_velocity_beads, combine_velocity, global_self_exclusion,
 * mobility_apply, set_forces, set_noise, get_elem_ids with element ids
 * numbered system after system); the single-mesh calls (spread,
 * local_velocity_points, the CSR exports, cheb_brownian, lanczos_bounds)
 * return GGEM_ERR_INVALID_ARG. */
int ggem_create_batch(const ggem_params* params, int32_t nsys, ggem_ctx** out);
/* counts[nsys] particles per system; xyz, f hold the systems one after the
 * other (3 * sum(counts) values, interleaved). */
int ggem_batch_set_particles(ggem_ctx* ctx, const int64_t* counts, const double* xyz, const double* f);
/* Fixed-order Chebyshev step of every system with its own bounds
 * eig_min[k], eig_max[k] (GGEM_CHEB_FIXED semantics of ggem_cheb_brownian,
 * device operator).  dw, y, disp: nrhs blocks of 3 * sum(counts); dw NULL
 * uses ggem_set_noise's vectors.  Replayed from a captured graph when
 * ggem_set_graph is on. */
int ggem_batch_cheb_brownian(ggem_ctx* ctx, int32_t N, const double* eig_min, const double* eig_max,
                             int32_t nrhs, const double* dw, double* y, double* disp);
/* Lanczos bounds of every system's mobility (ggem_lanczos_bounds per
 * system, start vector of system k from mt19937_64(seed + k)).  lmin, lmax:
 * nsys values each. */
int ggem_batch_lanczos_bounds(ggem_ctx* ctx, int32_t iters, uint64_t seed, double* lmin, double* lmax);

/* ---- host utilities (no GPU needed) ---------------------------------- */
/* The reference's Brownian noise: std::default_random_engine(seed) +
 * std::normal_distribution(0, sqrt(dt)), each value times sqrt(2)
 * (random_generator.C:43-67, copss.C:1565-1572); skip_vectors discarded
 * vectors of length n first (copss.C:1544-1555). */
int ggem_util_reference_noise(int64_t n, uint64_t seed, double dt, int64_t skip_vectors,
                              double* out);
/* The same generator as a stream that continues across calls (one
 * RandomGenerator for the whole run, brownian_system.h:292, copss.C:1389):
 * ggem_util_noise_draw gives the next sqrt(2) N(0, sqrt(dt)) vector of n
 * values (copss.C:1565-1572), ggem_util_noise_skip discards n_vectors
 * vectors of n N(0,1) values (the restart catch-up, copss.C:1544-1555).
 * Draw k after seeding equals ggem_util_reference_noise(n, seed, dt, k). */
typedef struct ggem_noise_stream ggem_noise_stream;
int ggem_util_noise_create(uint64_t seed, ggem_noise_stream** out);
int ggem_util_noise_destroy(ggem_noise_stream* s);
int ggem_util_noise_draw(ggem_noise_stream* s, int64_t n, double dt, double* out);
int ggem_util_noise_skip(ggem_noise_stream* s, int64_t n, int64_t n_vectors);
/* Random sequential addition of n points uniform in the box, rejecting any
 * pair closer than min_dist (minimum image on periodic axes);
 * std::mt19937_64(seed), uniform_real_distribution per axis x, y, z. */
int ggem_util_rsa_positions(const ggem_params* params, int64_t n, double min_dist, uint64_t seed,
                            double* xyz);
/* n i.i.d. N(0,1) values from std::mt19937_64(seed). */
int ggem_util_normal(int64_t n, uint64_t seed, double* out);
/* The (element, local node) items of the wall and inlet sides of an
 * n_elem[0] x n_elem[1] x n_elem[2] box grid for these options, as
 * ggem_assemble_elem_rhs visits them: items[2q] = element, items[2q+1] =
 * l | wall side mask << 8 | inlet side mask << 16 (bit s for hex side s:
 * 0 z-, 1 y-, 2 x+, 3 y+, 4 x-, 5 z+), elements and nodes ascending.  The
 * inlet masks are set for the undisturbed option only.  items may be NULL
 * to count. */
int ggem_util_rhs_wall_items(const int32_t* n_elem /* 3 */, const int32_t* periodic /* 3 */,
                             const ggem_rhs_opts* o, int64_t* m, int32_t* items /* 2m */);
/* JxW_face phi_l of every side s of a HEX20 element of size h[3] at its
 * 2 x 2 Gauss points (QGauss(2, SECOND) on the side, the side's first
 * reference axis fastest): tab[(s 4 + qp) 20 + l], the inlet traction of
 * compute_element_rhs (assemble_stokes.C:687-727). */
int ggem_util_hex20_traction_table(const double* h /* 3 */, double* tab /* 480 */);
/* IF table of a HEX20 element of size h[3] with nqp_1d^3 Gauss points:
 * IF[l*nqp + qp] = JxW[qp] * phi_l(qp), libMesh HEX20 node order. */
int ggem_util_hex20_if_table(const double* h, int32_t nqp_1d, double* IF /* 20*nqp */);
/* The FP32C pair coefficients (A, Bx) of G_reg = A I + Bx x x^T
 * (ggem_stokes.C:304-357) as a piecewise polynomial in s = r^2, the form
 * the FP32C kernels evaluate: n_intervals intervals of 16 floats (A's 8
 * monomial coefficients in y = 2 (s - mid)/d, then Bx's), the last one
 * zero (s beyond the cut-off, where min(alpha^2, xi^2) s >= 20 and both
 * coefficients are below e^-20 of their size); meta = {s1, 1/d1, 1/d2, d1,
 * d2, n1, n2, s_zero} (intervals of width d1 on [0, s1), d2 after, zero
 * from s_zero on).  The FP32C pair sums skip the pairs of the zero
 * interval.  table = NULL queries
 * n_intervals; cap is the table's size in floats. */
int ggem_util_fp32c_table(double eps, double alpha, double ksi, int32_t cap, float* table, float* meta,
                          int32_t* n_intervals);

/* ---- launch-shape autotuning -------------------------------------------- */
/* enable != 0: the first launch of the pair-tensor, pair-sum and mobility
 * kernels for a new signature (precision, operand count, log2 of the
 * particle count) times each candidate launch shape (block size, XCD block
 * renumbering) and keeps the fastest for the context; all shapes give
 * bit-identical results.  Off (the default): the default shape.  No
 * counterpart in the reference (CPU loops have no launch shape). */
int ggem_set_autotune(ggem_ctx* ctx, int32_t enable);
/* Pin the pick of one signature (as printed by ggem_tuning_report), e.g. to
 * replay a tuning table recorded in an earlier run; used while autotuning
 * is on.  A pick outside the kernel's candidate list makes that kernel's
 * launches fail with GGEM_ERR_HIP. */
int ggem_tune_set(ggem_ctx* ctx, const char* signature, int32_t pick);
/* The picks so far, one line per signature: "<signature> <pick> <ms per
 * candidate>...".  buf (cap bytes, NUL-terminated) may be NULL; *len gets
 * the full length. */
int ggem_tuning_report(ggem_ctx* ctx, char* buf, int64_t cap, int64_t* len);

/* ---- profiling ---------------------------------------------------------- */
/* When enabled, every kernel launch is bracketed by HIP events on the
 * context's stream.  ggem_profile_get sums the durations per kernel name
 * ("*": over all kernels). */
int ggem_profile_enable(ggem_ctx* ctx, int32_t enable);
int ggem_profile_reset(ggem_ctx* ctx);
int ggem_profile_get(ggem_ctx* ctx, const char* kernel, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
