/*
 * ggem_oracle.h -- C interface of the FP64 CPU restatement of the reference's
 * GGEM particle path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the timed CPU baseline ("port") of bench.py.  Only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load it; the product
 * library never links or calls it.
 *
 * Every function restates one reference function in plain FP64 with the same
 * arithmetic and the same enumeration order; the citation next to each one is
 * the reference file:line it follows (paths relative to the reference root).
 */
#ifndef GGEM_ORACLE_H
#define GGEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Point types, src/point_particle.h:36-41 */
enum { ORC_POLYMER_BEAD = 0, ORC_POINT_PARTICLE = 1, ORC_LAGRANGIAN_POINT = 2 };

typedef struct orc_params {
  double mu;              /* viscosity (viscosity_0); 1/(6 pi) non-dimensional */
  double alpha;           /* GGEM alpha */
  double ksi;             /* > 0: use this xi; <= 0: derive as the reference does */
  double br0;             /* bead radius (1 non-dimensional) */
  double ibm_beta;        /* Lagrangian points: xi = 1/(ibm_beta*hmin) */
  double hmin;
  double box_min[3];
  double box_max[3];
  int32_t periodic[3];
  int32_t n_elem[3];      /* uniform hex grid, ids = i + nx*(j + ny*k) */
  double search_radius_e; /* > 0: use it; <= 0: 0.5*h_max + 4/alpha */
  int32_t point_type;
  int32_t keep_stale_lists; /* point_particle.h:189-194 quirk on re-build */
} orc_params;

typedef struct orc_consts {
  double one_eight_pi_mu, alpha, alpha2, alpha3_pi_23, two_alpha_sqrt_pi, coeff_1;
  double ksi, ksi2, two_ksi_sqrt_pi, coeff_2;
  double search_radius_e, search_radius_p;
  double h[3], L[3];
} orc_consts;

/* ---- constants and pointwise kernels ---------------------------------- */
int orc_consts_get(const orc_params* p, orc_consts* c);
void orc_point_vector(const orc_params* p, const double* p0, const double* p1, double* d);
void orc_correct_position(const orc_params* p, double* x);
/* FixPoint::check_walls for one point: wall_type 0 slit, 1 spherical cavity */
void orc_check_walls(const orc_params* p, int wall_type, double radius, double* x, int32_t* count);
void orc_green_local_regularized(const orc_params* p, const double* x, double* G9);
void orc_green_unbounded_smoothed(const orc_params* p, const double* x, double a_or_k, double* G9);
double orc_smoothed_force_exp(const orc_params* p, const double* r);

/* ---- point mesh (neighbor lists) -------------------------------------- */
typedef struct orc_mesh orc_mesh;
orc_mesh* orc_mesh_create(const orc_params* p);
void orc_mesh_destroy(orc_mesh* m);
/* elem-elem list: 0 = literal O(n_e^2) enumeration, 1 = stencil enumeration */
int orc_mesh_build_elem_elem(orc_mesh* m, int method);
int64_t orc_mesh_elem_elem_len(orc_mesh* m, int64_t e);
const int32_t* orc_mesh_elem_elem(orc_mesh* m, int64_t e);
/* full reinit with positions (3n interleaved) and forces (3n interleaved) */
int orc_mesh_reinit(orc_mesh* m, int64_t n, const double* xyz, const double* f);
/* PointMesh::reinit(false): new positions and forces, lists and elements kept */
int orc_mesh_update_positions(orc_mesh* m, const double* xyz, const double* f);
int orc_mesh_set_forces(orc_mesh* m, const double* f);
int32_t orc_mesh_elem_of(orc_mesh* m, int64_t i);
int64_t orc_mesh_elem_point_len(orc_mesh* m, int64_t e);
const int32_t* orc_mesh_elem_point(orc_mesh* m, int64_t e);
int64_t orc_mesh_point_len(orc_mesh* m, int64_t i);
const int32_t* orc_mesh_point(orc_mesh* m, int64_t i);
int64_t orc_mesh_total_pairs(orc_mesh* m);
/* CSR export of all particle lists: offsets[n+1], ids[nnz] */
int orc_mesh_export_csr(orc_mesh* m, int64_t* offsets, int32_t* ids);

/* ---- GGEM sums (FP64, threaded over beads / elements / points) -------- */
int orc_local_velocity_beads(orc_mesh* m, double* u);                 /* 3n */
int orc_global_self_exclusion(orc_mesh* m, double* uex);              /* 3n */
int orc_combine_velocity(orc_mesh* m, const double* u_glob, double* U);/* 3n */
int orc_local_velocity_points(orc_mesh* m, int64_t np, const double* pts,
                              const int32_t* elem, double* u);
int32_t orc_locate_point(orc_mesh* m, const double* x);
int orc_exact_solution_infinite(orc_mesh* m, int64_t np, const double* pts, double* u);
/* spreading: IF[n_u*nqp] ((l*nqp + qp) as the reference), rho[n_e][nqp][3],
   Fe[n_e][3][n_u]; either output may be NULL */
int orc_spread(orc_mesh* m, int nqp_1d, int n_u, const double* IF, double* rho, double* Fe);
int orc_qp_coords(orc_mesh* m, int nqp_1d, int64_t e, double* q /* nqp*3 */);

/* ---- Chebyshev Brownian step ------------------------------------------ */
void orc_chebyshev_coefficients(int N, double da, double db, double* coef /* N+1 */);
typedef struct orc_cheb_info {
  int32_t converged, iterations, applies;
  double err1, err2;
} orc_cheb_info;
/* y = M^{-1/2} dw via the reference recurrence with the device-side mobility
   M f = U_loc(f) - coeff_1 f + f (pm_system_stokes.C:450-452 with U_glob = 0).
   mode 0: reference-compatible (estimators and stopping rule),
   mode 1: fixed order (exactly N terms, no estimator applies).
   disp (optional) = M y. */
int orc_cheb_brownian(orc_mesh* m, int N, double eig_min, double eig_max, double tol,
                      int mode, const double* dw, double* y, double* disp, orc_cheb_info* info);
int orc_mobility_apply(orc_mesh* m, const double* f, double* u);
/* dense M (3n x 3n row-major) for tiny n (test use) */
int orc_mobility_dense(orc_mesh* m, double* Mdense);

/* threads used by the threaded loops (0 = all) */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
