// ggem_internal.h -- context layout and launcher declarations shared by the
// HIP translation units of libcopss_ggem.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "copss_ggem.h"

namespace ggem {

// One row of the element stencil: cells (i+di, j+dj, k+dk) for di in
// [di_lo, di_hi], offsets taken as minimum-image representatives on periodic
// axes.  The rows cover every element whose centre lies strictly inside R_e
// of the target element's centre, the target itself included (the reference
// adds the own element first, point_mesh.C:688-692).
struct StencilRow {
  int32_t dj, dk, di_lo, di_hi;
};

// Grid and GGEM constants passed by value to every kernel.
struct Grid {
  double bmin[3];
  double L[3];
  double h[3];
  double half_L[3];
  int32_t n[3];
  int32_t periodic[3];
  int64_t n_elem;
};

struct GGEMConst {
  double eps;              // 1/(8 pi mu)
  double alpha, alpha2, alpha3_pi_23, two_alpha_sqrt_pi, coeff_1;
  double ksi, ksi2, two_ksi_sqrt_pi, coeff_2;
  int32_t point_type;
};

// Device pointers of the particle store in cell order ("slots").
struct Store {
  const double* xs;   // [3][n] SoA positions by slot
  const int32_t* cell_start;  // [n_elem + 1]
  const int32_t* slot_cell;   // [n] element of each slot
  const int32_t* perm;        // slot -> particle id
  int64_t n;
};

struct Profiler {
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  bool on = false;
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  struct Acc {
    std::string name;
    double ms;
    int64_t launches;
  };
  std::vector<Acc> acc;
};

}  // namespace ggem

struct ggem_ctx {
  ggem_params p;
  ggem_consts c;
  ggem::Grid grid;
  ggem::GGEMConst gc;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // stencil
  std::vector<ggem::StencilRow> rows_h;
  ggem::StencilRow* rows_d = nullptr;
  int32_t n_rows = 0;

  // particle store
  int64_t n = 0;
  int64_t cap = 0;
  double* x_in = nullptr;     // [3n] interleaved, particle order
  double* f_in = nullptr;     // [3n] interleaved, particle order
  int32_t* elem = nullptr;    // [n] element of each particle
  bool elem_valid = false;
  int32_t* perm = nullptr;    // [n] slot -> particle
  int32_t* slot_of = nullptr; // [n] particle -> slot
  int32_t* slot_cell = nullptr;
  double* xs = nullptr;       // [3][n] sorted SoA positions
  double* fs = nullptr;       // [3][n] sorted SoA forces (FP64)
  float* fs32 = nullptr;      // [3][n] sorted SoA forces (FP32)
  double* u_loc = nullptr;    // [3n] interleaved, particle order
  double* u_out = nullptr;    // [3n] interleaved, particle order
  int32_t* cell_count = nullptr;  // [n_elem]
  int32_t* cell_start = nullptr;  // [n_elem + 1]
  int32_t* scan_tmp = nullptr;
  int64_t scan_tmp_len = 0;
  int32_t* flag = nullptr;        // device error flag
  int64_t* pairs_d = nullptr;     // pair count reduction
  int64_t pairs = -1;
  bool built = false;

  // device-resident Brownian noise (particle order, interleaved)
  double* noise = nullptr;
  int64_t noise_cap = 0;
  int32_t noise_nrhs = 0;

  // generic scratch (grown on demand)
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  double* qscratch = nullptr;    // query points etc.
  size_t qscratch_bytes = 0;

  ggem::Profiler prof;
};

namespace ggem {

// error helpers
int fail(ggem_ctx* ctx, int code, const std::string& msg);
int hip_fail(ggem_ctx* ctx, hipError_t e, const char* what);
#define GGEM_HIP(ctx, expr)                                   \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) return ggem::hip_fail(ctx, _e, #expr); \
  } while (0)

// profiling brackets around a launch
void prof_begin(ggem_ctx* ctx, const char* name, hipEvent_t* a, hipEvent_t* b);
void prof_end(ggem_ctx* ctx, const char* name, hipEvent_t a, hipEvent_t b);

struct ProfScope {
  ggem_ctx* ctx;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(ggem_ctx* c, const char* n) : ctx(c), name(n) { prof_begin(ctx, name, &a, &b); }
  ~ProfScope() { prof_end(ctx, name, a, b); }
};

Store store_of(const ggem_ctx* ctx);

// ---- launchers (kernels_build.hip) ----
hipError_t launch_bin(ggem_ctx* ctx, bool hysteresis);
hipError_t launch_scan(ggem_ctx* ctx);
hipError_t launch_place(ggem_ctx* ctx);
hipError_t launch_gather_forces(ggem_ctx* ctx, const double* f_interleaved, double* fs, float* fs32,
                                int nrhs, int64_t stride_in, int64_t stride_out);
hipError_t launch_count_pairs(ggem_ctx* ctx, int64_t* out_d, int64_t* counts_d);
hipError_t launch_fill_csr(ggem_ctx* ctx, const int64_t* offsets_d, int32_t* ids_d);
hipError_t launch_elem_point_csr(ggem_ctx* ctx, int64_t e0, int64_t ne, int64_t* offsets_d,
                                 int32_t* ids_d, int pass);
hipError_t launch_locate(ggem_ctx* ctx, int64_t m, const double* pts, int32_t* elem);

// ---- launchers (kernels_pair.hip) ----
// Epilogue of the bead pair sum:
enum Epi : int {
  EPI_LOCAL = 0,    // out = U_loc (particle order, interleaved)
  EPI_MOBILITY = 1, // out = U_loc - coeff_1 f + f (+ type rules)
};
hipError_t launch_pair_beads(ggem_ctx* ctx, int precision, int epi, const double* u_glob,
                             double* out);
hipError_t launch_pair_points(ggem_ctx* ctx, int precision, int64_t m, const double* pts,
                              const int32_t* elem, double* u);

// Chebyshev-facing operator on slot-ordered SoA vectors: for r < nrhs,
// v = M x[r] (x, v: [nrhs][3][n] by slot), then the recurrence epilogue.
struct ChebEpi {
  int kind;            // 0: v = M x; 1: v = da M x + db x; 2: v = 2da M x + 2db x - xm1, y += c v
  double da, db;
  const double* coef;  // device pointer to the coefficient used by kind 2 (may be null)
  double c_host;
  const double* xm1;   // [nrhs][3][n] (kind 2)
  double* y;           // [nrhs][3][n] (kind 2)
};
hipError_t launch_apply_slots(ggem_ctx* ctx, int precision, int nrhs, const double* x, double* v,
                              const ChebEpi& epi);

// ---- launchers (kernels_spread.hip) ----
hipError_t launch_spread(ggem_ctx* ctx, int precision, int nqp_1d, int n_u, const double* IF_d,
                         double* rho, double* Fe);

// ---- launchers (kernels_vec.hip) ----
hipError_t launch_permute_in(ggem_ctx* ctx, int nrhs, const double* in_interleaved, double* out_slots);
hipError_t launch_permute_out(ggem_ctx* ctx, int nrhs, const double* in_slots, double* out_interleaved);
// partial dot products over [nrhs][3n] blocks: out[r] = sum a_r * b_r (deterministic)
hipError_t launch_dot(ggem_ctx* ctx, int nrhs, const double* a, const double* b, double* out_d,
                      double* partials);
hipError_t launch_axpby(ggem_ctx* ctx, int64_t len, double a, const double* x, double b, double* y);
hipError_t launch_fill(ggem_ctx* ctx, int64_t len, double v, double* y);

}  // namespace ggem
