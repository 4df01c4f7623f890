// ggem_internal.h -- context layout and launcher declarations shared by the
// HIP translation units of libcopss_ggem.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <initializer_list>
#include <map>
#include <random>
#include <string>
#include <utility>
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

// Grid and GGEM constants passed by value to every kernel.  Elements keep
// the reference's id e = i + nx (j + ny k) at the interface; inside, the
// cell-sorted particle store numbers cells brick by brick (8 x 8 x 8 cells
// per brick, bricks and the cells of a brick in x-major order) so that the
// particles of a compact 3D region occupy one contiguous slot range.
struct Grid {
  double bmin[3];
  double L[3];
  double h[3];
  double half_L[3];
  int32_t n[3];
  int32_t periodic[3];
  int32_t nb[3];    // bricks per axis, ceil(n / 8)
  int64_t n_elem;   // n[0] n[1] n[2]
  int64_t n_cells;  // brick-padded cell index range, 512 nb[0] nb[1] nb[2]
};
constexpr int kBrickBits = 3;  // 8 cells per brick edge

struct GGEMConst {
  double eps;  // 1/(8 pi mu)
  double alpha, alpha2, alpha3_pi_23, two_alpha_sqrt_pi, coeff_1;
  double ksi, ksi2, two_ksi_sqrt_pi, coeff_2;
  int32_t point_type;
};

// One suspension of a batched context: its grid, the first global element
// of its cells and its stencil rows.  Elements are numbered system by
// system, so the slots of a system form one contiguous range.
struct SysGrid {
  Grid g;
  int64_t elem_off;  // first element id of the system
  int64_t cell_off;  // first brick-order cell index of the system
  int32_t row_off, nrows;
  int32_t reach[3];  // largest stencil offset per axis, in cells
};

// Layout of the FP32C coefficient table (ggem_util_fp32c_table): n1
// intervals of width d1 on [0, s1), then n2 of width d2, then one zero
// interval; 16 floats per interval.
struct CoefTab {
  float s1, inv1, inv2, d1, d2;
  int32_t n1, n2, n;  // n = n1 + n2 + 1 intervals
};
constexpr int kTabMaxIntervals = 128;  // the kernels' LDS copy (8 KiB)

// Neighbour entries carry the particle slot in the low 26 bits and the
// periodic image of the pair in the high 6 bits: 2 bits per axis, 1 = add
// the box length, 2 = subtract it (the single shift of point_vector).
constexpr uint32_t kSlotBits = 26;
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1u;
constexpr int64_t kMaxParticles = int64_t(1) << kSlotBits;
constexpr int kSlice = 64;  // targets per neighbour-list slice (one wave)
// packed cell coordinates of the FP32C position records: i | j << 11 | k << 22
constexpr int kCellBitsX = 11, kCellBitsY = 11, kCellBitsZ = 10;
// most particles of one brick of cells for the staged FP32C entry layout
constexpr int kBoxMaxTargets = 65536;

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

// Launch-shape autotuner: for each kernel signature (kernel, precision,
// operand count, log2 size), the fastest of a fixed candidate list of
// launch shapes, all of which compute bit-identical results.
struct Tuner {
  bool on = false;
  std::map<std::string, int> pick;
  std::map<std::string, std::vector<float>> ms;  // measured ms per candidate
};

}  // namespace ggem

struct ggem_ctx {
  ggem_params p;
  ggem_consts c;
  ggem::Grid grid;
  ggem::GGEMConst gc;
  int device = 0;
  int32_t lds_block = 0;  // LDS bytes a workgroup may use (queried on first use)
  hipStream_t stream = nullptr;
  std::string err;
  // periodic shifts are taken from the neighbour entries unless the stencil
  // reaches half a box on some periodic axis; then every pair is reduced by
  // the explicit minimum-image test
  int exact_minimg = 0;

  // stencil
  std::vector<ggem::StencilRow> rows_h;
  ggem::StencilRow* rows_d = nullptr;
  int32_t n_rows = 0;
  int32_t reach[3] = {0, 0, 0};

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
  double4* x4 = nullptr;      // [n] sorted positions, one 32-byte record per slot
  double* fs = nullptr;       // [3][n] sorted SoA forces (FP64)
  float* fs32 = nullptr;      // [3][n] sorted SoA forces (FP32)
  float4* fg32 = nullptr;     // [n] forces by slot, gather copy (FP32)
  double4* fg64 = nullptr;    // [n] forces by slot, gather copy (FP64)
  double* u_loc = nullptr;    // [3n] interleaved, particle order
  double* u_out = nullptr;    // [3n] interleaved, particle order
  int32_t* pcell = nullptr;       // [n] brick-order cell of each particle
  int32_t* cell_count = nullptr;  // [n_cells]
  int32_t* cell_start = nullptr;  // [n_cells + 1], by brick-order cell index
  int32_t* scan_tmp = nullptr;
  int64_t scan_tmp_len = 0;
  int32_t* flag = nullptr;        // device error flag
  int64_t* pairs_d = nullptr;     // pair count reduction
  int64_t pairs = -1;
  bool built = false;
  // the slot-ordered positions (xs, x4, pc32) hold the particles' current
  // positions; cleared by ggem_move_particles until a build or a
  // ggem_update_positions copies them in
  bool slots_current = false;
  // positions changed since the lists were built (ggem_update_positions,
  // PointMesh::reinit(false)): the lists' periodic shifts no longer apply,
  // every pair takes the explicit minimum image
  bool stale_pos = false;
  // GGEM_COMPAT_STALE_LISTS: some particle kept its previous list this
  // build (point_particle.h:189-194), so the lists are not stencil lists
  bool stale_lists = false;
  // GGEM_COMPAT_STALE_LISTS: each particle's list of the last build as CSR
  // by particle id in the reference's order (the list it keeps when its new
  // one is empty), and the slots that kept theirs
  int64_t* keep_off = nullptr;   // [keep_n + 1]
  int32_t* keep_ids = nullptr;
  size_t keep_off_cap = 0, keep_ids_cap = 0;
  int64_t keep_n = -1;            // particles of the kept lists (-1: none)
  int64_t* keep_off2 = nullptr;   // the next build's lists, swapped in when complete
  int32_t* keep_ids2 = nullptr;
  size_t keep_off2_cap = 0, keep_ids2_cap = 0;
  int64_t* keep_cnt = nullptr;    // [n] list lengths by particle
  size_t keep_cnt_cap = 0;
  uint8_t* stale_slot = nullptr;  // [n] 1: the slot's list is its kept one
  size_t stale_slot_cap = 0;
  bool ell_valid = false;  // the sliced lists are filled (on demand: the staged FP32C sums do not read them)

  // particle neighbour lists, sliced ELL: slice q holds targets 64q..64q+63,
  // entry k of lane l at ell[slice_off[q] + 64 k + l]
  int32_t* nbr_cnt = nullptr;     // [n] list length per slot
  int64_t* slice_len = nullptr;   // [n/64 + 1] 64 x max length, scanned into slice_off
  int64_t* slice_off = nullptr;   // [n/64 + 2]
  int64_t* scan64_tmp = nullptr;
  int64_t scan64_tmp_len = 0;
  uint32_t* ell = nullptr;
  int64_t ell_cap = 0;
  int64_t ell_len = 0;
  // pair tensors, G = A I + Bx x x^T: FP64 keeps (A, w) per list entry
  // (same layout as ell) with w = sqrt(|Bx|) x_ij, Bx = sgn |Bx|; FP32C
  // forms x_ij from the cell-relative positions pc32 and (A, Bx) from the
  // coefficient table at r^2, in every pair sum
  float4* pc32 = nullptr;     // [n] by slot: FP32 position relative to its cell centre, packed cell (i, j, k)
  float4* tab32 = nullptr;    // FP32C coefficient table, 4 float4 per interval
  ggem::CoefTab tab{};
  double4* coef64 = nullptr;
  int64_t coef64_cap = 0;
  bool coef64_valid = false;
  // FP32C neighbour records staged per brick of 8 x 8 x 8 cells (the brick
  // widened by the stencil reach), built on the first FP32C pair sum after a
  // neighbour build when the lists carry periodic shifts; a brick without
  // records gathers from global memory
  int64_t* box_cnt = nullptr;     // [n_bricks] records of each staged brick, 0 otherwise
  int64_t* box_off = nullptr;     // [n_bricks + 1] first record of each brick
  int64_t* box_ecnt = nullptr;    // [n_bricks] entries of each brick's layout
  int64_t* box_eoff = nullptr;    // [n_bricks + 1] first entry of each brick
  size_t box_brick_cap = 0;       // (one allocation at box_cnt holds the four)
  int32_t* box_tgt = nullptr;     // [n] slots of each brick, longest lists first in staged bricks
  int64_t* box_base = nullptr;    // [n] entry 0 of the target at each position of box_tgt
  size_t box_tgt_cap = 0, box_base_cap = 0;
  uint32_t* box_slot = nullptr;   // [records] slot of each record
  size_t box_rec_cap = 0;
  float4* box_pos = nullptr;      // [records] cell-relative FP32 position, packed box cell
  size_t box_pos_cap = 0;
  uint16_t* box_idx = nullptr;    // [entries] the neighbour's record number (staged bricks)
  size_t box_idx_cap = 0;
  int32_t* brick_recs = nullptr;  // [n_bricks] particles in each brick's box (list count pass)
  size_t brick_recs_cap = 0;
  int32_t* box_max_d = nullptr;   // [3] largest record count of a staged brick, most targets of a brick, unstaged bricks with targets
  uint8_t* spread_unstaged = nullptr;  // [bricks] the brick spreading's boxes over its record table
  size_t spread_unstaged_cap = 0;
  int32_t box_max = 0;            // same on the host; 0 when no brick is staged
  bool box_built = false;         // box_* describe the current lists
  bool box_fp64 = false;          // box_idx holds every pair in the lists' slice layout (FP64)
  bool box_valid = false;         // ... and some brick is staged
  bool box_unstaged = false;      // ... and some brick with targets is not
  int bsign = 1;  // sign of Bx: +1 when xi > alpha

  // element particle lists (spreading), CSR over elements
  int64_t* elist_off = nullptr;   // [n_elem + 1]
  uint32_t* elist = nullptr;
  int64_t elist_cap = 0;
  bool elist_valid = false;

  // device-resident Brownian noise (particle order, interleaved)
  double* noise = nullptr;
  int64_t noise_cap = 0;
  int32_t noise_nrhs = 0;
  // the reference's noise engine, one for the whole run (host; sequential)
  std::default_random_engine noise_gen;
  std::vector<double> noise_h;

  // Brownian vectors kept on the device (GGEM_CHEB_KEEP): y and the
  // displacement M y, particle order [nrhs][3n]
  double* keep_y = nullptr;
  double* keep_d = nullptr;
  size_t keep_cap = 0;
  int32_t keep_nrhs = 0;
  bool keep_d_valid = false;
  // positions saved for a move (R0), unwrapped positions (ROUT) and the
  // periodic image counts, particle order
  double* x_saved = nullptr;
  double* rout = nullptr;
  int32_t* images = nullptr;
  int64_t motion_cap = 0;
  bool saved_valid = false;
  bool rout_valid = false;
  bool u_out_valid = false;  // u_out holds the velocity of the last combination

  // fixed-order Chebyshev step captured as a HIP graph; the recurrence
  // reads [da, db, c_0..c_N] from dpar, so one capture serves any bounds
  bool graph_on = false;
  hipGraphExec_t cheb_exec = nullptr;
  std::vector<uintptr_t> cheb_key;  // buffers and sizes of the capture
  double* dpar = nullptr;
  size_t dpar_cap = 0;

  // batched suspensions (nsys > 1): particles, elements and slots of all
  // systems in one store; the stencil rows of system k are
  // rows_d[row_off .. row_off + nrows)
  int32_t nsys = 1;
  std::vector<ggem::SysGrid> sys_h;
  ggem::SysGrid* sys_d = nullptr;
  std::vector<int64_t> part_off;     // [nsys + 1] particle ranges
  int64_t* part_off_d = nullptr;
  int32_t* sys_part = nullptr;       // [n] system of each particle
  int32_t* sys_slot = nullptr;       // [n] system of each slot
  int64_t sys_cap = 0;
  double* par_h = nullptr;           // pinned staging of per-system recurrence parameters
  int64_t* peek_d = nullptr;         // [kPeekMax] device values gathered for one readback
  int64_t* peek_h = nullptr;         // [kPeekMax] their pinned host copy
  int32_t cheb_terms = 1 << 30;      // terms of the last reference-mode Chebyshev call (its statistics read)
  size_t par_cap = 0;
  hipEvent_t par_ev = nullptr;

  // force-field fixes: bonds (incidences by bead, bond order), constant
  // forces and discretized wall particles
  int64_t* bond_off = nullptr;  // [bond_n + 1]
  int4* bond_inc = nullptr;     // (bead 1, bead 2, 1 for bead 2's incidence, 0)
  size_t bond_off_cap = 0, bond_inc_cap = 0;
  int64_t bond_n = -1;          // particles the bonds were set for
  int2* bond_list = nullptr;    // [n_bonds] (bead 1, bead 2) in bond order
  size_t bond_list_cap = 0;
  int64_t n_bonds = 0;
  int32_t* chain_flag = nullptr;  // device flag of ggem_check_chain
  double* vext = nullptr;         // [2 + partials] velocity extrema
  double* cforce = nullptr;     // [3n] constant forces (p_constant)
  size_t cforce_cap = 0;
  int64_t cforce_n = -1;
  double* wall_xyz = nullptr;   // [3 n_wall]
  size_t wall_cap = 0;
  int64_t n_wall = 0;

  // FE mesh of the global velocity: 20 node numbers per element (HEX20)
  int32_t* fe_conn = nullptr;
  int64_t fe_nodes = -1;  // -1 until ggem_set_fe_mesh
  double* fe_u = nullptr;  // [3 fe_nodes] nodal velocity, then nodal force
  size_t fe_u_cap = 0;
  int64_t* fe_inc_off = nullptr;  // [fe_nodes + 1] incidences of each node
  int32_t* fe_inc = nullptr;      // [20 n_elem] 20 e + l, by node, (e, l) ascending
  int32_t* fe_node_order = nullptr;  // [fe_nodes] nodes by the brick of their first element
  // periodic condensation: a min-face node sums the incidences of all its
  // images ((e, l) ascending), a max-face node of a periodic axis has none
  int64_t* fe_cinc_off = nullptr;  // [fe_nodes + 1]
  int32_t* fe_cinc = nullptr;      // [20 n_elem]

  // wall and inlet items of the right-hand side (element, l | masks), for
  // the options key rhs_key (-1: none built)
  int2* rhs_items = nullptr;
  size_t rhs_items_cap = 0;
  int64_t rhs_m = 0;
  int32_t rhs_key = -1;
  double* rhs_tab = nullptr;  // [6][4][20] JxW_face phi_l of every side's face points
  double* rhs_buf = nullptr;  // [9 rhs_m] item points, local and unbounded velocities
  size_t rhs_buf_cap = 0;
  int32_t* rhs_el = nullptr;  // [rhs_m] element of each item
  size_t rhs_el_cap = 0;

  // generic scratch (grown on demand)
  void* scratch = nullptr;

  ggem::Profiler prof;
  ggem::Tuner tune;
};

namespace ggem {

// Every pair takes the explicit minimum-image test instead of the lists'
// periodic shifts: the stencil reaches half a box, the positions moved
// since the build, or some lists are kept ones.
inline bool explicit_minimg(const ggem_ctx* c) { return c->exact_minimg || c->stale_pos || c->stale_lists; }

// error helpers
int fail(ggem_ctx* ctx, int code, const std::string& msg);
int hip_fail(ggem_ctx* ctx, hipError_t e, const char* what);
#define GGEM_HIP(ctx, expr)                                      \
  do {                                                           \
    hipError_t _e = (expr);                                      \
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

// Candidate index for `key`: the recorded pick, else (autotuning on, stream
// not capturing) each candidate is launched once to warm up and three times
// under events and the fastest is kept; dflt (the default shape) otherwise.
// launch(c) must be idempotent; a candidate whose launch fails is skipped.
int tune_pick(ggem_ctx* ctx, const std::string& key, int ncand, const std::function<hipError_t(int)>& launch,
              int dflt = 0);
// The recorded pick without tuning (dflt when none or autotuning is off).
int tune_lookup(const ggem_ctx* ctx, const std::string& key, int dflt = 0);
// "name/precision/.../log2 n" signature of a launch
std::string tune_key(const char* kernel, int precision, int extra, int64_t n);

// ---- launchers (kernels_build.hip) ----
hipError_t launch_bin(ggem_ctx* ctx, bool hysteresis);
hipError_t launch_scan(ggem_ctx* ctx);
hipError_t launch_place(ggem_ctx* ctx);
hipError_t launch_gather_forces(ggem_ctx* ctx);
// per-slot counts, per-slice lengths and the total pair count
hipError_t launch_count_lists(ggem_ctx* ctx);
// exclusive scan of int64 values; out[len] receives the total
hipError_t launch_scan64(ggem_ctx* ctx, const int64_t* in, int64_t len, int64_t* out);
hipError_t launch_fill_ell(ggem_ctx* ctx);
// batched contexts: system of each particle from part_off_d
hipError_t launch_part_sys(ggem_ctx* ctx);
hipError_t launch_elist_count(ggem_ctx* ctx);
hipError_t launch_elist_fill(ggem_ctx* ctx);
hipError_t launch_count_pairs(ggem_ctx* ctx, int64_t* counts_d);
hipError_t launch_fill_csr(ggem_ctx* ctx, const int64_t* offsets_d, int32_t* ids_d);
hipError_t launch_elem_point_csr(ggem_ctx* ctx, int64_t e0, int64_t ne, int64_t* offsets_d,
                                 int32_t* ids_d, int pass);
hipError_t launch_locate(ggem_ctx* ctx, int64_t m, const double* pts, int32_t* elem);
// slot positions and FP32C records from x_in, lists kept (PointMesh::reinit(false))
hipError_t launch_update_slots(ggem_ctx* ctx);
// GGEM_COMPAT_STALE_LISTS: slots with an empty new list take their kept
// one's length (stale_slot, flag), then the slice lengths and pair total
hipError_t launch_stale_patch(ggem_ctx* ctx, bool have_kept);
hipError_t launch_counts_by_particle(ggem_ctx* ctx, int64_t* out);
// every particle's effective list (kept or new) in the reference's order
hipError_t launch_fill_keep(ggem_ctx* ctx, const int64_t* old_off, const int32_t* old_ids, const int64_t* off,
                            int32_t* ids);

// ---- launchers (kernels_fe.hip) ----
// out[3p + d] = HEX20 interpolant of node_u[3 node + d] at particle p in its element
hipError_t launch_interp_hex20(ggem_ctx* ctx, const int32_t* conn, const double* node_u, double* out);
// out[3 node + d] = sum over the node's incidences (e, l) of Fe[e][l][d] (n_u = 20);
// condensed: over the incidences of the node's periodic images (none for a
// max-face node of a periodic axis)
hipError_t launch_gather_nodes(ggem_ctx* ctx, const double* Fe, double* out, bool condensed = false);
// f_in = FENE forces along chains of `chain` consecutive particles (chain > 1)
// + LJ forces of the walls of the non-periodic axes (walls != 0), from x_in
hipError_t launch_fix_forces(ggem_ctx* ctx, int32_t chain, double k, double b, int32_t walls, double eps,
                             double sigma, double rcut);

// ---- launchers (kernels_pair.hip) ----
// A vector operand of the pair sum: FP64 SoA by slot for the element-wise
// algebra plus a one-record-per-slot copy for the neighbour gathers.
struct SlotVec {
  double* soa;    // [3][n]
  float4* g32;    // [n]
  double4* g64;   // [n]
};

// Epilogue of the pair sum, per target slot s with sum m_s = sum_t G f_t:
enum Epi : int {
  EPI_LOCAL = 0,     // out[3p+c] = U_loc  (p = particle id, interleaved)
  EPI_MOBILITY = 1,  // out[3p+c] = U_glob + U_loc - coeff_1 f + f (type rules)
  EPI_CHEB = 2,      // Chebyshev step on slot vectors (ChebEpi)
};
struct ChebEpi {
  int kind;            // 0: v = M x; 1: v = db x + da M x; 2: v = 2db x + 2da M x - x_{k-1}, y += c v
  double da, db;
  double c;            // coefficient of v (kind 2)
  SlotVec xm1;         // x_{k-1} (kind 2); may alias v
  double* y;           // [nrhs][3][n] (kind 2)
  // Device-side parameters (used when dpar != nullptr, e.g. in a captured
  // graph): per system [da, db, c_0 .. c_N] with stride pstride; term is
  // the index of v's coefficient; sys gives the system of each slot
  // (nullptr: system 0).
  const double* dpar = nullptr;
  int pstride = 0;
  int term = 0;
  const int32_t* sys = nullptr;
  // kind 2: terms added to y, bit 0 x_{k-1}, bit 1 x_k, bit 2 v = x_{k+1}
  // (0: y untouched), with coefficients cm1, c0 and c (dpar: the terms
  // before term)
  int ymask = 4;
  double cm1 = 0.0, c0 = 0.0;
  // non-null: the launch does nothing while *skip is set (the reference
  // mode's recurrence after its stopping rule fired on the device)
  const int32_t* skip = nullptr;
};
// pair tensors of every list entry for the given precision
hipError_t launch_pair_tensors(ggem_ctx* ctx, int precision);
// FP32C staged records: box_cnt per brick (and box_max_d by a reduction)
hipError_t launch_box_count(ggem_ctx* ctx);
// records, target layout and record numbers of every brick, with box_off
// and box_eoff scanned from the counts
hipError_t launch_box_fill(ggem_ctx* ctx, int32_t max_records, int32_t max_targets);
// whether the context's FP32C lists can be staged per brick
bool box_supported(const ggem_ctx* ctx);
hipError_t launch_pair_beads(ggem_ctx* ctx, int precision, int epi, const double* u_glob,
                             double* out);
hipError_t launch_pair_points(ggem_ctx* ctx, int precision, int64_t m, const double* pts,
                              const int32_t* elem, double* u);
// v = M x (+ epilogue) for nrhs slot vectors (x, v: nrhs consecutive SlotVecs
// of stride 3n / n)
hipError_t launch_apply(ggem_ctx* ctx, int precision, int nrhs, const SlotVec& x, const SlotVec& v,
                        const ChebEpi& epi);

// ---- launchers (kernels_spread.hip) ----
// Fe[e][j][l] (node_major: Fe[e][l][j])
// Small device values (int32 or int64 each) read back in one kernel, one
// copy into pinned host memory and one synchronisation; out gets them as
// int64.
constexpr int kPeekMax = 8;
struct PeekArgs {
  const void* src[kPeekMax];
  int32_t bytes[kPeekMax];  // 4 or 8
  int32_t n;
  int64_t* dst;
};
hipError_t launch_peek(ggem_ctx* ctx, const PeekArgs& p);
int peek(ggem_ctx* ctx, std::initializer_list<std::pair<const void*, int32_t>> items, int64_t* out);

hipError_t launch_spread(ggem_ctx* ctx, int precision, int nqp_1d, int n_u, const double* IF_d,
                         double* rho, double* Fe, bool node_major = false);
// FP32C spreading by brick (no element lists): 0 = it does not apply (box
// over the packing or the LDS), 1 = an autotuner candidate, 2 = also the
// default of launch_spread
int spread_brick_mode(ggem_ctx* ctx, int precision, int nqp_1d);

// ---- launchers (kernels_vec.hip) ----
hipError_t launch_permute_in(ggem_ctx* ctx, int nrhs, const double* in_interleaved, double* out_slots);
hipError_t launch_permute_out(ggem_ctx* ctx, int nrhs, const double* in_slots, double* out_interleaved);
// gather copies of nrhs slot SoA vectors
hipError_t launch_make_gather(ggem_ctx* ctx, int nrhs, const double* soa, float4* g32, double4* g64);
// dpar[off .. off+cnt) = vals (kernel arguments, so safe inside a stream or graph)
hipError_t launch_set_params(ggem_ctx* ctx, double* dpar, int64_t off, const double* vals, int cnt);
// y = c0 x0 + c1 x1 over nrhs slot vectors (device parameters when dpar != nullptr)
hipError_t launch_cheb_init_y(ggem_ctx* ctx, int nrhs, const double* x0, const double* x1, double* y,
                              double c0, double c1, const double* dpar, int pstride, const int32_t* sys);
// out[r] = sum a_r * b_r over [nrhs][3n] blocks (deterministic)
hipError_t launch_dot(ggem_ctx* ctx, int nrhs, const double* a, const double* b, double* out_d,
                      double* partials);
hipError_t launch_axpby(ggem_ctx* ctx, int64_t len, double a, const double* x, double b, double* y);
// partials of one dot product of 3n slot vectors (512 doubles) or of two
// (1024), and the reference mode's per-term estimators and stopping rule
// over them (state st[8] and stop flag per right-hand side)
hipError_t launch_dot_partial(ggem_ctx* ctx, const double* a, const double* b, double* partials);
hipError_t launch_dot2_partial(ggem_ctx* ctx, const double* a1, const double* b1, const double* a2, const double* b2,
                               double* partials);
hipError_t launch_cheb_ref_init(ggem_ctx* ctx, double* st, int32_t* stop);
hipError_t launch_cheb_ref_term(ggem_ctx* ctx, double* st, int32_t* stop, const double* px, const double* pmy,
                                double xabs, double tol, int term, int check);
// batched contexts: out_d[k] = a . b over system k's slots ([3][n] vectors)
hipError_t launch_seg_dot(ggem_ctx* ctx, const double* a, const double* b, double* out_d);
// w -= ca[k] x + cb[k] y, out = sc[k] x on the slots of system k
hipError_t launch_seg_axpy2(ggem_ctx* ctx, const double* ca, const double* x, const double* cb, const double* y,
                            double* w);
hipError_t launch_seg_scale(ggem_ctx* ctx, const double* sc, const double* x, double* out);
hipError_t launch_fill(ggem_ctx* ctx, int64_t len, double v, double* y);

}  // namespace ggem
