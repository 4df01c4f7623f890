// kernels_build.hip -- (a) the neighbour build: element binning, counting
// sort of the particles into cell order (stable by particle id), the
// per-slot list lengths and sliced lists and the element lists of the
// spreading (built brick by brick from LDS-staged cell starts), and the CSR
// emission of the reference's lists for export.
//
// Reference: PointMesh<3>::reinit (point_mesh.C:245-289) with
// build_elem_point_containing_list (:573-648), build_elem_point_neighbor_list
// (:652-799) and build_point_point_neighbor_list (:803-848).  The element
// stencil replaces build_elem_elem_neighbor_list (:293-443); it is computed
// on the host once per context.
#include "ggem_device.h"

namespace ggem {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(int64_t n, int b = kBlock) { return unsigned((n + b - 1) / b); }

// Canonical element of a point: per axis floor((x - min)/h) clamped to
// [0, n-1]; -1 outside [min, max].
__device__ __forceinline__ int64_t locate_dev(const Grid& g, double x, double y, double z) {
  const double v[3] = {x, y, z};
  int64_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!(v[d] >= g.bmin[d]) || !(v[d] <= g.bmin[d] + g.L[d])) return -1;
    int64_t q = int64_t(floor((v[d] - g.bmin[d]) / g.h[d]));
    q = q < 0 ? 0 : q;
    q = q > g.n[d] - 1 ? g.n[d] - 1 : q;
    c[d] = q;
  }
  return c[0] + int64_t(g.n[0]) * (c[1] + int64_t(g.n[1]) * c[2]);
}

// brick-order cell index of reference element e
__device__ __forceinline__ int64_t elem_cell(const Grid& g, int64_t e) {
  int i, j, k;
  elem_ijk(g, e, i, j, k);
  return cell_index(g, i, j, k);
}

// Elem::contains_point of an axis-aligned hex with libMesh's 1e-6 tolerance
// on the reference element.
__device__ __forceinline__ bool contains_dev(const Grid& g, int64_t e, double x, double y,
                                             double z) {
  int ci, cj, ck;
  elem_ijk(g, e, ci, cj, ck);
  const double v[3] = {x, y, z};
  const int ijk[3] = {ci, cj, ck};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double cen = g.bmin[d] + (double(ijk[d]) + 0.5) * g.h[d];
    const double xi = 2.0 * (v[d] - cen) / g.h[d];
    if (!(fabs(xi) <= 1.0 + 1e-6)) return false;
  }
  return true;
}

// Grid, stencil rows and cell starts seen by one particle or slot: the
// context's own, or those of its system in a batched context (cell starts
// shifted to the system's first element, so element ids stay local).
struct SysView {
  Grid g;
  const StencilRow* rows;
  int nrows;
  const int32_t* cs;  // cell starts from the system's first cell
  int64_t eoff;       // first element id of the system
  int64_t coff;       // first cell index of the system
};

template <bool BATCH>
__device__ __forceinline__ SysView sys_view(const Grid& g, const StencilRow* rows, int nrows,
                                            const int32_t* cs, const SysGrid* sys, const int32_t* sys_of,
                                            int64_t i) {
  if constexpr (BATCH) {
    const SysGrid& S = sys[sys_of[i]];
    return SysView{S.g, rows + S.row_off, S.nrows, cs + S.cell_off, S.elem_off, S.cell_off};
  } else {
    return SysView{g, rows, nrows, cs, 0, 0};
  }
}

template <bool BATCH>
__global__ void k_bin(Grid g, const SysGrid* __restrict__ sys, const int32_t* __restrict__ sys_part,
                      int64_t n, const double* __restrict__ x, int32_t* __restrict__ elem, int use_prev,
                      int32_t* __restrict__ pcell, int32_t* __restrict__ count, int32_t* __restrict__ flag) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SysView v = sys_view<BATCH>(g, nullptr, 0, nullptr, sys, sys_part, i);
  const double px = x[3 * i], py = x[3 * i + 1], pz = x[3 * i + 2];
  int64_t e = -1;
  if (use_prev) {
    const int64_t old = int64_t(elem[i]) - v.eoff;
    if (old >= 0 && old < v.g.n_elem && contains_dev(v.g, old, px, py, pz)) e = old;
  }
  if (e < 0) e = locate_dev(v.g, px, py, pz);
  if (e < 0) {
    atomicOr(flag, 1);
    e = 0;  // keep later kernels in bounds; the host reports the error
  }
  const int64_t c = v.coff + elem_cell(v.g, e);
  elem[i] = int32_t(v.eoff + e);
  pcell[i] = int32_t(c);
  atomicAdd(&count[c], 1);
}

// system of each slot (batched contexts)
__global__ void k_slot_sys(int64_t n, const int32_t* __restrict__ perm, const int32_t* __restrict__ sys_part,
                           int32_t* __restrict__ sys_slot) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s < n) sys_slot[s] = sys_part[perm[s]];
}

// FP32C position records: each particle's FP64 offset from its cell centre
// rounded to FP32, the cell's (i, j, k) packed into the fourth word.  A
// pair displacement is then (cell difference) h + (offset difference), all
// terms small, in FP32.
template <bool BATCH>
__global__ void k_cell_rel(Grid g, const SysGrid* __restrict__ sys, const int32_t* __restrict__ sys_slot,
                           int64_t n, const int32_t* __restrict__ slot_cell, const double4* __restrict__ x4,
                           float4* __restrict__ pc) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const SysView v = sys_view<BATCH>(g, nullptr, 0, nullptr, sys, sys_slot, s);
  int ci, cj, ck;
  cell_ijk(v.g, slot_cell[s] - v.coff, ci, cj, ck);
  const double4 p = x4[s];
  const double cx = v.g.bmin[0] + (double(ci) + 0.5) * v.g.h[0];
  const double cy = v.g.bmin[1] + (double(cj) + 0.5) * v.g.h[1];
  const double cz = v.g.bmin[2] + (double(ck) + 0.5) * v.g.h[2];
  const uint32_t packed = uint32_t(ci) | (uint32_t(cj) << kCellBitsX) | (uint32_t(ck) << (kCellBitsX + kCellBitsY));
  pc[s] = make_float4(float(p.x - cx), float(p.y - cy), float(p.z - cz), __uint_as_float(packed));
}

// system of each particle from the particle ranges (binary search)
__global__ void k_part_sys(int64_t n, int nsys, const int64_t* __restrict__ part_off,
                           int32_t* __restrict__ sys_part) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {  // last k with part_off[k] <= i
    const int mid = (lo + hi + 1) >> 1;
    if (part_off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  sys_part[i] = lo;
}

// Exclusive scan of int32 counts in three phases: per-block scan with block
// totals, a single-block scan of the totals, and the carry-in add.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* sh, int32_t& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int32_t add = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  total = sh[kScanThreads - 1];
  const int32_t incl = sh[t];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan_tiles(const int32_t* __restrict__ in, int64_t len, int32_t* __restrict__ out,
                             int32_t* __restrict__ tile_sums) {
  __shared__ int32_t sh[kScanThreads];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  int32_t v[kScanItems];
  int32_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    v[q] = (base + q < len) ? in[base + q] : 0;
    s += v[q];
  }
  int32_t total;
  int32_t run = block_exclusive_scan(s, sh, total);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    if (base + q < len) out[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void k_scan_sums(int32_t* __restrict__ sums, int64_t ntiles) {
  __shared__ int32_t sh[kScanThreads];
  int32_t carry = 0;
  for (int64_t c0 = 0; c0 < ntiles; c0 += kScanThreads) {
    const int64_t i = c0 + threadIdx.x;
    const int32_t v = (i < ntiles) ? sums[i] : 0;
    int32_t total;
    const int32_t ex = block_exclusive_scan(v, sh, total);
    if (i < ntiles) sums[i] = ex + carry;
    carry += total;
    __syncthreads();
  }
}

__global__ void k_scan_add(int32_t* __restrict__ out, int64_t len, const int32_t* __restrict__ sums,
                           int32_t total_n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len) out[i] += sums[i / kScanTile];
  if (i == 0) out[len] = total_n;
}

// Place every particle into its cell's slot range (arbitrary order within
// the cell), then sort each cell by particle id and fill the slot arrays.
__global__ void k_place(int64_t n, const int32_t* __restrict__ pcell,
                        const int32_t* __restrict__ cell_start, int32_t* __restrict__ cursor,
                        int32_t* __restrict__ perm) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t c = pcell[i];
  const int32_t s = cell_start[c] + atomicAdd(&cursor[c], 1);
  perm[s] = int32_t(i);
}

// e runs over brick-order cell indices
__global__ void k_sort_cells(int64_t n_cells, int64_t n, const int32_t* __restrict__ cell_start,
                             int32_t* __restrict__ perm, int32_t* __restrict__ slot_of,
                             int32_t* __restrict__ slot_cell, const double* __restrict__ x,
                             double* __restrict__ xs, double4* __restrict__ x4) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_cells) return;
  const int32_t a = cell_start[e], b = cell_start[e + 1];
  for (int32_t s = a + 1; s < b; ++s) {  // insertion sort by particle id
    const int32_t v = perm[s];
    int32_t t = s - 1;
    while (t >= a && perm[t] > v) {
      perm[t + 1] = perm[t];
      --t;
    }
    perm[t + 1] = v;
  }
  for (int32_t s = a; s < b; ++s) {
    const int32_t p = perm[s];
    slot_of[p] = s;
    slot_cell[s] = int32_t(e);
    const double px = x[3 * int64_t(p)], py = x[3 * int64_t(p) + 1], pz = x[3 * int64_t(p) + 2];
    xs[s] = px;
    xs[n + s] = py;
    xs[2 * n + s] = pz;
    x4[s] = make_double4(px, py, pz, 0.0);
  }
}

// Forces (interleaved, particle order) -> slot SoA (FP64, FP32) and the
// per-slot gather copies.
__global__ void k_gather_forces(int64_t n, const int32_t* __restrict__ perm,
                                const double* __restrict__ f, double* __restrict__ fs,
                                float* __restrict__ fs32, float4* __restrict__ fg32,
                                double4* __restrict__ fg64) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t p = perm[s];
  const double fx = f[3 * p], fy = f[3 * p + 1], fz = f[3 * p + 2];
  fs[s] = fx;
  fs[n + s] = fy;
  fs[2 * n + s] = fz;
  fs32[s] = float(fx);
  fs32[n + s] = float(fy);
  fs32[2 * n + s] = float(fz);
  fg32[s] = make_float4(float(fx), float(fy), float(fz), 0.f);
  fg64[s] = make_double4(fx, fy, fz, 0.0);
}

// ---- per-brick list builds: the brick's box cells staged in LDS ----
//
// One workgroup per brick of 8 x 8 x 8 cells.  The cell starts of the
// brick's box (the brick widened by the stencil reach) go to LDS once:
// per box cell the first slot of its grid cell and the particles of the
// box cells before it.  Each target (slot or element) of the brick then
// walks its stencil rows through LDS (box_walk), in for_each_segment's
// order, so the lists and counts equal the per-slot walks over the global
// cell starts, which serve bricks whose box does not fit the LDS tables.
constexpr int kBrickThreads = 256;

struct BrickArgs {
  Grid g;
  const StencilRow* rows;
  int nrows;
  int32_t reach[3];
  const int32_t* cs;         // cell starts, brick order, all systems
  const int32_t* slot_cell;  // brick-order cell of each slot
  const SysGrid* sys;        // batched contexts (null: one system)
  const int32_t* sys_slot;
  int32_t cells_cap;         // box cells the LDS tables hold
  int32_t* recs;             // [n_bricks] particles in the box of each brick the count pass staged
  // slots whose list is the one their particle kept (GGEM_COMPAT_STALE_LISTS)
  const uint8_t* stale;      // [n] or null
  const int64_t* keep_off;   // kept lists by particle (reference order)
  const int32_t* keep_ids;
  const int32_t* slot_of;
  const int32_t* perm;
};

// grid, rows and box of the brick whose first slot is sb
template <bool BATCH>
__device__ __forceinline__ SysView brick_view(const BrickArgs& a, int64_t brick, int64_t sb, BoxGeom& B) {
  const SysView v = sys_view<BATCH>(a.g, a.rows, a.nrows, a.cs, a.sys, a.sys_slot, sb);
  if constexpr (BATCH) B.init(v.g, v.coff, a.sys[a.sys_slot[sb]].reach, brick);
  else B.init(v.g, 0, a.reach, brick);
  return v;
}

__device__ __forceinline__ bool brick_staged(const BrickArgs& a, const BoxGeom& B) {
  return B.nbc <= a.cells_cap && B.fits();
}

// Slot of every record of the box (record = particle of a box cell, in box
// order) after stage_box_cells: a stencil run of box cells q0..q1 is then
// records cst[q0] .. cst[q1 + 1] - 1.  False (nothing written) when the
// box holds more than cap records.  Uniform over the block; ends with a
// barrier when it writes.
constexpr int kBrickRecCap = 2048;

__device__ __forceinline__ bool stage_box_slots(const BoxGeom& B, const int32_t* first, const int32_t* cst,
                                                int32_t* rslot, int cap) {
  if (cst[B.nbc] > cap) return false;
  const int per = (B.nbc + kBrickThreads - 1) / kBrickThreads;
  const int q0 = min(B.nbc, int(threadIdx.x) * per), q1 = min(B.nbc, q0 + per);
  for (int q = q0; q < q1; ++q) {
    const int f = first[q], r0 = cst[q];
    for (int r = r0, e = cst[q + 1]; r < e; ++r) rslot[r] = f + (r - r0);
  }
  __syncthreads();
  return true;
}

// List lengths of the brick's slots (own element + stencil, minus itself).
template <bool BATCH>
__global__ void __launch_bounds__(kBrickThreads) k_brick_count(BrickArgs a, int32_t* __restrict__ cnt) {
  extern __shared__ int32_t cst[];  // [cells_cap + 1]
  __shared__ int32_t s_w[kBrickThreads / 64];
  const int64_t brick = blockIdx.x;
  const int64_t c0 = brick * kBrickCells;
  const int32_t sb = a.cs[c0], se = a.cs[c0 + kBrickCells];
  if (sb >= se) return;  // uniform over the block
  BoxGeom B;
  const SysView v = brick_view<BATCH>(a, brick, sb, B);
  if (brick_staged(a, B)) {
    stage_box_cells<kBrickThreads>(B, a.cs, nullptr, cst, s_w);
    if (threadIdx.x == 0) a.recs[brick] = cst[B.nbc];  // (the staged-record census reuses it)
    for (int32_t s = sb + int32_t(threadIdx.x); s < se; s += kBrickThreads) {
      int ci, cj, ck;
      cell_ijk(v.g, a.slot_cell[s] - v.coff, ci, cj, ck);
      int c = -1;
      box_walk(B, v.rows, v.nrows, ci, cj, ck, [&](int q0, int q1, uint32_t) { c += cst[q1 + 1] - cst[q0]; });
      cnt[s] = c;
    }
  } else {
    for (int32_t s = sb + int32_t(threadIdx.x); s < se; s += kBrickThreads) {
      int ci, cj, ck;
      cell_ijk(v.g, a.slot_cell[s] - v.coff, ci, cj, ck);
      int c = -1;
      for_each_segment(v.g, v.rows, v.nrows, v.cs, ci, cj, ck, [&](int32_t p, int32_t q, uint32_t) { c += q - p; });
      cnt[s] = c;
    }
  }
}

// Slice lengths (64 x the longest list of each wave of 64 slots) and the
// pair total from the list lengths (one atomic per block of 1024).
constexpr int kSliceBlock = 1024;
__global__ void __launch_bounds__(kSliceBlock) k_slice_len(const int32_t* __restrict__ cnt, int64_t n,
                                                           int64_t* __restrict__ slice_len,
                                                           unsigned long long* __restrict__ total) {
  __shared__ int64_t s_sum[kSliceBlock / 64];
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t c = s < n ? cnt[s] : 0;
  int64_t mx = c, sum = c;
  for (int off = 32; off > 0; off >>= 1) {
    mx = max(mx, (int64_t)__shfl_down(mx, off, 64));
    sum += __shfl_down(sum, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    const int64_t q = s >> 6;
    if (q * 64 < n) slice_len[q] = mx * 64;
    s_sum[threadIdx.x >> 6] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t b = 0;
    for (int w = 0; w < kSliceBlock / 64; ++w) b += s_sum[w];
    if (b) atomicAdd(total, (unsigned long long)b);
  }
}

// The lists of the brick's slots in the sliced layout.
template <bool BATCH>
__global__ void __launch_bounds__(kBrickThreads) k_brick_fill_ell(BrickArgs a, const int64_t* __restrict__ slice_off,
                                                                  uint32_t* __restrict__ ell) {
  extern __shared__ int32_t lds[];
  __shared__ int32_t s_w[kBrickThreads / 64];
  int32_t* first = lds;                  // [cells_cap]
  int32_t* cst = lds + a.cells_cap;      // [cells_cap + 1]
  int32_t* rslot = cst + a.cells_cap + 1;  // [kBrickRecCap]
  const int64_t brick = blockIdx.x;
  const int64_t c0 = brick * kBrickCells;
  const int32_t sb = a.cs[c0], se = a.cs[c0 + kBrickCells];
  if (sb >= se) return;
  BoxGeom B;
  const SysView v = brick_view<BATCH>(a, brick, sb, B);
  const bool staged = brick_staged(a, B);
  bool recs = false;
  if (staged) {
    stage_box_cells<kBrickThreads>(B, a.cs, first, cst, s_w);
    recs = stage_box_slots(B, first, cst, rslot, kBrickRecCap);
  }
  for (int32_t s = sb + int32_t(threadIdx.x); s < se; s += kBrickThreads) {
    int ci, cj, ck;
    cell_ijk(v.g, a.slot_cell[s] - v.coff, ci, cj, ck);
    uint32_t* out = ell + slice_off[s >> 6] + (s & 63);
    int64_t k = 0;
    if (a.stale && a.stale[s]) {  // the particle's kept list, slots by the current order
      const int32_t p = a.perm[s];
      for (int64_t q = a.keep_off[p], e = a.keep_off[p + 1]; q < e; ++q, ++k)
        out[k * kSlice] = uint32_t(a.slot_of[a.keep_ids[q]]);
      continue;
    }
    auto emit = [&](int32_t p, int32_t q, uint32_t code) {
      const uint32_t hi = code << kSlotBits;
      for (int32_t t = p; t < q; ++t) {
        if (t == s) continue;
        out[k * kSlice] = uint32_t(t) | hi;
        ++k;
      }
    };
    if (recs) {
      box_walk(B, v.rows, v.nrows, ci, cj, ck, [&](int q0, int q1, uint32_t code) {
        const uint32_t hi = code << kSlotBits;
        for (int r = cst[q0], e = cst[q1 + 1]; r < e; ++r) {
          const int32_t t = rslot[r];
          if (t == s) continue;
          out[k * kSlice] = uint32_t(t) | hi;
          ++k;
        }
      });
    } else if (staged) {
      box_walk(B, v.rows, v.nrows, ci, cj, ck, [&](int q0, int q1, uint32_t code) {
        for (int q = q0; q <= q1; ++q) emit(first[q], first[q] + (cst[q + 1] - cst[q]), code);
      });
    } else {
      for_each_segment(v.g, v.rows, v.nrows, v.cs, ci, cj, ck, emit);
    }
  }
}

// Element lists of the spreading by brick (one system): element = cell,
// the own element is part of its stencil and nothing is excluded.  Every
// brick has elements, particles or not.
__global__ void __launch_bounds__(kBrickThreads) k_brick_elist_count(BrickArgs a, int64_t* __restrict__ cnt) {
  extern __shared__ int32_t cst[];  // [cells_cap + 1]
  __shared__ int32_t s_w[kBrickThreads / 64];
  BoxGeom B;
  B.init(a.g, 0, a.reach, blockIdx.x);
  const bool staged = brick_staged(a, B);
  if (staged) stage_box_cells<kBrickThreads>(B, a.cs, nullptr, cst, s_w);
  for (int l = int(threadIdx.x); l < kBrickCells; l += kBrickThreads) {
    int ci, cj, ck;
    int64_t e;
    if (!brick_elem(B, l, ci, cj, ck, e)) continue;
    int64_t c = 0;
    if (staged)
      box_walk(B, a.rows, a.nrows, ci, cj, ck, [&](int q0, int q1, uint32_t) { c += cst[q1 + 1] - cst[q0]; });
    else
      for_each_segment(a.g, a.rows, a.nrows, a.cs, ci, cj, ck, [&](int32_t p, int32_t q, uint32_t) { c += q - p; });
    cnt[e] = c;
  }
}

// The lists of a block's consecutive elements form one contiguous output
// range: they are assembled in LDS and written out with coalesced stores
// (direct stores when the range exceeds the LDS buffer).
constexpr int kElistLds = 10240;

__global__ void __launch_bounds__(256) k_elist_fill(Grid g, const StencilRow* __restrict__ rows, int nrows,
                                                   const int32_t* __restrict__ cell_start,
                                                   const int64_t* __restrict__ off,
                                                   uint32_t* __restrict__ elist) {
  __shared__ uint32_t buf[kElistLds];
  const int64_t e0 = int64_t(blockIdx.x) * blockDim.x;
  const int64_t e_end = min(e0 + int64_t(blockDim.x), g.n_elem);
  const int64_t base = off[e0];
  const int64_t total = off[e_end] - base;
  const bool staged = total <= kElistLds;
  const int64_t e = e0 + threadIdx.x;
  if (e < g.n_elem) {
    int ci, cj, ck;
    elem_ijk(g, e, ci, cj, ck);
    const int64_t my = off[e] - base;
    int64_t k = 0;
    for_each_segment(g, rows, nrows, cell_start, ci, cj, ck, [&](int32_t a, int32_t b, uint32_t code) {
      const uint32_t hi = code << kSlotBits;
      for (int32_t t = a; t < b; ++t) {
        if (staged) buf[my + k] = uint32_t(t) | hi;
        else elist[base + my + k] = uint32_t(t) | hi;
        ++k;
      }
    });
  }
  if (!staged) return;  // block-uniform
  __syncthreads();
  for (int64_t i = threadIdx.x; i < total; i += blockDim.x) elist[base + i] = buf[i];
}

// Exclusive scan of int64 values (three phases, as the int32 scan).
constexpr int kScan64Threads = 256;
constexpr int kScan64Items = 4;
constexpr int kScan64Tile = kScan64Threads * kScan64Items;

__device__ __forceinline__ int64_t block_exclusive_scan64(int64_t v, int64_t* sh, int64_t& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScan64Threads; off <<= 1) {
    const int64_t add = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  total = sh[kScan64Threads - 1];
  const int64_t incl = sh[t];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan64_tiles(const int64_t* __restrict__ in, int64_t len, int64_t* __restrict__ out,
                               int64_t* __restrict__ tile_sums) {
  __shared__ int64_t sh[kScan64Threads];
  const int64_t base = int64_t(blockIdx.x) * kScan64Tile + int64_t(threadIdx.x) * kScan64Items;
  int64_t v[kScan64Items];
  int64_t s = 0;
#pragma unroll
  for (int q = 0; q < kScan64Items; ++q) {
    v[q] = (base + q < len) ? in[base + q] : 0;
    s += v[q];
  }
  int64_t total;
  int64_t run = block_exclusive_scan64(s, sh, total);
#pragma unroll
  for (int q = 0; q < kScan64Items; ++q) {
    if (base + q < len) out[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void k_scan64_sums(int64_t* __restrict__ sums, int64_t ntiles) {
  __shared__ int64_t sh[kScan64Threads];
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < ntiles; c0 += kScan64Threads) {
    const int64_t i = c0 + threadIdx.x;
    const int64_t v = (i < ntiles) ? sums[i] : 0;
    int64_t total;
    const int64_t ex = block_exclusive_scan64(v, sh, total);
    if (i < ntiles) sums[i] = ex + carry;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[ntiles] = carry;
}

__global__ void k_scan64_add(int64_t* __restrict__ out, int64_t len, const int64_t* __restrict__ sums,
                             int64_t ntiles) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len) out[i] += sums[i / kScan64Tile];
  if (i == 0) out[len] = sums[ntiles];
}

// Neighbour count of every particle (own element + stencil, minus itself).
__global__ void k_count_pairs(Grid g, const StencilRow* __restrict__ rows, int nrows,
                              const int32_t* __restrict__ cell_start,
                              const int32_t* __restrict__ slot_cell,
                              const int32_t* __restrict__ perm, int64_t n,
                              int64_t* __restrict__ counts, unsigned long long* __restrict__ total) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t cnt = 0;
  if (s < n) {
    int ci, cj, ck;
    cell_ijk(g, slot_cell[s], ci, cj, ck);
    for_each_segment(g, rows, nrows, cell_start, ci, cj, ck,
                     [&](int32_t a, int32_t b, uint32_t) { cnt += b - a; });
    cnt -= 1;
    if (counts) counts[perm[s]] = cnt;
  }
  // wave reduction then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(total, (unsigned long long)cnt);
}

// Reference order of a particle list: own element's particles (ascending id,
// itself removed), then the other stencil elements by ascending element id,
// each by ascending particle id.  Slots within a cell are in particle-id
// order; the tail (cells of the stencil other than the own) is sorted by
// (element id, particle id).  c: brick-order index of cell (ci, cj, ck).
__device__ void emit_list(const Grid& g, const StencilRow* rows, int nrows, const int32_t* cell_start,
                          int ci, int cj, int ck, int64_t c, int32_t self_slot, const int32_t* perm,
                          const int32_t* elem, int32_t* out) {
  const int32_t oa = cell_start[c], ob = cell_start[c + 1];
  int64_t w = 0;
  for (int32_t t = oa; t < ob; ++t)
    if (t != self_slot) out[w++] = t;
  const int64_t head = w;
  for_each_segment(g, rows, nrows, cell_start, ci, cj, ck, [&](int32_t a, int32_t b, uint32_t) {
    for (int32_t t = a; t < b; ++t)
      if (t < oa || t >= ob) out[w++] = t;
  });
  auto key = [&](int32_t t) {
    const int32_t p = perm[t];
    return (uint64_t(uint32_t(elem[p])) << 32) | uint32_t(p);
  };
  for (int64_t q = head + 1; q < w; ++q) {  // insertion sort of the tail
    const int32_t v = out[q];
    const uint64_t kv = key(v);
    int64_t t = q - 1;
    while (t >= head && key(out[t]) > kv) {
      out[t + 1] = out[t];
      --t;
    }
    out[t + 1] = v;
  }
  for (int64_t q = 0; q < w; ++q) out[q] = perm[out[q]];
}

__global__ void k_fill_csr(Grid g, const StencilRow* __restrict__ rows, int nrows,
                           const int32_t* __restrict__ cell_start,
                           const int32_t* __restrict__ slot_cell, const int32_t* __restrict__ perm,
                           const int32_t* __restrict__ elem, int64_t n, const int64_t* __restrict__ offsets,
                           int32_t* __restrict__ ids) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int32_t p = perm[s];
  const int64_t c = slot_cell[s];
  int ci, cj, ck;
  cell_ijk(g, c, ci, cj, ck);
  emit_list(g, rows, nrows, cell_start, ci, cj, ck, c, int32_t(s), perm, elem, ids + offsets[p]);
}

// e runs over reference element ids
__global__ void k_elem_point(Grid g, const StencilRow* __restrict__ rows, int nrows,
                             const int32_t* __restrict__ cell_start,
                             const int32_t* __restrict__ perm, const int32_t* __restrict__ elem,
                             int64_t e0, int64_t ne, int64_t* __restrict__ offsets, int32_t* __restrict__ ids,
                             int pass) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= ne) return;
  const int64_t e = e0 + q;
  int ci, cj, ck;
  elem_ijk(g, e, ci, cj, ck);
  if (pass == 0) {
    int64_t cnt = 0;
    for_each_segment(g, rows, nrows, cell_start, ci, cj, ck,
                     [&](int32_t a, int32_t b, uint32_t) { cnt += b - a; });
    offsets[q + 1] = cnt;
  } else {
    emit_list(g, rows, nrows, cell_start, ci, cj, ck, cell_index(g, ci, cj, ck), -1, perm, elem,
              ids + offsets[q]);
  }
}

// PointMesh::reinit(false) (point_mesh.C:245-289 without the list
// update): the slot positions take the particles' new positions; the FP32C
// records are relative to the centre of the cell now holding the particle
// (floor, clamped), the lists, elements and slot order stay.
__global__ void k_update_slots(Grid g, int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ x,
                               double* __restrict__ xs, double4* __restrict__ x4, float4* __restrict__ pc) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int64_t p = perm[s];
  const double v[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
  int c[3];
  float o[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double q = floor((v[d] - g.bmin[d]) / g.h[d]);
    q = q < 0.0 ? 0.0 : (q > double(g.n[d] - 1) ? double(g.n[d] - 1) : q);
    c[d] = int(q);
    o[d] = float(v[d] - (g.bmin[d] + (q + 0.5) * g.h[d]));
    xs[d * n + s] = v[d];
  }
  x4[s] = make_double4(v[0], v[1], v[2], 0.0);
  const uint32_t packed = uint32_t(c[0]) | (uint32_t(c[1]) << kCellBitsX) | (uint32_t(c[2]) << (kCellBitsX + kCellBitsY));
  pc[s] = make_float4(o[0], o[1], o[2], __uint_as_float(packed));
}

// PointParticle::set_neighbor_list (point_particle.h:189-194): a slot whose
// new list is empty keeps its particle's previous list (keep_off null: no
// previous lists).
__global__ void k_stale_patch(int64_t n, const int32_t* __restrict__ perm, const int64_t* __restrict__ keep_off,
                              int32_t* __restrict__ cnt, uint8_t* __restrict__ stale, int32_t* __restrict__ flag) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint8_t st = 0;
  if (keep_off && cnt[s] == 0) {
    const int64_t p = perm[s];
    const int64_t m = keep_off[p + 1] - keep_off[p];
    if (m > 0) {
      cnt[s] = int32_t(m);
      st = 1;
      atomicOr(flag, 1);
    }
  }
  stale[s] = st;
}

// list length of every particle, by particle id
__global__ void k_counts_by_particle(int64_t n, const int32_t* __restrict__ slot_of, const int32_t* __restrict__ cnt,
                                     int64_t* __restrict__ out) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < n) out[p] = cnt[slot_of[p]];
}

// every particle's list in the reference's order: the kept one, or the
// stencil list of this build
__global__ void k_fill_keep(Grid g, const StencilRow* __restrict__ rows, int nrows,
                            const int32_t* __restrict__ cell_start, const int32_t* __restrict__ slot_cell,
                            const int32_t* __restrict__ perm, const int32_t* __restrict__ elem, int64_t n,
                            const uint8_t* __restrict__ stale, const int64_t* __restrict__ old_off,
                            const int32_t* __restrict__ old_ids, const int64_t* __restrict__ off,
                            int32_t* __restrict__ ids) {
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int32_t p = perm[s];
  if (stale[s]) {
    for (int64_t q = old_off[p], w = off[p]; q < old_off[p + 1]; ++q, ++w) ids[w] = old_ids[q];
    return;
  }
  const int64_t c = slot_cell[s];
  int ci, cj, ck;
  cell_ijk(g, c, ci, cj, ck);
  emit_list(g, rows, nrows, cell_start, ci, cj, ck, c, int32_t(s), perm, elem, ids + off[p]);
}

__global__ void k_locate(Grid g, int64_t m, const double* __restrict__ pts, int32_t* __restrict__ elem) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m) return;
  elem[q] = int32_t(locate_dev(g, pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]));
}

}  // namespace

hipError_t launch_bin(ggem_ctx* ctx, bool hysteresis) {
  hipError_t e = hipMemsetAsync(ctx->cell_count, 0, sizeof(int32_t) * ctx->grid.n_cells, ctx->stream);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(ctx->flag, 0, sizeof(int32_t), ctx->stream);
  if (e != hipSuccess) return e;
  if (ctx->n == 0) return hipSuccess;
  ProfScope ps(ctx, "bin");
  if (ctx->nsys > 1)
    hipLaunchKernelGGL(k_bin<true>, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->grid,
                       ctx->sys_d, ctx->sys_part, ctx->n, ctx->x_in, ctx->elem, hysteresis ? 1 : 0,
                       ctx->pcell, ctx->cell_count, ctx->flag);
  else
    hipLaunchKernelGGL(k_bin<false>, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->grid,
                       nullptr, nullptr, ctx->n, ctx->x_in, ctx->elem, hysteresis ? 1 : 0, ctx->pcell,
                       ctx->cell_count, ctx->flag);
  return hipGetLastError();
}

hipError_t launch_scan(ggem_ctx* ctx) {
  const int64_t len = ctx->grid.n_cells;
  const int64_t ntiles = (len + kScanTile - 1) / kScanTile;
  ProfScope ps(ctx, "scan");
  hipLaunchKernelGGL(k_scan_tiles, dim3(unsigned(ntiles)), dim3(kScanThreads), 0, ctx->stream,
                     ctx->cell_count, len, ctx->cell_start, ctx->scan_tmp);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, ctx->stream, ctx->scan_tmp, ntiles);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(len)), dim3(kBlock), 0, ctx->stream, ctx->cell_start,
                     len, ctx->scan_tmp, int32_t(ctx->n));
  return hipGetLastError();
}

hipError_t launch_place(ggem_ctx* ctx) {
  hipError_t e = hipMemsetAsync(ctx->cell_count, 0, sizeof(int32_t) * ctx->grid.n_cells, ctx->stream);
  if (e != hipSuccess) return e;
  ProfScope ps(ctx, "place_sort");
  if (ctx->n > 0)
    hipLaunchKernelGGL(k_place, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n,
                       ctx->pcell, ctx->cell_start, ctx->cell_count, ctx->perm);
  hipLaunchKernelGGL(k_sort_cells, dim3(blocks_for(ctx->grid.n_cells)), dim3(kBlock), 0, ctx->stream,
                     ctx->grid.n_cells, ctx->n, ctx->cell_start, ctx->perm, ctx->slot_of,
                     ctx->slot_cell, ctx->x_in, ctx->xs, ctx->x4);
  if (ctx->n == 0) return hipGetLastError();
  if (ctx->nsys > 1) {
    hipLaunchKernelGGL(k_slot_sys, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->perm,
                       ctx->sys_part, ctx->sys_slot);
    hipLaunchKernelGGL(k_cell_rel<true>, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->grid,
                       ctx->sys_d, ctx->sys_slot, ctx->n, ctx->slot_cell, ctx->x4, ctx->pc32);
  } else {
    hipLaunchKernelGGL(k_cell_rel<false>, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->grid,
                       nullptr, nullptr, ctx->n, ctx->slot_cell, ctx->x4, ctx->pc32);
  }
  return hipGetLastError();
}

hipError_t launch_gather_forces(ggem_ctx* ctx) {
  if (ctx->n == 0) return hipSuccess;
  ProfScope ps(ctx, "gather_forces");
  hipLaunchKernelGGL(k_gather_forces, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n,
                     ctx->perm, ctx->f_in, ctx->fs, ctx->fs32, ctx->fg32, ctx->fg64);
  return hipGetLastError();
}

BrickArgs brick_args(const ggem_ctx* ctx) {
  BrickArgs a;
  a.g = ctx->grid;
  a.rows = ctx->rows_d;
  a.nrows = ctx->n_rows;
  for (int d = 0; d < 3; ++d) a.reach[d] = ctx->reach[d];
  a.cs = ctx->cell_start;
  a.slot_cell = ctx->slot_cell;
  a.sys = ctx->nsys > 1 ? ctx->sys_d : nullptr;
  a.sys_slot = ctx->nsys > 1 ? ctx->sys_slot : nullptr;
  // LDS tables sized for the largest box of any system
  auto cells = [](const int32_t* r) {
    return ((1 << kBrickBits) + 2 * r[0]) * ((1 << kBrickBits) + 2 * r[1]) * ((1 << kBrickBits) + 2 * r[2]);
  };
  int cap = cells(ctx->reach);
  for (int k = 0; k < ctx->nsys && ctx->nsys > 1; ++k) cap = std::max(cap, cells(ctx->sys_h[k].reach));
  a.cells_cap = std::min(cap, kBoxMaxCells);
  a.recs = ctx->brick_recs;
  a.stale = ctx->stale_lists ? ctx->stale_slot : nullptr;
  a.keep_off = ctx->keep_off;
  a.keep_ids = ctx->keep_ids;
  a.slot_of = ctx->slot_of;
  a.perm = ctx->perm;
  return a;
}

unsigned brick_count(const ggem_ctx* ctx) { return unsigned(ctx->grid.n_cells >> (3 * kBrickBits)); }

hipError_t launch_count_lists(ggem_ctx* ctx) {
  hipError_t e = hipMemsetAsync(ctx->pairs_d, 0, sizeof(int64_t), ctx->stream);
  if (e != hipSuccess || ctx->n == 0) return e;
  ProfScope ps(ctx, "count_lists");
  const BrickArgs a = brick_args(ctx);
  const size_t lds = sizeof(int32_t) * (size_t(a.cells_cap) + 1);
  if (a.sys)
    hipLaunchKernelGGL(k_brick_count<true>, dim3(brick_count(ctx)), dim3(kBrickThreads), lds, ctx->stream, a,
                       ctx->nbr_cnt);
  else
    hipLaunchKernelGGL(k_brick_count<false>, dim3(brick_count(ctx)), dim3(kBrickThreads), lds, ctx->stream, a,
                       ctx->nbr_cnt);
  hipLaunchKernelGGL(k_slice_len, dim3(blocks_for(ctx->n, kSliceBlock)), dim3(kSliceBlock), 0, ctx->stream,
                     ctx->nbr_cnt, ctx->n,
                     ctx->slice_len, reinterpret_cast<unsigned long long*>(ctx->pairs_d));
  return hipGetLastError();
}

hipError_t launch_scan64(ggem_ctx* ctx, const int64_t* in, int64_t len, int64_t* out) {
  const int64_t ntiles = (len + kScan64Tile - 1) / kScan64Tile;
  if (ctx->scan64_tmp_len < ntiles + 1) return hipErrorInvalidValue;
  ProfScope ps(ctx, "scan64");
  if (ntiles > 0)
    hipLaunchKernelGGL(k_scan64_tiles, dim3(unsigned(ntiles)), dim3(kScan64Threads), 0, ctx->stream, in,
                       len, out, ctx->scan64_tmp);
  hipLaunchKernelGGL(k_scan64_sums, dim3(1), dim3(kScan64Threads), 0, ctx->stream, ctx->scan64_tmp,
                     ntiles);
  hipLaunchKernelGGL(k_scan64_add, dim3(blocks_for(len + 1)), dim3(kBlock), 0, ctx->stream, out, len,
                     ctx->scan64_tmp, ntiles);
  return hipGetLastError();
}

hipError_t launch_fill_ell(ggem_ctx* ctx) {
  if (ctx->n == 0) return hipSuccess;
  ProfScope ps(ctx, "fill_ell");
  const BrickArgs a = brick_args(ctx);
  const size_t lds = sizeof(int32_t) * (2 * size_t(a.cells_cap) + 1 + kBrickRecCap);
  if (a.sys)
    hipLaunchKernelGGL(k_brick_fill_ell<true>, dim3(brick_count(ctx)), dim3(kBrickThreads), lds, ctx->stream, a,
                       ctx->slice_off, ctx->ell);
  else
    hipLaunchKernelGGL(k_brick_fill_ell<false>, dim3(brick_count(ctx)), dim3(kBrickThreads), lds, ctx->stream,
                       a, ctx->slice_off, ctx->ell);
  return hipGetLastError();
}

hipError_t launch_part_sys(ggem_ctx* ctx) {
  if (ctx->n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_part_sys, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->nsys,
                     ctx->part_off_d, ctx->sys_part);
  return hipGetLastError();
}

hipError_t launch_elist_count(ggem_ctx* ctx) {
  ProfScope ps(ctx, "elist_count");
  const BrickArgs a = brick_args(ctx);
  const size_t lds = sizeof(int32_t) * (size_t(a.cells_cap) + 1);
  hipLaunchKernelGGL(k_brick_elist_count, dim3(brick_count(ctx)), dim3(kBrickThreads), lds, ctx->stream, a,
                     ctx->elist_off);
  return hipGetLastError();
}

hipError_t launch_elist_fill(ggem_ctx* ctx) {
  ProfScope ps(ctx, "elist_fill");
  hipLaunchKernelGGL(k_elist_fill, dim3(blocks_for(ctx->grid.n_elem)), dim3(kBlock), 0, ctx->stream,
                     ctx->grid, ctx->rows_d, ctx->n_rows, ctx->cell_start, ctx->elist_off, ctx->elist);
  return hipGetLastError();
}

hipError_t launch_count_pairs(ggem_ctx* ctx, int64_t* counts_d) {
  hipError_t e = hipMemsetAsync(ctx->pairs_d, 0, sizeof(int64_t), ctx->stream);
  if (e != hipSuccess || ctx->n == 0) return e;
  hipLaunchKernelGGL(k_count_pairs, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->grid,
                     ctx->rows_d, ctx->n_rows, ctx->cell_start, ctx->slot_cell, ctx->perm, ctx->n,
                     counts_d, reinterpret_cast<unsigned long long*>(ctx->pairs_d));
  return hipGetLastError();
}

hipError_t launch_fill_csr(ggem_ctx* ctx, const int64_t* offsets_d, int32_t* ids_d) {
  if (ctx->n == 0) return hipSuccess;
  ProfScope ps(ctx, "fill_csr");
  hipLaunchKernelGGL(k_fill_csr, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->grid,
                     ctx->rows_d, ctx->n_rows, ctx->cell_start, ctx->slot_cell, ctx->perm, ctx->elem, ctx->n,
                     offsets_d, ids_d);
  return hipGetLastError();
}

hipError_t launch_elem_point_csr(ggem_ctx* ctx, int64_t e0, int64_t ne, int64_t* offsets_d,
                                 int32_t* ids_d, int pass) {
  if (ne == 0) return hipSuccess;
  hipLaunchKernelGGL(k_elem_point, dim3(blocks_for(ne)), dim3(kBlock), 0, ctx->stream, ctx->grid,
                     ctx->rows_d, ctx->n_rows, ctx->cell_start, ctx->perm, ctx->elem, e0, ne, offsets_d, ids_d,
                     pass);
  return hipGetLastError();
}

hipError_t launch_update_slots(ggem_ctx* ctx) {
  if (ctx->n == 0) return hipSuccess;
  ProfScope ps(ctx, "update_slots");
  hipLaunchKernelGGL(k_update_slots, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->grid, ctx->n,
                     ctx->perm, ctx->x_in, ctx->xs, ctx->x4, ctx->pc32);
  return hipGetLastError();
}

hipError_t launch_stale_patch(ggem_ctx* ctx, bool have_kept) {
  hipError_t e = hipMemsetAsync(ctx->flag, 0, sizeof(int32_t), ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(ctx->pairs_d, 0, sizeof(int64_t), ctx->stream);
  if (e != hipSuccess || ctx->n == 0) return e;
  ProfScope ps(ctx, "stale_lists");
  hipLaunchKernelGGL(k_stale_patch, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n, ctx->perm,
                     have_kept ? ctx->keep_off : nullptr, ctx->nbr_cnt, ctx->stale_slot, ctx->flag);
  hipLaunchKernelGGL(k_slice_len, dim3(blocks_for(ctx->n, kSliceBlock)), dim3(kSliceBlock), 0, ctx->stream,
                     ctx->nbr_cnt, ctx->n, ctx->slice_len, reinterpret_cast<unsigned long long*>(ctx->pairs_d));
  return hipGetLastError();
}

hipError_t launch_counts_by_particle(ggem_ctx* ctx, int64_t* out) {
  if (ctx->n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_counts_by_particle, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->n,
                     ctx->slot_of, ctx->nbr_cnt, out);
  return hipGetLastError();
}

hipError_t launch_fill_keep(ggem_ctx* ctx, const int64_t* old_off, const int32_t* old_ids, const int64_t* off,
                            int32_t* ids) {
  if (ctx->n == 0) return hipSuccess;
  ProfScope ps(ctx, "stale_lists");
  hipLaunchKernelGGL(k_fill_keep, dim3(blocks_for(ctx->n)), dim3(kBlock), 0, ctx->stream, ctx->grid, ctx->rows_d,
                     ctx->n_rows, ctx->cell_start, ctx->slot_cell, ctx->perm, ctx->elem, ctx->n, ctx->stale_slot,
                     old_off, old_ids, off, ids);
  return hipGetLastError();
}

hipError_t launch_locate(ggem_ctx* ctx, int64_t m, const double* pts, int32_t* elem) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_locate, dim3(blocks_for(m)), dim3(kBlock), 0, ctx->stream, ctx->grid, m, pts,
                     elem);
  return hipGetLastError();
}

}  // namespace ggem
