// ggem_api.hip -- the C-ABI of libcopss_ggem.so: context life cycle,
// particles, the neighbour build, local velocities, spreading, device noise
// and the event profiler.  The Chebyshev and Lanczos drivers live in
// ggem_cheb.hip.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ggem_host.h"

namespace {

thread_local std::string g_err;  // errors without a context
const double PI = 3.14159265358979323846;

int grow_scan64(ggem_ctx* ctx, int64_t len);

}  // namespace

namespace ggem {

int fail(ggem_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  else g_err = msg;
  return code;
}

int hip_fail(ggem_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, GGEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void prof_begin(ggem_ctx* ctx, const char*, hipEvent_t* a, hipEvent_t* b) {
  if (!ctx->prof.on) return;
  auto get = [&]() {
    hipEvent_t ev;
    if (!ctx->prof.pool.empty()) {
      ev = ctx->prof.pool.back();
      ctx->prof.pool.pop_back();
    } else if (hipEventCreate(&ev) != hipSuccess) {
      ev = nullptr;
    }
    return ev;
  };
  *a = get();
  *b = get();
  if (*a) (void)hipEventRecord(*a, ctx->stream);
}

void prof_end(ggem_ctx* ctx, const char* name, hipEvent_t a, hipEvent_t b) {
  if (!ctx->prof.on || !a || !b) return;
  (void)hipEventRecord(b, ctx->stream);
  ctx->prof.pending.push_back({name, a, b});
}

static void prof_collect(ggem_ctx* ctx) {
  if (ctx->prof.pending.empty()) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& r : ctx->prof.pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    bool found = false;
    for (auto& acc : ctx->prof.acc)
      if (acc.name == r.name) {
        acc.ms += ms;
        acc.launches += 1;
        found = true;
        break;
      }
    if (!found) ctx->prof.acc.push_back({r.name, double(ms), 1});
    ctx->prof.pool.push_back(r.a);
    ctx->prof.pool.push_back(r.b);
  }
  ctx->prof.pending.clear();
}

int tune_lookup(const ggem_ctx* ctx, const std::string& key, int dflt) {
  if (!ctx->tune.on) return dflt;
  auto it = ctx->tune.pick.find(key);
  return it == ctx->tune.pick.end() ? dflt : it->second;
}

std::string tune_key(const char* kernel, int precision, int extra, int64_t n) {
  int lg = 0;
  while ((int64_t(1) << (lg + 1)) <= n) ++lg;
  return std::string(kernel) + (precision == GGEM_FP64 ? "/fp64/" : "/fp32c/") + std::to_string(extra) + "/" +
         std::to_string(lg);
}

int tune_pick(ggem_ctx* ctx, const std::string& key, int ncand, const std::function<hipError_t(int)>& launch,
              int dflt) {
  Tuner& T = ctx->tune;
  if (!T.on) return dflt;
  auto it = T.pick.find(key);
  if (it != T.pick.end()) return it->second;
  if (ncand <= 1) return dflt;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return dflt;
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess) return dflt;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return dflt;
  }
  const bool prof = ctx->prof.on;
  ctx->prof.on = false;
  std::vector<float> ms(ncand, 1e30f);
  for (int c = 0; c < ncand; ++c) {
    if (launch(c) != hipSuccess) continue;  // warm-up; a shape that cannot launch is skipped
    bool ok = hipEventRecord(a, ctx->stream) == hipSuccess;
    for (int r = 0; r < 3 && ok; ++r) ok = launch(c) == hipSuccess;
    ok = ok && hipEventRecord(b, ctx->stream) == hipSuccess && hipEventSynchronize(b) == hipSuccess;
    float t = 0.f;
    if (ok && hipEventElapsedTime(&t, a, b) == hipSuccess) ms[c] = t / 3.f;
  }
  (void)hipGetLastError();
  ctx->prof.on = prof;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  int best = dflt;
  for (int c = 0; c < ncand; ++c)
    if (ms[c] < ms[best]) best = c;
  T.pick[key] = best;
  T.ms[key] = ms;
  return best;
}

// Staged neighbour records of the FP32C lists (per brick of cells): counts,
// their scans, one synchronisation for the sizes, then records, target
// layout and record numbers.
int prepare_box(ggem_ctx* ctx) {
  ctx->box_valid = false;
  ctx->box_unstaged = false;
  ctx->box_max = 0;
  // FP64 stores every pair's record number at its list position (the
  // tensors stay in list order); FP32C the kept pairs' in the brick layout
  ctx->box_fp64 = ctx->p.precision == GGEM_FP64;
  if (!box_supported(ctx)) return GGEM_OK;
  const size_t nb = size_t(ctx->grid.n_cells >> (3 * kBrickBits));
  if (ctx->box_brick_cap < nb + 1) {
    if (ctx->box_cnt) (void)hipFree(ctx->box_cnt);
    ctx->box_cnt = ctx->box_off = ctx->box_ecnt = ctx->box_eoff = nullptr;
    ctx->box_brick_cap = 0;
    GGEM_HIP(ctx, hipMalloc(&ctx->box_cnt, sizeof(int64_t) * 4 * (nb + 1)));
    ctx->box_brick_cap = nb + 1;
  }
  ctx->box_off = ctx->box_cnt + ctx->box_brick_cap;
  ctx->box_ecnt = ctx->box_cnt + 2 * ctx->box_brick_cap;
  ctx->box_eoff = ctx->box_cnt + 3 * ctx->box_brick_cap;
  if (!ctx->box_max_d) GGEM_HIP(ctx, hipMalloc(&ctx->box_max_d, 3 * sizeof(int32_t)));
  int rc;
  if ((rc = ensure(ctx, &ctx->box_tgt, &ctx->box_tgt_cap, size_t(ctx->n)))) return rc;
  if ((rc = ensure(ctx, &ctx->box_base, &ctx->box_base_cap, size_t(ctx->n)))) return rc;
  if ((rc = grow_scan64(ctx, int64_t(nb) + 1))) return rc;
  hipError_t e = launch_box_count(ctx);
  if (e != hipSuccess) return hip_fail(ctx, e, "box_count");
  if ((e = launch_scan64(ctx, ctx->box_cnt, int64_t(nb), ctx->box_off)) != hipSuccess ||
      (e = launch_scan64(ctx, ctx->box_ecnt, int64_t(nb), ctx->box_eoff)) != hipSuccess)
    return hip_fail(ctx, e, "box scan");
  int64_t total[2] = {0, 0};
  int64_t mx[3] = {0, 0, 0};
  {
    int64_t v[5];
    if ((rc = peek(ctx, {{ctx->box_off + nb, 8}, {ctx->box_eoff + nb, 8}, {ctx->box_max_d, 4}, {ctx->box_max_d + 1, 4},
                         {ctx->box_max_d + 2, 4}}, v)))
      return rc;
    total[0] = v[0];
    total[1] = v[1];
    mx[0] = v[2];
    mx[1] = v[3];
    mx[2] = v[4];
  }
  // no brick staged, or a brick too crowded for the entry layout: the
  // list-order kernels serve
  if (total[0] <= 0 || mx[0] <= 0 || mx[1] > kBoxMaxTargets) return GGEM_OK;
  if ((rc = ensure(ctx, &ctx->box_slot, &ctx->box_rec_cap, size_t(total[0])))) return rc;
  if ((rc = ensure(ctx, &ctx->box_pos, &ctx->box_pos_cap, size_t(total[0])))) return rc;
  const size_t ne = size_t(std::max<int64_t>(ctx->box_fp64 ? ctx->ell_len : total[1], 1));
  if ((rc = ensure(ctx, &ctx->box_idx, &ctx->box_idx_cap, ne))) return rc;
  // bricks that are not staged fill their tensors from the lists
  ctx->box_unstaged = mx[2] > 0;
  if (ctx->box_unstaged && (rc = ensure_ell(ctx))) return rc;
  if ((e = launch_box_fill(ctx, int32_t(mx[0]), int32_t(mx[1]))) != hipSuccess) return hip_fail(ctx, e, "box_fill");
  ctx->box_max = int32_t(mx[0]);
  ctx->box_valid = true;
  return GGEM_OK;
}

int ensure_ell(ggem_ctx* ctx) {
  if (ctx->ell_valid) return GGEM_OK;
  // slice offsets (the scan of the slice lengths of the count pass), one
  // synchronisation for the size, then the lists
  const int64_t nsl = (ctx->n + kSlice - 1) / kSlice;
  hipError_t e = launch_scan64(ctx, ctx->slice_len, nsl, ctx->slice_off);
  if (e != hipSuccess) return hip_fail(ctx, e, "scan64");
  int64_t len = 0;
  int rc = peek(ctx, {{ctx->slice_off + nsl, 8}}, &len);
  if (rc) return rc;
  ctx->ell_len = len;
  size_t cap = size_t(ctx->ell_cap);
  if ((rc = ensure(ctx, &ctx->ell, &cap, size_t(std::max<int64_t>(len, 1))))) return rc;
  ctx->ell_cap = int64_t(cap);
  if ((e = launch_fill_ell(ctx)) != hipSuccess) return hip_fail(ctx, e, "fill_ell");
  ctx->ell_valid = true;
  return GGEM_OK;
}

int ensure_tensors(ggem_ctx* ctx, int nrhs) {
  int rc;
  if (ctx->p.precision == GGEM_FP64) {
    if ((rc = ensure_ell(ctx))) return rc;
    if (!ctx->coef64_valid) {
      size_t cap = size_t(ctx->coef64_cap);
      if ((rc = ensure(ctx, &ctx->coef64, &cap, size_t(std::max<int64_t>(ctx->ell_len, 1))))) return rc;
      ctx->coef64_cap = int64_t(cap);
      hipError_t e = launch_pair_tensors(ctx, GGEM_FP64);
      if (e != hipSuccess) return hip_fail(ctx, e, "pair tensors");
      ctx->coef64_valid = true;
    }
    // the staged-record kernel serves one right-hand side
    if (!ctx->box_built) {
      if ((rc = prepare_box(ctx))) return rc;
      ctx->box_built = true;
    }
    return GGEM_OK;
  }
  // FP32C keeps no tensors (the pair sums evaluate the coefficient table);
  // the staged-brick records serve one right-hand side
  if (!ctx->box_built) {
    if ((rc = prepare_box(ctx))) return rc;
    ctx->box_built = true;
  }
  // the list-order kernels read the lists: several operands, no staged
  // bricks (or some unstaged), and the autotuner, which times every shape
  if (nrhs > 1 || !ctx->box_valid || ctx->box_unstaged || ctx->tune.on) return ensure_ell(ctx);
  return GGEM_OK;
}

}  // namespace ggem

using namespace ggem;

namespace {

// GGEMStokes::set_mu / set_alpha / set_ksi (ggem_stokes.C:39-114) and the
// search radii of copss.C:986-992, :1074-1075.
int derive(const ggem_params* p, ggem_consts* c, std::string* why) {
  std::memset(c, 0, sizeof(*c));
  for (int d = 0; d < 3; ++d) {
    if (p->n_elem[d] <= 0) {
      *why = "n_elem must be positive";
      return GGEM_ERR_INVALID_ARG;
    }
    if (!(p->box_max[d] > p->box_min[d])) {
      *why = "box_max must exceed box_min";
      return GGEM_ERR_INVALID_ARG;
    }
    c->L[d] = p->box_max[d] - p->box_min[d];
    c->h[d] = c->L[d] / double(p->n_elem[d]);
  }
  if ((int64_t)p->n_elem[0] * p->n_elem[1] * p->n_elem[2] >= INT32_MAX) {
    *why = "element count exceeds int32";
    return GGEM_ERR_CAPACITY;
  }
  if (p->n_elem[0] > (1 << kCellBitsX) || p->n_elem[1] > (1 << kCellBitsY) || p->n_elem[2] > (1 << kCellBitsZ)) {
    *why = "more than 2048 x 2048 x 1024 elements per axis";
    return GGEM_ERR_CAPACITY;
  }
  if (!(p->mu > 0) || !(p->alpha > 0)) {
    *why = "mu and alpha must be positive";
    return GGEM_ERR_INVALID_ARG;
  }
  if (p->point_type < GGEM_POLYMER_BEAD || p->point_type > GGEM_LAGRANGIAN_POINT) {
    *why = "undefined point type";  // ggem_stokes.C:98-103
    return GGEM_ERR_INVALID_ARG;
  }
  if (p->precision != GGEM_FP64 && p->precision != GGEM_FP32C) {
    *why = "precision must be GGEM_FP64 or GGEM_FP32C";
    return GGEM_ERR_INVALID_ARG;
  }
  const double sqrt_pi = std::sqrt(PI), pi_23 = std::pow(PI, 3. / 2.);
  c->one_eight_pi_mu = 1. / (8. * PI * p->mu);
  c->alpha = p->alpha;
  c->alpha2 = p->alpha * p->alpha;
  c->alpha3_pi_23 = (c->alpha2 * p->alpha) / pi_23;
  c->two_alpha_sqrt_pi = 2. * p->alpha / sqrt_pi;
  c->coeff_1 = c->one_eight_pi_mu * 4.0 * p->alpha / sqrt_pi;
  double ksi = p->ksi;
  if (!(ksi > 0)) {
    if (p->point_type == GGEM_LAGRANGIAN_POINT) {
      if (!(p->ibm_beta > 0) || !(p->hmin > 0)) {
        *why = "Lagrangian points need ibm_beta and hmin";  // ggem_stokes.C:67-74
        return GGEM_ERR_INVALID_ARG;
      }
      ksi = 1.0 / (p->ibm_beta * p->hmin);
    } else {
      if (!(p->br0 > 0)) {
        *why = "br0 must be positive";
        return GGEM_ERR_INVALID_ARG;
      }
      ksi = sqrt_pi / (3. * p->br0);
    }
  }
  c->ksi = ksi;
  c->ksi2 = ksi * ksi;
  c->two_ksi_sqrt_pi = 2. * ksi / sqrt_pi;
  c->coeff_2 = c->one_eight_pi_mu * 4.0 * (ksi - p->alpha) / sqrt_pi;
  const double hmax = std::max(std::max(c->h[0], c->h[1]), c->h[2]);
  c->search_radius_p = 4. / p->alpha;
  c->search_radius_e = p->search_radius_e > 0 ? p->search_radius_e : 0.5 * hmax + 4. / p->alpha;
  return GGEM_OK;
}

// The element stencil: every cell offset whose centre distance is strictly
// below R_e (build_elem_elem_neighbor_list's test, point_mesh.C:344-352),
// offsets taken as the distinct minimum-image representatives on periodic
// axes, d in (-n/2, n/2].  Offsets within 1e-9 R_e of R_e are rejected: the
// reference compares centroids assembled from mesh nodes, and its rounding
// would decide such a tie.
int build_stencil(const ggem_params* p, const ggem_consts* c, std::vector<StencilRow>* rows,
                  int32_t* ncells, int* reach, std::string* why) {
  const double Re = c->search_radius_e;
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    const int n = p->n_elem[d];
    const int w = int(std::ceil(Re / c->h[d]));
    if (p->periodic[d]) {
      lo[d] = -((n + 1) / 2) + 1;  // -ceil(n/2) + 1
      hi[d] = n / 2;
    } else {
      lo[d] = -(n - 1);
      hi[d] = n - 1;
    }
    lo[d] = std::max(lo[d], -w);
    hi[d] = std::min(hi[d], w);
    reach[d] = 0;
  }
  rows->clear();
  int64_t cells = 0;
  for (int dk = lo[2]; dk <= hi[2]; ++dk)
    for (int dj = lo[1]; dj <= hi[1]; ++dj) {
      int dmin = 1 << 30, dmax = -(1 << 30);
      for (int di = lo[0]; di <= hi[0]; ++di) {
        const double dx = di * c->h[0], dy = dj * c->h[1], dz = dk * c->h[2];
        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (std::fabs(dist - Re) <= 1e-9 * Re) {
          char buf[160];
          std::snprintf(buf, sizeof(buf),
                        "element search radius %.17g ties the centre distance of offset (%d,%d,%d)",
                        Re, di, dj, dk);
          *why = buf;
          return GGEM_ERR_STENCIL_TIE;
        }
        if (dist < Re) {
          dmin = std::min(dmin, di);
          dmax = std::max(dmax, di);
        }
      }
      if (dmin <= dmax) {
        rows->push_back({dj, dk, dmin, dmax});
        cells += dmax - dmin + 1;
        reach[0] = std::max(reach[0], std::max(-dmin, dmax));
        reach[1] = std::max(reach[1], std::abs(dj));
        reach[2] = std::max(reach[2], std::abs(dk));
      }
    }
  *ncells = int32_t(cells);
  return GGEM_OK;
}

void free_all(ggem_ctx* ctx) {
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(ctx->rows_d);
  f(ctx->x_in);
  f(ctx->f_in);
  f(ctx->elem);
  f(ctx->perm);
  f(ctx->slot_of);
  f(ctx->slot_cell);
  f(ctx->xs);
  f(ctx->x4);
  f(ctx->fs);
  f(ctx->fs32);
  f(ctx->fg32);
  f(ctx->fg64);
  f(ctx->u_loc);
  f(ctx->u_out);
  f(ctx->cell_count);
  f(ctx->pcell);
  f(ctx->cell_start);
  f(ctx->scan_tmp);
  f(ctx->flag);
  f(ctx->pairs_d);
  f(ctx->nbr_cnt);
  f(ctx->slice_len);
  f(ctx->slice_off);
  f(ctx->scan64_tmp);
  f(ctx->ell);
  f(ctx->elist_off);
  f(ctx->elist);
  f(ctx->tab32);
  f(ctx->box_cnt);
  f(ctx->box_tgt);
  f(ctx->box_base);
  f(ctx->box_slot);
  f(ctx->box_pos);
  f(ctx->box_idx);
  f(ctx->box_max_d);
  f(ctx->spread_unstaged);
  f(ctx->brick_recs);
  f(ctx->pc32);
  f(ctx->coef64);
  f(ctx->noise);
  f(ctx->keep_y);
  f(ctx->keep_d);
  f(ctx->keep_off);
  f(ctx->keep_ids);
  f(ctx->keep_off2);
  f(ctx->keep_ids2);
  f(ctx->keep_cnt);
  f(ctx->stale_slot);
  f(ctx->bond_off);
  f(ctx->bond_inc);
  f(ctx->bond_list);
  f(ctx->chain_flag);
  f(ctx->vext);
  f(ctx->cforce);
  f(ctx->wall_xyz);
  f(ctx->x_saved);
  f(ctx->rout);
  f(ctx->images);
  f(ctx->dpar);
  f(ctx->sys_d);
  f(ctx->part_off_d);
  f(ctx->sys_part);
  f(ctx->sys_slot);
  f(ctx->fe_conn);
  f(ctx->fe_u);
  f(ctx->fe_inc_off);
  f(ctx->fe_inc);
  f(ctx->fe_node_order);
  f(ctx->fe_cinc_off);
  f(ctx->fe_cinc);
  f(ctx->rhs_items);
  f(ctx->rhs_tab);
  f(ctx->rhs_buf);
  f(ctx->rhs_el);
  if (ctx->par_h) (void)hipHostFree(ctx->par_h);
  ctx->par_h = nullptr;
  f(ctx->peek_d);
  ctx->peek_d = nullptr;
  if (ctx->peek_h) (void)hipHostFree(ctx->peek_h);
  ctx->peek_h = nullptr;
  if (ctx->par_ev) (void)hipEventDestroy(ctx->par_ev);
  ctx->par_ev = nullptr;
  if (ctx->cheb_exec) (void)hipGraphExecDestroy(ctx->cheb_exec);
  ctx->cheb_exec = nullptr;
  if (ctx->scratch) {
    Pools& P = pools(ctx);
    f(P.vec);
    f(P.g32);
    f(P.g64);
    f(P.io);
    f(P.io2);
    f(P.ids);
    f(P.offs);
    f(P.red);
    f(P.fe);
    f(P.rho);
    f(P.ifd);
    f(P.cref);
    f(P.cstop);
    delete &P;
    ctx->scratch = nullptr;
  }
  for (auto& r : ctx->prof.pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto ev : ctx->prof.pool) (void)hipEventDestroy(ev);
  ctx->prof.pending.clear();
  ctx->prof.pool.clear();
}

int grow_scan64(ggem_ctx* ctx, int64_t len) {
  const int64_t need = (len + 1023) / 1024 + 2;
  if (ctx->scan64_tmp_len >= need) return GGEM_OK;
  if (ctx->scan64_tmp) (void)hipFree(ctx->scan64_tmp);
  ctx->scan64_tmp = nullptr;
  ctx->scan64_tmp_len = 0;
  GGEM_HIP(ctx, hipMalloc(&ctx->scan64_tmp, sizeof(int64_t) * need));
  ctx->scan64_tmp_len = need;
  return GGEM_OK;
}

}  // namespace

int ggem::peek(ggem_ctx* ctx, std::initializer_list<std::pair<const void*, int32_t>> items, int64_t* out) {
  if (items.size() > size_t(kPeekMax)) return fail(ctx, GGEM_ERR_INVALID_ARG, "too many values to read back");
  if (!ctx->peek_d) GGEM_HIP(ctx, hipMalloc(&ctx->peek_d, sizeof(int64_t) * kPeekMax));
  if (!ctx->peek_h) GGEM_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->peek_h), sizeof(int64_t) * kPeekMax));
  PeekArgs p{};
  p.n = int32_t(items.size());
  p.dst = ctx->peek_d;
  int i = 0;
  for (const auto& it : items) {
    p.src[i] = it.first;
    p.bytes[i] = it.second;
    ++i;
  }
  hipError_t e = launch_peek(ctx, p);
  if (e != hipSuccess) return hip_fail(ctx, e, "readback");
  GGEM_HIP(ctx, hipMemcpyAsync(ctx->peek_h, ctx->peek_d, sizeof(int64_t) * p.n, hipMemcpyDeviceToHost, ctx->stream));
  const int rc = sync(ctx);
  if (rc) return rc;
  for (int k = 0; k < p.n; ++k) out[k] = ctx->peek_h[k];
  return GGEM_OK;
}

// Element particle lists for the spreading, built on first use after a
// neighbour build.
int ggem::ensure_elist(ggem_ctx* ctx) {
  if (ctx->elist_valid) return GGEM_OK;
  const int64_t ne = ctx->grid.n_elem;
  hipError_t e = launch_elist_count(ctx);
  if (e == hipSuccess) e = launch_scan64(ctx, ctx->elist_off, ne, ctx->elist_off);
  if (e != hipSuccess) return hip_fail(ctx, e, "element lists");
  int64_t total = 0;
  int rc = peek(ctx, {{ctx->elist_off + ne, 8}}, &total);
  if (rc) return rc;
  size_t cap = size_t(ctx->elist_cap);
  if ((rc = ensure(ctx, &ctx->elist, &cap, size_t(std::max<int64_t>(total, 1))))) return rc;
  ctx->elist_cap = int64_t(cap);
  if ((e = launch_elist_fill(ctx)) != hipSuccess) return hip_fail(ctx, e, "element list fill");
  ctx->elist_valid = true;
  return GGEM_OK;
}

namespace {

int set_particles_impl(ggem_ctx* ctx, int64_t n, const double* xyz, const double* f);

// A context over nsys suspensions (nsys == 1: the plain context).  All
// systems share the physical constants and the precision; grids, boxes
// and stencils are per system.
int create_impl(const ggem_params* params, int32_t nsys, ggem_ctx** out) {
  if (!out) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  if (!params) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null params");
  std::vector<ggem_consts> cs(nsys);
  std::vector<SysGrid> sg(nsys);
  std::vector<StencilRow> rows;
  std::string why;
  int exact = 0;
  int64_t ne_total = 0, nc_total = 0;
  int32_t ncells0 = 0;
  for (int32_t k = 0; k < nsys; ++k) {
    const ggem_params* pk = params + k;
    int rc = derive(pk, &cs[k], &why);
    if (rc) return fail(nullptr, rc, nsys > 1 ? "system " + std::to_string(k) + ": " + why : why);
    if (k > 0 && (pk->mu != params->mu || pk->alpha != params->alpha || cs[k].ksi != cs[0].ksi ||
                  pk->point_type != params->point_type || pk->precision != params->precision ||
                  pk->device != params->device || pk->compat_flags != params->compat_flags ||
                  (pk->compat_flags & GGEM_COMPAT_STALE_LISTS)))
      return fail(nullptr, GGEM_ERR_INVALID_ARG,
                  "system " + std::to_string(k) + ": a batch shares mu, alpha, xi, point type, precision, "
                  "flags and device");
    std::vector<StencilRow> rk;
    int32_t ncells = 0;
    int reach[3];
    rc = build_stencil(pk, &cs[k], &rk, &ncells, reach, &why);
    if (rc) return fail(nullptr, rc, nsys > 1 ? "system " + std::to_string(k) + ": " + why : why);
    if (k == 0) ncells0 = ncells;
    Grid& g = sg[k].g;
    for (int d = 0; d < 3; ++d) {
      g.bmin[d] = pk->box_min[d];
      g.L[d] = cs[k].L[d];
      g.h[d] = cs[k].h[d];
      g.half_L[d] = cs[k].L[d] / 2.0;
      g.n[d] = pk->n_elem[d];
      g.periodic[d] = pk->periodic[d] ? 1 : 0;
      // a neighbour entry's single shift equals the minimum image while every
      // pair stays inside half a box: (reach + 1) cells <= n/2
      if (g.periodic[d] && 2 * (reach[d] + 1) > pk->n_elem[d]) exact = 1;
    }
    g.n_elem = int64_t(pk->n_elem[0]) * pk->n_elem[1] * pk->n_elem[2];
    for (int d = 0; d < 3; ++d) g.nb[d] = (pk->n_elem[d] + (1 << kBrickBits) - 1) >> kBrickBits;
    g.n_cells = (int64_t(g.nb[0]) * g.nb[1] * g.nb[2]) << (3 * kBrickBits);
    sg[k].elem_off = ne_total;
    sg[k].cell_off = nc_total;
    sg[k].row_off = int32_t(rows.size());
    sg[k].nrows = int32_t(rk.size());
    for (int d = 0; d < 3; ++d) sg[k].reach[d] = reach[d];
    ne_total += g.n_elem;
    nc_total += g.n_cells;
    rows.insert(rows.end(), rk.begin(), rk.end());
  }
  if (ne_total >= INT32_MAX || nc_total >= INT32_MAX)
    return fail(nullptr, GGEM_ERR_CAPACITY, "element count exceeds int32");
  const ggem_consts& c = cs[0];
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(nullptr, GGEM_ERR_NO_DEVICE, "no HIP device visible");
  if (params->device < 0 || params->device >= ndev)
    return fail(nullptr, GGEM_ERR_INVALID_ARG, "device ordinal out of range");

  ggem_ctx* ctx = new ggem_ctx();
  ctx->p = *params;
  ctx->c = c;
  ctx->c.stencil_rows = sg[0].nrows;
  ctx->c.stencil_cells = ncells0;
  ctx->device = params->device;
  ctx->scratch = new Pools();
  ctx->exact_minimg = exact;
  ctx->nsys = nsys;
  ctx->sys_h = sg;
  ctx->part_off.assign(nsys + 1, 0);
  // the whole store is one element range; per-system views come from sys_h
  ctx->grid = sg[0].g;
  ctx->grid.n_elem = ne_total;
  ctx->grid.n_cells = nc_total;
  ctx->gc.eps = c.one_eight_pi_mu;
  ctx->gc.alpha = c.alpha;
  ctx->gc.alpha2 = c.alpha2;
  ctx->gc.alpha3_pi_23 = c.alpha3_pi_23;
  ctx->gc.two_alpha_sqrt_pi = c.two_alpha_sqrt_pi;
  ctx->gc.coeff_1 = c.coeff_1;
  ctx->gc.ksi = c.ksi;
  ctx->gc.ksi2 = c.ksi2;
  ctx->gc.two_ksi_sqrt_pi = c.two_ksi_sqrt_pi;
  ctx->gc.coeff_2 = c.coeff_2;
  ctx->gc.point_type = params->point_type;
  ctx->bsign = c.ksi >= c.alpha ? 1 : -1;  // sign of (c1 - c2): that of xi - alpha
  ctx->rows_h = rows;
  ctx->n_rows = int32_t(rows.size());
  for (int d = 0; d < 3; ++d) ctx->reach[d] = sg[0].reach[d];

  auto bail = [&](hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    free_all(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GGEM_ERR_HIP;
  };
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return bail(e, "hipSetDevice");
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(e, "hipStreamCreate");
  const int64_t ne = ctx->grid.n_elem, nc = ctx->grid.n_cells;
  const int64_t ntiles = (nc + 2047) / 2048;
  if ((e = hipMalloc(&ctx->rows_d, sizeof(StencilRow) * std::max<size_t>(rows.size(), 1))) != hipSuccess ||
      (e = hipMalloc(&ctx->cell_count, sizeof(int32_t) * nc)) != hipSuccess ||
      (e = hipMalloc(&ctx->cell_start, sizeof(int32_t) * (nc + 1))) != hipSuccess ||
      (e = hipMalloc(&ctx->scan_tmp, sizeof(int32_t) * (ntiles + 1))) != hipSuccess ||
      (e = hipMalloc(&ctx->flag, sizeof(int32_t))) != hipSuccess ||
      (e = hipMalloc(&ctx->pairs_d, sizeof(int64_t))) != hipSuccess ||
      (e = hipMalloc(&ctx->elist_off, sizeof(int64_t) * (ne + 1))) != hipSuccess)
    return bail(e, "hipMalloc");
  if (nsys > 1) {
    if ((e = hipMalloc(&ctx->sys_d, sizeof(SysGrid) * nsys)) != hipSuccess ||
        (e = hipMalloc(&ctx->part_off_d, sizeof(int64_t) * (nsys + 1))) != hipSuccess)
      return bail(e, "hipMalloc");
    if ((e = hipMemcpy(ctx->sys_d, sg.data(), sizeof(SysGrid) * nsys, hipMemcpyHostToDevice)) != hipSuccess)
      return bail(e, "hipMemcpy");
  }
  if (!rows.empty()) {
    e = hipMemcpy(ctx->rows_d, rows.data(), sizeof(StencilRow) * rows.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(e, "hipMemcpy");
  }
  {  // FP32C coefficient table of (A, Bx) in r^2 (shared by all systems: same mu, alpha, xi)
    float meta[8];
    int32_t nint = 0;
    if (ggem_util_fp32c_table(c.one_eight_pi_mu, c.alpha, c.ksi, 0, nullptr, meta, &nint) != GGEM_OK ||
        nint > kTabMaxIntervals)
      return bail(hipErrorInvalidValue, "FP32C coefficient table");
    std::vector<float> tab(size_t(16) * nint);
    if (ggem_util_fp32c_table(c.one_eight_pi_mu, c.alpha, c.ksi, int32_t(tab.size()), tab.data(), meta, &nint) !=
        GGEM_OK)
      return bail(hipErrorInvalidValue, "FP32C coefficient table");
    if ((e = hipMalloc(&ctx->tab32, sizeof(float) * tab.size())) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMemcpy(ctx->tab32, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice)) != hipSuccess)
      return bail(e, "hipMemcpy");
    ctx->tab = CoefTab{meta[0], meta[1], meta[2], meta[3], meta[4], int32_t(meta[5]), int32_t(meta[6]), nint};
  }
  if (grow_scan64(ctx, ne + 1) != GGEM_OK) return bail(hipErrorOutOfMemory, "hipMalloc");
  *out = ctx;
  return GGEM_OK;
}

}  // namespace

extern "C" {

const char* ggem_last_error(const ggem_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return g_err.c_str();
}

int ggem_create(const ggem_params* params, ggem_ctx** out) { return create_impl(params, 1, out); }

int ggem_create_batch(const ggem_params* params, int32_t nsys, ggem_ctx** out) {
  if (nsys < 1) {
    if (out) *out = nullptr;
    return fail(nullptr, GGEM_ERR_INVALID_ARG, "nsys must be positive");
  }
  return create_impl(params, nsys, out);
}

int ggem_destroy(ggem_ctx* ctx) {
  if (!ctx) return GGEM_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_all(ctx);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return GGEM_OK;
}

int ggem_get_consts(const ggem_ctx* ctx, ggem_consts* out) {
  if (!ctx || !out) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null argument");
  *out = ctx->c;
  return GGEM_OK;
}

int ggem_set_precision(ggem_ctx* ctx, int32_t precision) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (precision != GGEM_FP64 && precision != GGEM_FP32C)
    return fail(ctx, GGEM_ERR_INVALID_ARG, "precision must be GGEM_FP64 or GGEM_FP32C");
  if (precision != ctx->p.precision) ctx->box_built = false;  // the staged records' layout is per precision
  ctx->p.precision = precision;
  return GGEM_OK;
}

int ggem_sync(ggem_ctx* ctx) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  return sync(ctx);
}

int ggem_set_particles(ggem_ctx* ctx, int64_t n, const double* xyz, const double* f) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (ctx->nsys > 1) return fail(ctx, GGEM_ERR_INVALID_ARG, "batched context: use ggem_batch_set_particles");
  return set_particles_impl(ctx, n, xyz, f);
}

int ggem_batch_set_particles(ggem_ctx* ctx, const int64_t* counts, const double* xyz, const double* f) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!counts) return fail(ctx, GGEM_ERR_INVALID_ARG, "null counts");
  std::vector<int64_t> off(ctx->nsys + 1, 0);
  for (int32_t k = 0; k < ctx->nsys; ++k) {
    if (counts[k] < 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "negative particle count");
    off[k + 1] = off[k] + counts[k];
  }
  const int64_t n = off[ctx->nsys];
  // element ids of the previous build belong to the previous partition
  const bool same = off == ctx->part_off;
  int rc = set_particles_impl(ctx, n, xyz, f);
  if (rc) return rc;
  if (!same) ctx->elem_valid = false;
  ctx->part_off = off;
  if (ctx->nsys == 1) return GGEM_OK;
  if (n > ctx->sys_cap) {
    if (ctx->sys_part) (void)hipFree(ctx->sys_part);
    if (ctx->sys_slot) (void)hipFree(ctx->sys_slot);
    ctx->sys_part = ctx->sys_slot = nullptr;
    ctx->sys_cap = 0;
    GGEM_HIP(ctx, hipMalloc(&ctx->sys_part, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->sys_slot, sizeof(int32_t) * n));
    ctx->sys_cap = n;
  }
  GGEM_HIP(ctx, hipMemcpyAsync(ctx->part_off_d, ctx->part_off.data(), sizeof(int64_t) * (ctx->nsys + 1),
                               hipMemcpyHostToDevice, ctx->stream));
  hipError_t e = launch_part_sys(ctx);
  if (e != hipSuccess) return hip_fail(ctx, e, "part_sys");
  return sync(ctx);
}

}  // extern "C"

namespace {

int set_particles_impl(ggem_ctx* ctx, int64_t n, const double* xyz, const double* f) {
  if (n < 0 || (n > 0 && !xyz)) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid particle arrays");
  if (n >= kMaxParticles) return fail(ctx, GGEM_ERR_CAPACITY, "particle count exceeds 2^26");
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  if (n != ctx->n) ctx->elem_valid = false;
  if (n > ctx->cap) {
    auto f_ = [](void* p) {
      if (p) (void)hipFree(p);
    };
    f_(ctx->x_in);
    f_(ctx->f_in);
    f_(ctx->elem);
    f_(ctx->perm);
    f_(ctx->slot_of);
    f_(ctx->slot_cell);
    f_(ctx->xs);
    f_(ctx->x4);
    f_(ctx->fs);
    f_(ctx->fs32);
    f_(ctx->fg32);
    f_(ctx->fg64);
    f_(ctx->u_loc);
    f_(ctx->u_out);
    f_(ctx->nbr_cnt);
    f_(ctx->slice_len);
    f_(ctx->slice_off);
    f_(ctx->pc32);
    f_(ctx->pcell);
    ctx->pc32 = nullptr;
    ctx->pcell = nullptr;
    ctx->x_in = ctx->f_in = ctx->xs = ctx->fs = ctx->u_loc = ctx->u_out = nullptr;
    ctx->elem = ctx->perm = ctx->slot_of = ctx->slot_cell = ctx->nbr_cnt = nullptr;
    ctx->x4 = nullptr;
    ctx->fs32 = nullptr;
    ctx->fg32 = nullptr;
    ctx->fg64 = nullptr;
    ctx->slice_len = ctx->slice_off = nullptr;
    ctx->cap = 0;
    const size_t n3 = size_t(3) * n;
    const size_t nsl = size_t(n + kSlice - 1) / kSlice;
    GGEM_HIP(ctx, hipMalloc(&ctx->x_in, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->f_in, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->elem, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->perm, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->slot_of, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->slot_cell, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->xs, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->x4, sizeof(double4) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->pc32, sizeof(float4) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->pcell, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->fs, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->fs32, sizeof(float) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->fg32, sizeof(float4) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->fg64, sizeof(double4) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->u_loc, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->u_out, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->nbr_cnt, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->slice_len, sizeof(int64_t) * (nsl + 1)));
    GGEM_HIP(ctx, hipMalloc(&ctx->slice_off, sizeof(int64_t) * (nsl + 2)));
    int rc = grow_scan64(ctx, int64_t(nsl) + 1);
    if (rc) return rc;
    ctx->cap = n;
    ctx->elem_valid = false;
  }
  if (n != ctx->n) {
    ctx->saved_valid = ctx->rout_valid = ctx->keep_d_valid = false;
    ctx->keep_nrhs = 0;
  }
  ctx->u_out_valid = false;
  ctx->n = n;
  ctx->built = false;
  ctx->slots_current = false;
  ctx->elist_valid = false;
  ctx->coef64_valid = false;
  ctx->box_built = false;
  ctx->pairs = -1;
  if (n > 0) {
    int rc = h2d(ctx, ctx->x_in, xyz, sizeof(double) * 3 * n);
    if (rc) return rc;
    if (f) rc = h2d(ctx, ctx->f_in, f, sizeof(double) * 3 * n);
    else GGEM_HIP(ctx, hipMemsetAsync(ctx->f_in, 0, sizeof(double) * 3 * n, ctx->stream));
    if (rc) return rc;
  }
  return sync(ctx);
}

}  // namespace

namespace {

// GGEM_COMPAT_STALE_LISTS after a build: a particle whose new list is
// empty keeps its previous one (point_particle.h:189-194); then every
// particle's effective list is stored by particle id in the reference's
// order (the lists the next build keeps, and the CSR export).
int keep_lists(ggem_ctx* ctx) {
  const int64_t n = ctx->n;
  int rc = ensure(ctx, &ctx->stale_slot, &ctx->stale_slot_cap, size_t(std::max<int64_t>(n, 1)));
  if (rc) return rc;
  const bool have = ctx->keep_n == n && ctx->keep_off != nullptr;
  hipError_t e = launch_stale_patch(ctx, have);
  if (e != hipSuccess) return hip_fail(ctx, e, "stale lists");
  int64_t v[2] = {0, 0};
  if ((rc = peek(ctx, {{ctx->flag, 4}, {ctx->pairs_d, 8}}, v))) return rc;
  ctx->stale_lists = v[0] != 0;
  ctx->pairs = v[1];
  if ((rc = ensure(ctx, &ctx->keep_cnt, &ctx->keep_cnt_cap, size_t(std::max<int64_t>(n, 1))))) return rc;
  if ((rc = ensure(ctx, &ctx->keep_off2, &ctx->keep_off2_cap, size_t(n + 1)))) return rc;
  if ((rc = ensure(ctx, &ctx->keep_ids2, &ctx->keep_ids2_cap, size_t(std::max<int64_t>(v[1], 1))))) return rc;
  if ((rc = grow_scan64(ctx, n + 1))) return rc;
  if ((e = launch_counts_by_particle(ctx, ctx->keep_cnt)) != hipSuccess ||
      (e = launch_scan64(ctx, ctx->keep_cnt, n, ctx->keep_off2)) != hipSuccess ||
      (e = launch_fill_keep(ctx, ctx->keep_off, ctx->keep_ids, ctx->keep_off2, ctx->keep_ids2)) != hipSuccess)
    return hip_fail(ctx, e, "kept lists");
  std::swap(ctx->keep_off, ctx->keep_off2);
  std::swap(ctx->keep_off_cap, ctx->keep_off2_cap);
  std::swap(ctx->keep_ids, ctx->keep_ids2);
  std::swap(ctx->keep_ids_cap, ctx->keep_ids2_cap);
  ctx->keep_n = n;
  return GGEM_OK;
}

}  // namespace

extern "C" {

int ggem_set_forces(ggem_ctx* ctx, const double* f) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!f && ctx->n > 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "null forces");
  if (ctx->n == 0) return GGEM_OK;
  int rc = h2d(ctx, ctx->f_in, f, sizeof(double) * 3 * ctx->n);
  if (rc) return rc;
  if (ctx->built) {
    hipError_t e = launch_gather_forces(ctx);
    if (e != hipSuccess) return hip_fail(ctx, e, "gather_forces");
  }
  return GGEM_OK;
}

int ggem_build_neighbors(ggem_ctx* ctx) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  const bool hyst = (ctx->p.compat_flags & GGEM_COMPAT_ELEM_HYSTERESIS) && ctx->elem_valid;
  ctx->built = false;
  ctx->elist_valid = false;
  ctx->coef64_valid = false;
  ctx->box_built = false;
  hipError_t e = launch_bin(ctx, hyst);
  if (e != hipSuccess) return hip_fail(ctx, e, "bin");
  if ((e = launch_scan(ctx)) != hipSuccess) return hip_fail(ctx, e, "scan");
  if ((e = launch_place(ctx)) != hipSuccess) return hip_fail(ctx, e, "place");
  if ((e = launch_gather_forces(ctx)) != hipSuccess) return hip_fail(ctx, e, "gather_forces");
  {
    int rc0 = ensure(ctx, &ctx->brick_recs, &ctx->brick_recs_cap,
                     size_t(std::max<int64_t>(ctx->grid.n_cells >> (3 * kBrickBits), 1)));
    if (rc0) return rc0;
  }
  if ((e = launch_count_lists(ctx)) != hipSuccess) return hip_fail(ctx, e, "count");
  // one synchronisation: the out-of-domain flag and the pair count (the
  // sliced layout is laid out and filled by ensure_ell when a kernel that
  // reads it is about to run)
  int64_t v[2] = {0, 0};
  int rc = peek(ctx, {{ctx->flag, 4}, {ctx->pairs_d, 8}}, v);
  if (rc) return rc;
  const int64_t flag = v[0], pairs = v[1];
  if (flag) {
    ctx->elem_valid = false;
    // point_mesh.C:602-612: a particle no element contains is fatal
    return fail(ctx, GGEM_ERR_OUT_OF_DOMAIN, "a particle lies outside the mesh");
  }
  ctx->pairs = pairs;
  ctx->ell_len = -1;
  ctx->ell_valid = false;
  ctx->elem_valid = true;
  ctx->built = true;
  ctx->slots_current = true;
  ctx->stale_pos = false;
  ctx->stale_lists = false;
  if ((ctx->p.compat_flags & GGEM_COMPAT_STALE_LISTS) && (rc = keep_lists(ctx))) {
    ctx->built = false;
    return rc;
  }
  return GGEM_OK;
}

// PointMesh<3>::reinit(false) (point_mesh.C:245-289): new positions (and
// forces, zero when NULL: the reinit zeroes them, :249-251) with the lists,
// the element ids and the element lists of the last build kept; every pair
// then takes the explicit minimum image, as reinit_neighbor_vector
// recomputes r_ij = point_vector(x_i, x_j) (:220-241).  xyz NULL: the
// context's current positions (after ggem_move_particles).
int ggem_update_positions(ggem_ctx* ctx, const double* xyz, const double* f) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!ctx->built) return fail(ctx, GGEM_ERR_NOT_READY, "no lists: call ggem_build_neighbors first");
  int rc = single_only(ctx, "ggem_update_positions");
  if (rc) return rc;
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  const size_t bytes = sizeof(double) * 3 * size_t(ctx->n);
  if (bytes) {
    if (xyz && (rc = h2d(ctx, ctx->x_in, xyz, bytes))) return rc;
    if (f) rc = h2d(ctx, ctx->f_in, f, bytes);
    else GGEM_HIP(ctx, hipMemsetAsync(ctx->f_in, 0, bytes, ctx->stream));
    if (rc) return rc;
  }
  hipError_t e = launch_update_slots(ctx);
  if (e == hipSuccess) e = launch_gather_forces(ctx);
  if (e != hipSuccess) return hip_fail(ctx, e, "update positions");
  ctx->slots_current = true;
  ctx->stale_pos = true;
  ctx->coef64_valid = false;
  ctx->box_built = false;
  ctx->u_out_valid = false;
  return xyz || f ? sync(ctx) : GGEM_OK;  // the host arrays may be reused on return
}

int ggem_get_elem_ids(ggem_ctx* ctx, int32_t* elem) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!elem) return fail(ctx, GGEM_ERR_INVALID_ARG, "null output");
  return d2h(ctx, elem, ctx->elem, sizeof(int32_t) * ctx->n);
}

int ggem_neighbor_pairs(ggem_ctx* ctx, int64_t* npairs) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (!npairs) return fail(ctx, GGEM_ERR_INVALID_ARG, "null output");
  *npairs = ctx->pairs;
  return GGEM_OK;
}

int ggem_get_neighbor_csr(ggem_ctx* ctx, int64_t* offsets, int32_t* ids) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_get_neighbor_csr"))) return rc;
  if (!offsets) return fail(ctx, GGEM_ERR_INVALID_ARG, "null offsets");
  Pools& P = pools(ctx);
  const int64_t n = ctx->n;
  if ((ctx->p.compat_flags & GGEM_COMPAT_STALE_LISTS) && ctx->keep_n == n) {  // the effective lists
    if ((rc = d2h(ctx, offsets, ctx->keep_off, sizeof(int64_t) * (n + 1)))) return rc;
    if (ids && offsets[n] > 0) return d2h(ctx, ids, ctx->keep_ids, sizeof(int32_t) * offsets[n]);
    return GGEM_OK;
  }
  if ((rc = ensure(ctx, &P.offs, &P.offs_cap, size_t(n + 1)))) return rc;
  hipError_t e = launch_count_pairs(ctx, P.offs + 1);
  if (e != hipSuccess) return hip_fail(ctx, e, "count");
  std::vector<int64_t> cnt(n + 1, 0);
  if (n > 0 && (rc = d2h(ctx, cnt.data() + 1, P.offs + 1, sizeof(int64_t) * n))) return rc;
  offsets[0] = 0;
  for (int64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + cnt[i + 1];
  if (!ids) return GGEM_OK;
  const int64_t nnz = offsets[n];
  if ((rc = ensure(ctx, &P.ids, &P.ids_cap, size_t(std::max<int64_t>(nnz, 1))))) return rc;
  if ((rc = h2d(ctx, P.offs, offsets, sizeof(int64_t) * (n + 1)))) return rc;
  if ((e = launch_fill_csr(ctx, P.offs, P.ids)) != hipSuccess) return hip_fail(ctx, e, "fill_csr");
  if (nnz > 0) return d2h(ctx, ids, P.ids, sizeof(int32_t) * nnz);
  return sync(ctx);
}

int ggem_get_elem_point_csr(ggem_ctx* ctx, int64_t e0, int64_t ne, int64_t* offsets, int32_t* ids) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_get_elem_point_csr"))) return rc;
  if (!offsets || e0 < 0 || ne < 0 || e0 + ne > ctx->grid.n_elem)
    return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid element range");
  Pools& P = pools(ctx);
  if ((rc = ensure(ctx, &P.offs, &P.offs_cap, size_t(ne + 1)))) return rc;
  hipError_t e = launch_elem_point_csr(ctx, e0, ne, P.offs, nullptr, 0);
  if (e != hipSuccess) return hip_fail(ctx, e, "elem_point count");
  std::vector<int64_t> cnt(ne + 1, 0);
  if (ne > 0 && (rc = d2h(ctx, cnt.data() + 1, P.offs + 1, sizeof(int64_t) * ne))) return rc;
  offsets[0] = 0;
  for (int64_t i = 0; i < ne; ++i) offsets[i + 1] = offsets[i] + cnt[i + 1];
  if (!ids) return GGEM_OK;
  const int64_t nnz = offsets[ne];
  if ((rc = ensure(ctx, &P.ids, &P.ids_cap, size_t(std::max<int64_t>(nnz, 1))))) return rc;
  if ((rc = h2d(ctx, P.offs, offsets, sizeof(int64_t) * (ne + 1)))) return rc;
  if ((e = launch_elem_point_csr(ctx, e0, ne, P.offs, P.ids, 1)) != hipSuccess)
    return hip_fail(ctx, e, "elem_point fill");
  if (nnz > 0) return d2h(ctx, ids, P.ids, sizeof(int32_t) * nnz);
  return sync(ctx);
}

int ggem_local_velocity_beads(ggem_ctx* ctx, double* u) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = ensure_tensors(ctx))) return rc;
  hipError_t e = launch_pair_beads(ctx, ctx->p.precision, EPI_LOCAL, nullptr, ctx->u_loc);
  if (e != hipSuccess) return hip_fail(ctx, e, "pair_beads");
  if (u && ctx->n > 0) return d2h(ctx, u, ctx->u_loc, sizeof(double) * 3 * ctx->n);
  return GGEM_OK;
}

int ggem_global_self_exclusion(ggem_ctx* ctx, double* uex) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  const int64_t len = 3 * ctx->n;
  ctx->u_out_valid = false;
  hipError_t e = launch_fill(ctx, len, 0.0, ctx->u_out);
  if (e == hipSuccess) e = launch_axpby(ctx, len, ctx->c.coeff_1, ctx->f_in, 0.0, ctx->u_out);
  if (e != hipSuccess) return hip_fail(ctx, e, "self_exclusion");
  if (uex && len > 0) return d2h(ctx, uex, ctx->u_out, sizeof(double) * len);
  return GGEM_OK;
}

int ggem_combine_velocity(ggem_ctx* ctx, const double* u_glob, double* U) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  const int64_t len = 3 * ctx->n;
  if ((rc = ensure_tensors(ctx))) return rc;
  Pools& P = pools(ctx);
  const double* ug = nullptr;
  if (u_glob && len > 0) {
    if ((rc = ensure(ctx, &P.io, &P.io_cap, size_t(len)))) return rc;
    if ((rc = h2d(ctx, P.io, u_glob, sizeof(double) * len))) return rc;
    ug = P.io;
  }
  hipError_t e = launch_pair_beads(ctx, ctx->p.precision, EPI_MOBILITY, ug, ctx->u_out);
  if (e != hipSuccess) return hip_fail(ctx, e, "pair_beads");
  ctx->u_out_valid = true;
  if (U && len > 0) return d2h(ctx, U, ctx->u_out, sizeof(double) * len);
  return GGEM_OK;
}

int ggem_set_fe_mesh(ggem_ctx* ctx, int64_t n_nodes, const int32_t* elem_nodes) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (ctx->nsys > 1) return fail(ctx, GGEM_ERR_INVALID_ARG, "ggem_set_fe_mesh needs a single-system context");
  const int64_t ne = ctx->grid.n_elem;
  if (n_nodes < 1 || n_nodes > INT32_MAX || !elem_nodes)
    return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid node count or connectivity");
  // each incidence is stored as the int32 number 20 e + l
  if (20 * ne > int64_t(INT32_MAX))
    return fail(ctx, GGEM_ERR_CAPACITY, "the FE mesh needs 20 * n_elem <= 2^31 - 1 (" + std::to_string(ne) +
                                            " elements)");
  for (int64_t k = 0; k < 20 * ne; ++k)
    if (elem_nodes[k] < 0 || elem_nodes[k] >= n_nodes) {
      char buf[128];
      std::snprintf(buf, sizeof(buf), "node %d of element %lld is out of range", int(k % 20), (long long)(k / 20));
      return fail(ctx, GGEM_ERR_INVALID_ARG, buf);
    }
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  for (void* p : {static_cast<void*>(ctx->fe_conn), static_cast<void*>(ctx->fe_inc_off),
                  static_cast<void*>(ctx->fe_inc), static_cast<void*>(ctx->fe_node_order),
                  static_cast<void*>(ctx->fe_cinc_off), static_cast<void*>(ctx->fe_cinc)})
    if (p) (void)hipFree(p);
  ctx->fe_conn = nullptr;
  ctx->fe_inc_off = nullptr;
  ctx->fe_inc = nullptr;
  ctx->fe_node_order = nullptr;
  ctx->fe_cinc_off = nullptr;
  ctx->fe_cinc = nullptr;
  ctx->fe_nodes = -1;
  // incidences of each node, (element, local node) ascending: the order in
  // which the nodal force sums the element vectors
  std::vector<int64_t> off(n_nodes + 1, 0);
  for (int64_t k = 0; k < 20 * ne; ++k) ++off[elem_nodes[k] + 1];
  for (int64_t v = 0; v < n_nodes; ++v) off[v + 1] += off[v];
  std::vector<int32_t> inc(20 * ne);
  {
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    for (int64_t k = 0; k < 20 * ne; ++k) inc[pos[elem_nodes[k]]++] = int32_t(k);
  }
  // nodes in the order of the 8 x 8 x 8-element brick of their first
  // element, so that the gather of the nodal force finds a brick's element
  // vectors in L2 for all of its nodes
  std::vector<int32_t> order(n_nodes);
  {
    const int nx = ctx->grid.n[0], ny = ctx->grid.n[1];
    const int64_t bx = (nx + 7) / 8, by = (ny + 7) / 8, bz = (ctx->grid.n[2] + 7) / 8;
    std::vector<int64_t> key(n_nodes), start(bx * by * bz + 1, 0);
    for (int64_t v = 0; v < n_nodes; ++v) {
      const int64_t e = off[v] < off[v + 1] ? inc[off[v]] / 20 : 0;
      const int64_t i = e % nx, j = (e / nx) % ny, k = e / (int64_t(nx) * ny);
      key[v] = (i >> 3) + bx * ((j >> 3) + by * (k >> 3));
      ++start[key[v] + 1];
    }
    for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
    for (int64_t v = 0; v < n_nodes; ++v) order[start[key[v]]++] = int32_t(v);
  }
  // Periodic condensation (constrain_element_vector, assemble_stokes.C:508,
  // with the periodic boundaries of copss.C:1294-1330 on the dof map): each
  // node sits at a point of the half-element lattice; a node on the max face
  // of a periodic axis is constrained to the node at its image with every
  // such coordinate moved to the min face, which then sums the incidences of
  // all its images in (element, local node) order, the order in which the
  // element loop adds them; the constrained node's row stays zero.
  std::vector<int64_t> coff;
  std::vector<int32_t> cinc;
  const bool any_periodic = ctx->grid.periodic[0] || ctx->grid.periodic[1] || ctx->grid.periodic[2];
  if (any_periodic) {
    const int nx = ctx->grid.n[0], ny = ctx->grid.n[1], nz = ctx->grid.n[2];
    const int64_t m0 = 2 * int64_t(nx) + 1, m1 = 2 * int64_t(ny) + 1, m2 = 2 * int64_t(nz) + 1;
    static const int kLat[20][3] = {{0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2},
                                    {0, 2, 2}, {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 1},
                                    {2, 2, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 2}, {1, 2, 2}, {0, 1, 2}};
    std::vector<int32_t> at(size_t(m0 * m1 * m2), -1);
    std::vector<int64_t> lat(n_nodes, -1);
    for (int64_t k = 0; k < 20 * ne; ++k) {
      const int64_t e = k / 20;
      const int l = int(k % 20);
      const int64_t p0 = 2 * (e % nx) + kLat[l][0], p1 = 2 * ((e / nx) % ny) + kLat[l][1],
                    p2 = 2 * (e / (int64_t(nx) * ny)) + kLat[l][2];
      const int64_t idx = p0 + m0 * (p1 + m1 * p2);
      const int32_t v = elem_nodes[k];
      if (at[idx] < 0) at[idx] = v;
      if (lat[v] < 0) lat[v] = idx;
    }
    std::vector<int32_t> rep(n_nodes);
    for (int64_t v = 0; v < n_nodes; ++v) {
      if (lat[v] < 0) {
        rep[v] = int32_t(v);
        continue;
      }
      int64_t p[3] = {lat[v] % m0, (lat[v] / m0) % m1, lat[v] / (m0 * m1)};
      const int64_t mx[3] = {m0 - 1, m1 - 1, m2 - 1};
      for (int d = 0; d < 3; ++d)
        if (ctx->grid.periodic[d] && p[d] == mx[d]) p[d] = 0;
      rep[v] = at[p[0] + m0 * (p[1] + m1 * p[2])];
    }
    coff.assign(n_nodes + 1, 0);
    for (int64_t k = 0; k < 20 * ne; ++k) ++coff[rep[elem_nodes[k]] + 1];
    for (int64_t v = 0; v < n_nodes; ++v) coff[v + 1] += coff[v];
    cinc.resize(20 * ne);
    std::vector<int64_t> pos(coff.begin(), coff.end() - 1);
    for (int64_t k = 0; k < 20 * ne; ++k) cinc[pos[rep[elem_nodes[k]]]++] = int32_t(k);
  }
  GGEM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->fe_conn), sizeof(int32_t) * 20 * ne));
  GGEM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->fe_inc), sizeof(int32_t) * 20 * ne));
  GGEM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->fe_node_order), sizeof(int32_t) * n_nodes));
  GGEM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->fe_inc_off), sizeof(int64_t) * (n_nodes + 1)));
  int rc = h2d(ctx, ctx->fe_conn, elem_nodes, sizeof(int32_t) * 20 * ne);
  if (rc) return rc;
  if ((rc = h2d(ctx, ctx->fe_inc, inc.data(), sizeof(int32_t) * 20 * ne))) return rc;
  if ((rc = h2d(ctx, ctx->fe_inc_off, off.data(), sizeof(int64_t) * (n_nodes + 1)))) return rc;
  if ((rc = h2d(ctx, ctx->fe_node_order, order.data(), sizeof(int32_t) * n_nodes))) return rc;
  if (any_periodic) {
    GGEM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->fe_cinc), sizeof(int32_t) * 20 * ne));
    GGEM_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&ctx->fe_cinc_off), sizeof(int64_t) * (n_nodes + 1)));
    if ((rc = h2d(ctx, ctx->fe_cinc, cinc.data(), sizeof(int32_t) * 20 * ne))) return rc;
    if ((rc = h2d(ctx, ctx->fe_cinc_off, coff.data(), sizeof(int64_t) * (n_nodes + 1)))) return rc;
  }
  if ((rc = ensure(ctx, &ctx->fe_u, &ctx->fe_u_cap, size_t(3 * n_nodes)))) return rc;
  ctx->fe_nodes = n_nodes;
  return sync(ctx);  // the host tables above are freed on return
}

namespace {
// U_glob of every bead into the device array out (particle order)
int interp_global(ggem_ctx* ctx, const double* node_u, double* out) {
  if (ctx->fe_nodes < 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "no FE mesh (call ggem_set_fe_mesh)");
  if (!node_u) return fail(ctx, GGEM_ERR_INVALID_ARG, "null nodal velocity");
  int rc = h2d(ctx, ctx->fe_u, node_u, sizeof(double) * 3 * ctx->fe_nodes);
  if (rc) return rc;
  hipError_t e = launch_interp_hex20(ctx, ctx->fe_conn, ctx->fe_u, out);
  if (e != hipSuccess) return hip_fail(ctx, e, "interp_hex20");
  return GGEM_OK;
}
}  // namespace

int ggem_global_velocity_beads(ggem_ctx* ctx, const double* node_u, double* u_glob) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_global_velocity_beads"))) return rc;
  if (!u_glob && ctx->n > 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "null output");
  const int64_t len = 3 * ctx->n;
  Pools& P = pools(ctx);
  if ((rc = ensure(ctx, &P.io2, &P.io2_cap, size_t(len)))) return rc;
  if ((rc = interp_global(ctx, node_u, P.io2))) return rc;
  return len > 0 ? d2h(ctx, u_glob, P.io2, sizeof(double) * len) : GGEM_OK;
}

int ggem_combine_velocity_fe(ggem_ctx* ctx, const double* node_u, double* U) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_combine_velocity_fe"))) return rc;
  const int64_t len = 3 * ctx->n;
  if ((rc = ensure_tensors(ctx))) return rc;
  Pools& P = pools(ctx);
  if ((rc = ensure(ctx, &P.io, &P.io_cap, size_t(len)))) return rc;
  if ((rc = interp_global(ctx, node_u, P.io))) return rc;
  hipError_t e = launch_pair_beads(ctx, ctx->p.precision, EPI_MOBILITY, P.io, ctx->u_out);
  if (e != hipSuccess) return hip_fail(ctx, e, "pair_beads");
  ctx->u_out_valid = true;
  if (U && len > 0) return d2h(ctx, U, ctx->u_out, sizeof(double) * len);
  return GGEM_OK;
}

int ggem_fix_forces(ggem_ctx* ctx, const ggem_fix_params* fp, double* f) {
  if (!ctx || !fp) return fail(ctx, GGEM_ERR_INVALID_ARG, "null argument");
  if (ctx->nsys > 1) return fail(ctx, GGEM_ERR_INVALID_ARG, "ggem_fix_forces is not available for a batched context");
  if (fp->chain_beads > 1 && (ctx->n % fp->chain_beads != 0 || !(fp->fene_bmax > 0)))
    return fail(ctx, GGEM_ERR_INVALID_ARG, "chains: n must be a multiple of chain_beads and fene_bmax > 0");
  if (fp->walls && !(fp->lj_sigma > 0 && fp->lj_rcut >= 0))
    return fail(ctx, GGEM_ERR_INVALID_ARG, "walls: lj_sigma > 0 and lj_rcut >= 0 required");
  if (ctx->n == 0) return GGEM_OK;
  hipError_t e = launch_fix_forces(ctx, fp->chain_beads, fp->fene_k, fp->fene_bmax, fp->walls, fp->lj_eps,
                                   fp->lj_sigma, fp->lj_rcut);
  if (e != hipSuccess) return hip_fail(ctx, e, "fix_forces");
  if (ctx->built && (e = launch_gather_forces(ctx)) != hipSuccess) return hip_fail(ctx, e, "gather_forces");
  if (f) return d2h(ctx, f, ctx->f_in, sizeof(double) * 3 * ctx->n);
  return GGEM_OK;
}

int ggem_spread_nodes(ggem_ctx* ctx, int32_t nqp_1d, const double* IF, double* F_node) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_spread_nodes"))) return rc;
  if (ctx->fe_nodes < 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "no FE mesh (call ggem_set_fe_mesh)");
  if (nqp_1d != 2 && nqp_1d != 3) return fail(ctx, GGEM_ERR_INVALID_ARG, "nqp_1d must be 2 or 3");
  if (!IF || !F_node) return fail(ctx, GGEM_ERR_INVALID_ARG, "null IF table or output");
  if ((rc = ensure_spread(ctx, nqp_1d))) return rc;
  const int n_u = 20, nqp = nqp_1d * nqp_1d * nqp_1d;
  const int64_t ne = ctx->grid.n_elem;
  Pools& P = pools(ctx);
  if ((rc = ensure(ctx, &P.ifd, &P.ifd_cap, size_t(n_u * nqp)))) return rc;
  if ((rc = ensure(ctx, &P.fe, &P.fe_cap, size_t(ne) * 3 * n_u))) return rc;
  if ((rc = h2d(ctx, P.ifd, IF, sizeof(double) * n_u * nqp))) return rc;
  hipError_t e = launch_spread(ctx, ctx->p.precision, nqp_1d, n_u, P.ifd, nullptr, P.fe, true);
  if (e != hipSuccess) return hip_fail(ctx, e, "spread");
  if ((e = launch_gather_nodes(ctx, P.fe, ctx->fe_u)) != hipSuccess) return hip_fail(ctx, e, "gather_nodes");
  return d2h(ctx, F_node, ctx->fe_u, sizeof(double) * 3 * ctx->fe_nodes);
}

int ggem_local_velocity_points(ggem_ctx* ctx, int64_t m, const double* pts, const int32_t* elem,
                               double* u) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_local_velocity_points"))) return rc;
  if (m < 0 || (m > 0 && (!pts || !u))) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid point arrays");
  if (m == 0) return GGEM_OK;
  if (elem) {
    for (int64_t q = 0; q < m; ++q)
      if (elem[q] < 0 || elem[q] >= ctx->grid.n_elem) {
        char buf[128];
        std::snprintf(buf, sizeof(buf), "invalid element id %d for point %lld", elem[q], (long long)q);
        return fail(ctx, GGEM_ERR_INVALID_ELEM, buf);  // ggem_stokes.C:369-380
      }
  }
  Pools& P = pools(ctx);
  if ((rc = ensure(ctx, &P.io, &P.io_cap, size_t(3 * m)))) return rc;
  if ((rc = ensure(ctx, &P.io2, &P.io2_cap, size_t(3 * m)))) return rc;
  if ((rc = ensure(ctx, &P.ids, &P.ids_cap, size_t(m)))) return rc;
  if ((rc = h2d(ctx, P.io, pts, sizeof(double) * 3 * m))) return rc;
  hipError_t e;
  if (elem) {
    if ((rc = h2d(ctx, P.ids, elem, sizeof(int32_t) * m))) return rc;
  } else {
    if ((e = launch_locate(ctx, m, P.io, P.ids)) != hipSuccess) return hip_fail(ctx, e, "locate");
    std::vector<int32_t> el(m);
    if ((rc = d2h(ctx, el.data(), P.ids, sizeof(int32_t) * m))) return rc;
    for (int64_t q = 0; q < m; ++q)
      if (el[q] < 0) return fail(ctx, GGEM_ERR_OUT_OF_DOMAIN, "a point lies outside the mesh");
  }
  if ((e = launch_pair_points(ctx, ctx->p.precision, m, P.io, P.ids, P.io2)) != hipSuccess)
    return hip_fail(ctx, e, "pair_points");
  return d2h(ctx, u, P.io2, sizeof(double) * 3 * m);
}

int ggem_spread(ggem_ctx* ctx, int32_t nqp_1d, int32_t n_u, const double* IF, double* rho, double* Fe) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_spread"))) return rc;
  if (nqp_1d != 2 && nqp_1d != 3) return fail(ctx, GGEM_ERR_INVALID_ARG, "nqp_1d must be 2 or 3");
  if (n_u < 1 || n_u > 27 || !IF) return fail(ctx, GGEM_ERR_INVALID_ARG, "n_u must be in [1, 27] with an IF table");
  if ((rc = ensure_spread(ctx, nqp_1d))) return rc;
  const int nqp = nqp_1d * nqp_1d * nqp_1d;
  const int64_t ne = ctx->grid.n_elem;
  Pools& P = pools(ctx);
  if ((rc = ensure(ctx, &P.ifd, &P.ifd_cap, size_t(n_u * nqp)))) return rc;
  if ((rc = ensure(ctx, &P.fe, &P.fe_cap, size_t(ne) * 3 * n_u))) return rc;
  if (rho && (rc = ensure(ctx, &P.rho, &P.rho_cap, size_t(ne) * nqp * 3))) return rc;
  if ((rc = h2d(ctx, P.ifd, IF, sizeof(double) * n_u * nqp))) return rc;
  hipError_t e = launch_spread(ctx, ctx->p.precision, nqp_1d, n_u, P.ifd, rho ? P.rho : nullptr, P.fe);
  if (e != hipSuccess) return hip_fail(ctx, e, "spread");
  if (rho && (rc = d2h(ctx, rho, P.rho, sizeof(double) * ne * nqp * 3))) return rc;
  if (Fe && (rc = d2h(ctx, Fe, P.fe, sizeof(double) * ne * 3 * n_u))) return rc;
  return GGEM_OK;
}

int ggem_set_noise(ggem_ctx* ctx, int32_t nrhs, const double* dw) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (nrhs < 1 || !dw) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid noise arguments");
  const int64_t len = int64_t(nrhs) * 3 * ctx->n;
  if (len > ctx->noise_cap) {
    if (ctx->noise) (void)hipFree(ctx->noise);
    ctx->noise = nullptr;
    ctx->noise_cap = 0;
    GGEM_HIP(ctx, hipMalloc(&ctx->noise, sizeof(double) * std::max<int64_t>(len, 1)));
    ctx->noise_cap = len;
  }
  ctx->noise_nrhs = nrhs;
  int rc = h2d(ctx, ctx->noise, dw, sizeof(double) * len);
  if (rc) return rc;
  return sync(ctx);
}

int ggem_mobility_apply(ggem_ctx* ctx, const double* f, double* u) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (f && (rc = ggem_set_forces(ctx, f))) return rc;
  return ggem_combine_velocity(ctx, nullptr, u);
}

int ggem_set_graph(ggem_ctx* ctx, int32_t enable) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  ctx->graph_on = enable != 0;
  return GGEM_OK;
}

int ggem_set_autotune(ggem_ctx* ctx, int32_t enable) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  ctx->tune.on = enable != 0;
  return GGEM_OK;
}

int ggem_tune_set(ggem_ctx* ctx, const char* signature, int32_t pick) {
  if (!ctx || !signature || pick < 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid tuning entry");
  ctx->tune.pick[signature] = pick;
  ctx->tune.ms.erase(signature);
  return GGEM_OK;
}

int ggem_tuning_report(ggem_ctx* ctx, char* buf, int64_t cap, int64_t* len) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  std::string out;
  for (const auto& kv : ctx->tune.pick) {
    out += kv.first + " " + std::to_string(kv.second);
    auto m = ctx->tune.ms.find(kv.first);
    if (m != ctx->tune.ms.end())
      for (float t : m->second) {
        char b[32];
        std::snprintf(b, sizeof(b), " %.5f", double(t));
        out += b;
      }
    out += "\n";
  }
  if (len) *len = int64_t(out.size());
  if (buf && cap > 0) {
    const size_t k = std::min<size_t>(out.size(), size_t(cap - 1));
    std::memcpy(buf, out.data(), k);
    buf[k] = 0;
  }
  return GGEM_OK;
}

int ggem_profile_enable(ggem_ctx* ctx, int32_t enable) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  ctx->prof.on = enable != 0;
  return GGEM_OK;
}

int ggem_profile_reset(ggem_ctx* ctx) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  prof_collect(ctx);
  ctx->prof.acc.clear();
  return GGEM_OK;
}

int ggem_profile_get(ggem_ctx* ctx, const char* kernel, double* total_ms, int64_t* launches) {
  if (!ctx || !kernel) return fail(ctx, GGEM_ERR_INVALID_ARG, "null argument");
  prof_collect(ctx);
  double ms = 0;
  int64_t cnt = 0;
  const bool all = std::string(kernel) == "*";
  for (auto& a : ctx->prof.acc)
    if (all || a.name == kernel) {
      ms += a.ms;
      cnt += a.launches;
    }
  if (total_ms) *total_ms = ms;
  if (launches) *launches = cnt;
  return GGEM_OK;
}

}  // extern "C"
