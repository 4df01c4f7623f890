// ggem_api.hip -- the C-ABI of libcopss_ggem.so: context life cycle, the
// host side of every entry point, the Chebyshev and Lanczos drivers and the
// event profiler.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ggem_internal.h"

namespace {

thread_local std::string g_err;  // errors without a context
const double PI = 3.14159265358979323846;

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

Store store_of(const ggem_ctx* ctx) {
  Store s;
  s.xs = ctx->xs;
  s.cell_start = ctx->cell_start;
  s.slot_cell = ctx->slot_cell;
  s.perm = ctx->perm;
  s.n = ctx->n;
  return s;
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
                  int32_t* ncells, std::string* why) {
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
      }
    }
  *ncells = int32_t(cells);
  return GGEM_OK;
}

template <class T>
int ensure(ggem_ctx* ctx, T** ptr, size_t* cap_elems, size_t elems) {
  if (*ptr && *cap_elems >= elems) return GGEM_OK;
  if (*ptr) {
    (void)hipFree(*ptr);
    *ptr = nullptr;
    *cap_elems = 0;
  }
  if (elems == 0) return GGEM_OK;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), elems * sizeof(T));
  if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
  *cap_elems = elems;
  return GGEM_OK;
}

// Device scratch pools (grown on demand, freed with the context).
struct Pools {
  double* vec = nullptr;
  size_t vec_cap = 0;
  double* io = nullptr;
  size_t io_cap = 0;
  double* io2 = nullptr;
  size_t io2_cap = 0;
  int32_t* ids = nullptr;
  size_t ids_cap = 0;
  int64_t* offs = nullptr;
  size_t offs_cap = 0;
  double* red = nullptr;
  size_t red_cap = 0;
  double* fe = nullptr;
  size_t fe_cap = 0;
  double* rho = nullptr;
  size_t rho_cap = 0;
  double* ifd = nullptr;
  size_t ifd_cap = 0;
};

Pools& pools(ggem_ctx* ctx) { return *reinterpret_cast<Pools*>(ctx->scratch); }

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
  f(ctx->fs);
  f(ctx->fs32);
  f(ctx->u_loc);
  f(ctx->u_out);
  f(ctx->cell_count);
  f(ctx->cell_start);
  f(ctx->scan_tmp);
  f(ctx->flag);
  f(ctx->pairs_d);
  f(ctx->noise);
  if (ctx->scratch) {
    Pools& P = pools(ctx);
    f(P.vec);
    f(P.io);
    f(P.io2);
    f(P.ids);
    f(P.offs);
    f(P.red);
    f(P.fe);
    f(P.rho);
    f(P.ifd);
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

int check_ready(ggem_ctx* ctx) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!ctx->built) return fail(ctx, GGEM_ERR_NOT_READY, "call ggem_build_neighbors after setting particles");
  return GGEM_OK;
}

int sync(ggem_ctx* ctx) {
  GGEM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GGEM_OK;
}

int d2h(ggem_ctx* ctx, void* host, const void* dev, size_t bytes) {
  GGEM_HIP(ctx, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
  return sync(ctx);
}

int h2d(ggem_ctx* ctx, void* dev, const void* host, size_t bytes) {
  GGEM_HIP(ctx, hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  return GGEM_OK;
}

}  // namespace

extern "C" {

const char* ggem_last_error(const ggem_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return g_err.c_str();
}

int ggem_create(const ggem_params* params, ggem_ctx** out) {
  if (!out) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null output pointer");
  *out = nullptr;
  if (!params) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null params");
  ggem_consts c;
  std::string why;
  int rc = derive(params, &c, &why);
  if (rc) return fail(nullptr, rc, why);
  std::vector<StencilRow> rows;
  int32_t ncells = 0;
  rc = build_stencil(params, &c, &rows, &ncells, &why);
  if (rc) return fail(nullptr, rc, why);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(nullptr, GGEM_ERR_NO_DEVICE, "no HIP device visible");
  if (params->device < 0 || params->device >= ndev)
    return fail(nullptr, GGEM_ERR_INVALID_ARG, "device ordinal out of range");

  ggem_ctx* ctx = new ggem_ctx();
  ctx->p = *params;
  ctx->c = c;
  ctx->c.stencil_rows = int32_t(rows.size());
  ctx->c.stencil_cells = ncells;
  ctx->device = params->device;
  ctx->scratch = new Pools();
  for (int d = 0; d < 3; ++d) {
    ctx->grid.bmin[d] = params->box_min[d];
    ctx->grid.L[d] = c.L[d];
    ctx->grid.h[d] = c.h[d];
    ctx->grid.half_L[d] = c.L[d] / 2.0;
    ctx->grid.n[d] = params->n_elem[d];
    ctx->grid.periodic[d] = params->periodic[d] ? 1 : 0;
  }
  ctx->grid.n_elem = int64_t(params->n_elem[0]) * params->n_elem[1] * params->n_elem[2];
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
  ctx->rows_h = rows;
  ctx->n_rows = int32_t(rows.size());

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
  const int64_t ne = ctx->grid.n_elem;
  const int64_t ntiles = (ne + 2047) / 2048;
  if ((e = hipMalloc(&ctx->rows_d, sizeof(StencilRow) * std::max<size_t>(rows.size(), 1))) != hipSuccess ||
      (e = hipMalloc(&ctx->cell_count, sizeof(int32_t) * ne)) != hipSuccess ||
      (e = hipMalloc(&ctx->cell_start, sizeof(int32_t) * (ne + 1))) != hipSuccess ||
      (e = hipMalloc(&ctx->scan_tmp, sizeof(int32_t) * (ntiles + 1))) != hipSuccess ||
      (e = hipMalloc(&ctx->flag, sizeof(int32_t))) != hipSuccess ||
      (e = hipMalloc(&ctx->pairs_d, sizeof(int64_t))) != hipSuccess)
    return bail(e, "hipMalloc");
  if (!rows.empty()) {
    e = hipMemcpy(ctx->rows_d, rows.data(), sizeof(StencilRow) * rows.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(e, "hipMemcpy");
  }
  *out = ctx;
  return GGEM_OK;
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
  ctx->p.precision = precision;
  return GGEM_OK;
}

int ggem_sync(ggem_ctx* ctx) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  return sync(ctx);
}

int ggem_set_particles(ggem_ctx* ctx, int64_t n, const double* xyz, const double* f) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (n < 0 || (n > 0 && !xyz)) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid particle arrays");
  if (n >= INT32_MAX) return fail(ctx, GGEM_ERR_CAPACITY, "particle count exceeds int32");
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
    f_(ctx->fs);
    f_(ctx->fs32);
    f_(ctx->u_loc);
    f_(ctx->u_out);
    ctx->x_in = ctx->f_in = ctx->xs = ctx->fs = ctx->u_loc = ctx->u_out = nullptr;
    ctx->elem = ctx->perm = ctx->slot_of = ctx->slot_cell = nullptr;
    ctx->fs32 = nullptr;
    ctx->cap = 0;
    const size_t n3 = size_t(3) * n;
    GGEM_HIP(ctx, hipMalloc(&ctx->x_in, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->f_in, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->elem, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->perm, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->slot_of, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->slot_cell, sizeof(int32_t) * n));
    GGEM_HIP(ctx, hipMalloc(&ctx->xs, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->fs, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->fs32, sizeof(float) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->u_loc, sizeof(double) * n3));
    GGEM_HIP(ctx, hipMalloc(&ctx->u_out, sizeof(double) * n3));
    ctx->cap = n;
    ctx->elem_valid = false;
  }
  ctx->n = n;
  ctx->built = false;
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

int ggem_set_forces(ggem_ctx* ctx, const double* f) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!f && ctx->n > 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "null forces");
  if (ctx->n == 0) return GGEM_OK;
  int rc = h2d(ctx, ctx->f_in, f, sizeof(double) * 3 * ctx->n);
  if (rc) return rc;
  if (ctx->built) {
    hipError_t e = launch_gather_forces(ctx, ctx->f_in, ctx->fs, ctx->fs32, 1, 0, 0);
    if (e != hipSuccess) return hip_fail(ctx, e, "gather_forces");
  }
  return GGEM_OK;
}

int ggem_build_neighbors(ggem_ctx* ctx) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  const bool hyst = (ctx->p.compat_flags & GGEM_COMPAT_ELEM_HYSTERESIS) && ctx->elem_valid;
  hipError_t e = launch_bin(ctx, hyst);
  if (e != hipSuccess) return hip_fail(ctx, e, "bin");
  int32_t flag = 0;
  int rc = d2h(ctx, &flag, ctx->flag, sizeof(int32_t));
  if (rc) return rc;
  if (flag) {
    ctx->elem_valid = false;
    ctx->built = false;
    // point_mesh.C:602-612: a particle no element contains is fatal
    return fail(ctx, GGEM_ERR_OUT_OF_DOMAIN, "a particle lies outside the mesh");
  }
  if ((e = launch_scan(ctx)) != hipSuccess) return hip_fail(ctx, e, "scan");
  if ((e = launch_place(ctx)) != hipSuccess) return hip_fail(ctx, e, "place");
  if ((e = launch_gather_forces(ctx, ctx->f_in, ctx->fs, ctx->fs32, 1, 0, 0)) != hipSuccess)
    return hip_fail(ctx, e, "gather_forces");
  if ((e = launch_count_pairs(ctx, ctx->pairs_d, nullptr)) != hipSuccess) return hip_fail(ctx, e, "count");
  int64_t pairs = 0;
  rc = d2h(ctx, &pairs, ctx->pairs_d, sizeof(int64_t));
  if (rc) return rc;
  ctx->pairs = pairs;
  ctx->elem_valid = true;
  ctx->built = true;
  return GGEM_OK;
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
  if (!offsets) return fail(ctx, GGEM_ERR_INVALID_ARG, "null offsets");
  Pools& P = pools(ctx);
  const int64_t n = ctx->n;
  if ((rc = ensure(ctx, &P.offs, &P.offs_cap, size_t(n + 1)))) return rc;
  hipError_t e = launch_count_pairs(ctx, ctx->pairs_d, P.offs + 1);
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
  hipError_t e = launch_pair_beads(ctx, ctx->p.precision, EPI_LOCAL, nullptr, ctx->u_loc);
  if (e != hipSuccess) return hip_fail(ctx, e, "pair_beads");
  if (u && ctx->n > 0) return d2h(ctx, u, ctx->u_loc, sizeof(double) * 3 * ctx->n);
  return GGEM_OK;
}

int ggem_global_self_exclusion(ggem_ctx* ctx, double* uex) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  const int64_t len = 3 * ctx->n;
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
  Pools& P = pools(ctx);
  const double* ug = nullptr;
  if (u_glob && len > 0) {
    if ((rc = ensure(ctx, &P.io, &P.io_cap, size_t(len)))) return rc;
    if ((rc = h2d(ctx, P.io, u_glob, sizeof(double) * len))) return rc;
    ug = P.io;
  }
  hipError_t e = launch_pair_beads(ctx, ctx->p.precision, EPI_MOBILITY, ug, ctx->u_out);
  if (e != hipSuccess) return hip_fail(ctx, e, "pair_beads");
  if (U && len > 0) return d2h(ctx, U, ctx->u_out, sizeof(double) * len);
  return GGEM_OK;
}

int ggem_local_velocity_points(ggem_ctx* ctx, int64_t m, const double* pts, const int32_t* elem,
                               double* u) {
  int rc = check_ready(ctx);
  if (rc) return rc;
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
  if (nqp_1d != 2 && nqp_1d != 3) return fail(ctx, GGEM_ERR_INVALID_ARG, "nqp_1d must be 2 or 3");
  if (n_u < 1 || n_u > 27 || !IF) return fail(ctx, GGEM_ERR_INVALID_ARG, "n_u must be in [1, 27] with an IF table");
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

}  // extern "C"

// ---------------------------------------------------------------------------
// Chebyshev driver
namespace {

// Chebyshev::chebyshev_coefficients, "Gauss" nodes (chebyshev.C:44-151)
void cheb_coefficients(int N, double da, double db, std::vector<double>* coef) {
  const double RN = double(N);
  std::vector<double> x(N + 1), w(N + 1);
  for (int i = 0; i <= N; ++i) {
    const double j = double(i);
    x[i] = std::cos((2. * j + 1.) * PI / (2. * RN + 2.));
    w[i] = PI / (RN + 1.);
  }
  coef->assign(N + 1, 0.0);
  for (int k = 0; k <= N; ++k) {
    double s = 0.0;
    for (int j = 0; j <= N; ++j) {
      const double tmpt = PI * double(2. * j + 1.) * double(k) / (2. * RN + 2.);
      const double fj = 1.0 / std::sqrt((x[j] - db) / da);
      s += w[j] * std::cos(tmpt) * fj;
    }
    (*coef)[k] = s * (2. / (PI * (k == 0 ? 2.0 : 1.0)));
  }
}

// Mobility application on slot-ordered vectors: device operator or a host
// operator reached through the particle-order layout.
struct Applier {
  ggem_ctx* ctx;
  const ggem_op* op;
  int precision;
  double* io_dev;  // 3n particle-order staging (host op)
  std::vector<double> hf, hu;
  int operator()(int nrhs, const double* x, double* v, const ChebEpi& epi) {
    if (!op) {
      hipError_t e = launch_apply_slots(ctx, precision, nrhs, x, v, epi);
      return e == hipSuccess ? GGEM_OK : hip_fail(ctx, e, "apply");
    }
    const int64_t n = ctx->n, len = 3 * n;
    hf.resize(len);
    hu.resize(len);
    for (int r = 0; r < nrhs; ++r) {
      hipError_t e = launch_permute_out(ctx, 1, x + r * len, io_dev);
      if (e != hipSuccess) return hip_fail(ctx, e, "permute");
      int rc = d2h(ctx, hf.data(), io_dev, sizeof(double) * len);
      if (rc) return rc;
      if (op->apply(op->user, hf.data(), hu.data()) != 0)
        return fail(ctx, GGEM_ERR_INVALID_ARG, "operator callback failed");
      if ((rc = h2d(ctx, io_dev, hu.data(), sizeof(double) * len))) return rc;
      double* vr = v + r * len;
      const double* xr = x + r * len;
      // v_r = M x_r, then the epilogue on the device vectors
      if (epi.kind == 0) {
        e = launch_permute_in(ctx, 1, io_dev, vr);
      } else {
        // M x lands in the second half of the staging buffer; v may alias
        // x_{k-1}, so it is written only after x_{k-1} has been consumed.
        double* mx = io_dev + len;
        e = launch_permute_in(ctx, 1, io_dev, mx);
        if (e == hipSuccess) {
          if (epi.kind == 1) {
            e = launch_axpby(ctx, len, epi.db, xr, epi.da, mx);  // db x + da M x
          } else {
            e = launch_axpby(ctx, len, 2. * epi.db, xr, 2. * epi.da, mx);  // 2db x + 2da M x
            if (e == hipSuccess) e = launch_axpby(ctx, len, -1.0, epi.xm1 + r * len, 1.0, mx);
            if (e == hipSuccess) e = launch_axpby(ctx, len, epi.c_host, mx, 1.0, epi.y + r * len);
          }
          if (e == hipSuccess)
            e = hipMemcpyAsync(vr, mx, sizeof(double) * len, hipMemcpyDeviceToDevice, ctx->stream);
        }
      }
      if (e != hipSuccess) return hip_fail(ctx, e, "host operator epilogue");
    }
    return GGEM_OK;
  }
};

int dot_host(ggem_ctx* ctx, int nrhs, const double* a, const double* b, double* out) {
  Pools& P = pools(ctx);
  int rc = ensure(ctx, &P.red, &P.red_cap, size_t(nrhs) * 513);
  if (rc) return rc;
  hipError_t e = launch_dot(ctx, nrhs, a, b, P.red + size_t(nrhs) * 512, P.red);
  if (e != hipSuccess) return hip_fail(ctx, e, "dot");
  return d2h(ctx, out, P.red + size_t(nrhs) * 512, sizeof(double) * nrhs);
}

}  // namespace

extern "C" int ggem_cheb_brownian(ggem_ctx* ctx, const ggem_op* op, int32_t N, double eig_min,
                                  double eig_max, double tol, int32_t mode, int32_t nrhs,
                                  const double* dw, double* y, double* disp, ggem_cheb_info* info) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (N < 1 || nrhs < 1 || !(eig_max > eig_min) || (mode != GGEM_CHEB_REFERENCE && mode != GGEM_CHEB_FIXED))
    return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid Chebyshev arguments");
  if (!dw && (ctx->noise_nrhs != nrhs || !ctx->noise))
    return fail(ctx, GGEM_ERR_NOT_READY, "no device noise of this nrhs: call ggem_set_noise");
  if (op && !op->apply) return fail(ctx, GGEM_ERR_INVALID_ARG, "operator without apply");
  const int64_t n = ctx->n, len = 3 * n;
  if (n == 0) return GGEM_OK;
  const double d_eig = eig_max - eig_min;
  const double da = 2.0 / d_eig;
  const double db = -(eig_max + eig_min) / d_eig;
  std::vector<double> X;
  cheb_coefficients(N, da, db, &X);

  // device vectors [nrhs][3][n] by slot: x0, x1, y, dwm, plus staging
  Pools& P = pools(ctx);
  const size_t blk = size_t(nrhs) * len;
  if ((rc = ensure(ctx, &P.vec, &P.vec_cap, 4 * blk))) return rc;
  if ((rc = ensure(ctx, &P.io, &P.io_cap, std::max(blk, size_t(2 * len))))) return rc;
  double* x0 = P.vec;
  double* x1 = P.vec + blk;
  double* yv = P.vec + 2 * blk;
  double* dwm = P.vec + 3 * blk;
  const double* dw_dev = ctx->noise;
  if (dw) {
    if ((rc = h2d(ctx, P.io, dw, sizeof(double) * blk))) return rc;
    dw_dev = P.io;
  }
  hipError_t e = launch_permute_in(ctx, nrhs, dw_dev, x0);
  if (e != hipSuccess) return hip_fail(ctx, e, "permute");
  Applier apply{ctx, op, ctx->p.precision, P.io, {}, {}};

  std::vector<ggem_cheb_info> inf(nrhs);
  for (auto& i : inf) i = ggem_cheb_info{0, 1, 0, 0, 0.0, 0.0};
  ChebEpi plain{0, 0.0, 0.0, nullptr, 0.0, nullptr, nullptr};

  if (mode == GGEM_CHEB_FIXED) {
    ChebEpi init{1, da, db, nullptr, 0.0, nullptr, nullptr};
    if ((rc = apply(nrhs, x0, x1, init))) return rc;  // x1 = db x0 + da M x0
    if ((e = launch_fill(ctx, blk, 0.0, yv)) != hipSuccess) return hip_fail(ctx, e, "fill");
    if ((e = launch_axpby(ctx, blk, X[0], x0, 1.0, yv)) != hipSuccess) return hip_fail(ctx, e, "axpy");
    if ((e = launch_axpby(ctx, blk, X[1], x1, 1.0, yv)) != hipSuccess) return hip_fail(ctx, e, "axpy");
    for (int i = 1; i < N; ++i) {
      // x_{k+1} = 2da M x_k + 2db x_k - x_{k-1} into x0's storage; y += c x_{k+1}
      ChebEpi st{2, da, db, nullptr, X[i + 1], x0, yv};
      if ((rc = apply(nrhs, x1, x0, st))) return rc;
      std::swap(x0, x1);
    }
    for (auto& i : inf) {
      i.converged = 1;
      i.terms = N;
      i.applies = N + (disp ? 1 : 0);
    }
  } else {
    // reference-compatible: one right-hand side at a time
    std::vector<double> h(1);
    for (int r = 0; r < nrhs; ++r) {
      double* a0 = x0 + r * len;
      double* a1 = x1 + r * len;
      double* ay = yv + r * len;
      double* am = dwm + r * len;
      int applies = 0;
      if ((rc = apply(1, a0, a1, plain))) return rc;  // x1 = M x0
      ++applies;
      double aux2, aux1, aux3, aux4;
      if ((rc = dot_host(ctx, 1, a1, a0, &aux2))) return rc;
      if ((e = launch_axpby(ctx, len, db, a0, da, a1)) != hipSuccess) return hip_fail(ctx, e, "axpby");
      if ((e = launch_fill(ctx, len, 0.0, ay)) != hipSuccess) return hip_fail(ctx, e, "fill");
      if ((e = launch_axpby(ctx, len, X[0], a0, 1.0, ay)) != hipSuccess) return hip_fail(ctx, e, "axpy");
      if ((rc = dot_host(ctx, 1, ay, ay, &aux4))) return rc;
      aux4 = std::sqrt(aux4);
      if ((e = launch_axpby(ctx, len, X[1], a1, 1.0, ay)) != hipSuccess) return hip_fail(ctx, e, "axpy");
      if ((rc = dot_host(ctx, 1, a1, a1, &aux3))) return rc;
      aux3 = std::sqrt(aux3) * std::fabs(X[1]);
      if ((rc = apply(1, ay, am, plain))) return rc;
      ++applies;
      if ((rc = dot_host(ctx, 1, am, am, &aux1))) return rc;
      double err1 = 0, err2 = 0;
      if (aux2 != 0.0) err1 = std::sqrt(std::fabs(aux1 - aux2) / aux2);
      if (aux4 != 0.0) err2 = aux3 / aux4;
      bool conv = false;
      int terms = 1;
      for (int i = 1; i < N; ++i) {
        if ((rc = dot_host(ctx, 1, ay, ay, &aux4))) return rc;
        aux4 = std::sqrt(aux4);
        ChebEpi st{2, da, db, nullptr, X[i + 1], a0, ay};
        if ((rc = apply(1, a1, a0, st))) return rc;  // a0 <- x_{k+1}; y += c x_{k+1}
        ++applies;
        terms = i + 1;
        if ((rc = dot_host(ctx, 1, a0, a0, &aux3))) return rc;
        aux3 = std::sqrt(aux3) * std::fabs(X[i + 1]);
        if ((rc = apply(1, ay, am, plain))) return rc;
        ++applies;
        if ((rc = dot_host(ctx, 1, am, am, &aux1))) return rc;
        if (aux2 != 0.0) err1 = std::sqrt(std::fabs(aux1 - aux2) / aux2);
        if (aux4 != 0.0) err2 = aux3 / aux4;
        if (err1 <= tol) {
          conv = true;
          break;
        }
        if (err2 <= 1E-9) break;
        std::swap(a0, a1);
      }
      inf[r] = ggem_cheb_info{conv ? 1 : 0, terms, applies + (disp ? 1 : 0), 0, err1, err2};
    }
  }
  if (y) {
    if ((e = launch_permute_out(ctx, nrhs, yv, P.io)) != hipSuccess) return hip_fail(ctx, e, "permute");
    if ((rc = d2h(ctx, y, P.io, sizeof(double) * blk))) return rc;
  }
  if (disp) {  // hi_ewald: disp = M y (copss.C:1686)
    if ((rc = apply(nrhs, yv, dwm, plain))) return rc;
    if ((e = launch_permute_out(ctx, nrhs, dwm, P.io)) != hipSuccess) return hip_fail(ctx, e, "permute");
    if ((rc = d2h(ctx, disp, P.io, sizeof(double) * blk))) return rc;
  }
  if (info)
    for (int r = 0; r < nrhs; ++r) info[r] = inf[r];
  return sync(ctx);
}

// ---------------------------------------------------------------------------
// Lanczos extreme eigenvalues
namespace {

// number of eigenvalues of the symmetric tridiagonal (a, b) below x (Sturm)
int sturm_count(const std::vector<double>& a, const std::vector<double>& b, double x) {
  int cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < a.size(); ++i) {
    const double bb = i ? b[i - 1] * b[i - 1] : 0.0;
    q = a[i] - x - (i ? bb / q : 0.0);
    if (q == 0.0) q = 1e-300;
    if (q < 0) ++cnt;
  }
  return cnt;
}

double tridiag_eig(const std::vector<double>& a, const std::vector<double>& b, int k) {
  // k-th smallest eigenvalue (0-based) by bisection on the Gershgorin interval
  double lo = 1e300, hi = -1e300;
  for (size_t i = 0; i < a.size(); ++i) {
    const double r = (i ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < a.size() ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid == lo || mid == hi) break;
    if (sturm_count(a, b, mid) > k) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

}  // namespace

extern "C" int ggem_lanczos_bounds(ggem_ctx* ctx, const ggem_op* op, int32_t iters, uint64_t seed,
                                   double* lmin, double* lmax) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (iters < 1 || !lmin || !lmax) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid Lanczos arguments");
  if (op && !op->apply) return fail(ctx, GGEM_ERR_INVALID_ARG, "operator without apply");
  const int64_t len = 3 * ctx->n;
  if (len == 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "no particles");
  Pools& P = pools(ctx);
  if ((rc = ensure(ctx, &P.vec, &P.vec_cap, 4 * size_t(len)))) return rc;
  if ((rc = ensure(ctx, &P.io, &P.io_cap, size_t(2 * len)))) return rc;
  double* vprev = P.vec;
  double* vcur = P.vec + len;
  double* w = P.vec + 2 * len;
  std::vector<double> h(len);
  std::mt19937_64 gen(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  double s = 0;
  for (auto& v : h) {
    v = nd(gen);
    s += v * v;
  }
  s = 1.0 / std::sqrt(s);
  for (auto& v : h) v *= s;
  if ((rc = h2d(ctx, vcur, h.data(), sizeof(double) * len))) return rc;
  hipError_t e = launch_fill(ctx, len, 0.0, vprev);
  if (e != hipSuccess) return hip_fail(ctx, e, "fill");
  Applier apply{ctx, op, ctx->p.precision, P.io, {}, {}};
  ChebEpi plain{0, 0.0, 0.0, nullptr, 0.0, nullptr, nullptr};
  std::vector<double> alpha, beta;
  double bprev = 0.0;
  const int kmax = int(std::min<int64_t>(iters, len));
  for (int j = 0; j < kmax; ++j) {
    if ((rc = apply(1, vcur, w, plain))) return rc;
    double a;
    if ((rc = dot_host(ctx, 1, w, vcur, &a))) return rc;
    alpha.push_back(a);
    if ((e = launch_axpby(ctx, len, -a, vcur, 1.0, w)) != hipSuccess) return hip_fail(ctx, e, "axpby");
    if ((e = launch_axpby(ctx, len, -bprev, vprev, 1.0, w)) != hipSuccess) return hip_fail(ctx, e, "axpby");
    double b2;
    if ((rc = dot_host(ctx, 1, w, w, &b2))) return rc;
    const double b = std::sqrt(b2);
    if (j + 1 == kmax || b <= 1e-12 * std::fabs(a)) break;
    beta.push_back(b);
    // vprev <- vcur; vcur <- w / b
    std::swap(vprev, vcur);
    if ((e = launch_fill(ctx, len, 0.0, vcur)) != hipSuccess) return hip_fail(ctx, e, "fill");
    if ((e = launch_axpby(ctx, len, 1.0 / b, w, 0.0, vcur)) != hipSuccess) return hip_fail(ctx, e, "axpby");
    bprev = b;
  }
  beta.resize(alpha.size() > 0 ? alpha.size() - 1 : 0);
  *lmin = tridiag_eig(alpha, beta, 0);
  *lmax = tridiag_eig(alpha, beta, int(alpha.size()) - 1);
  return GGEM_OK;
}

// ---------------------------------------------------------------------------
extern "C" {

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
  for (auto& a : ctx->prof.acc)
    if (a.name == kernel) {
      ms = a.ms;
      cnt = a.launches;
    }
  if (total_ms) *total_ms = ms;
  if (launches) *launches = cnt;
  return GGEM_OK;
}

}  // extern "C"
