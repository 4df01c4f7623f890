// ggem_cheb.hip -- the Brownian-step drivers of the C-ABI: the Chebyshev
// approximation of M^{-1/2} dw with the reference's estimators and stopping
// rule or at fixed order, and the Lanczos extreme-eigenvalue estimate.
//
// Vectors live on the device in slot order (FP64, [nrhs][3][n]) with a
// per-slot gather record for the neighbour reads of the next application;
// the mobility application and the three-term update are one kernel
// (kernels_pair.hip).
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "ggem_host.h"

using namespace ggem;

namespace {

const double PI = 3.14159265358979323846;

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

// A set of nrhs slot vectors carved out of the pools.
SlotVec slot_vec(Pools& P, int which, int nrhs, int64_t n, int precision) {
  const size_t blk = size_t(nrhs) * 3 * n, gblk = size_t(nrhs) * n;
  SlotVec v;
  v.soa = P.vec + which * blk;
  v.g32 = precision == GGEM_FP32C ? P.g32 + which * gblk : nullptr;
  v.g64 = precision == GGEM_FP64 ? P.g64 + which * gblk : nullptr;
  return v;
}

SlotVec offset(const SlotVec& v, int r, int64_t n) {
  return SlotVec{v.soa + int64_t(r) * 3 * n, v.g32 ? v.g32 + int64_t(r) * n : nullptr,
                 v.g64 ? v.g64 + int64_t(r) * n : nullptr};
}

int reserve(ggem_ctx* ctx, int nvec, int nrhs) {
  Pools& P = pools(ctx);
  const int64_t n = ctx->n;
  int rc = ensure(ctx, &P.vec, &P.vec_cap, size_t(nvec) * nrhs * 3 * n);
  if (!rc) rc = ensure(ctx, &P.g32, &P.g32_cap, size_t(nvec) * nrhs * n);
  if (!rc) rc = ensure(ctx, &P.g64, &P.g64_cap, size_t(nvec) * nrhs * n);
  if (!rc) rc = ensure(ctx, &P.io, &P.io_cap, std::max(size_t(nrhs) * 3 * n, size_t(6 * n)));
  return rc;
}

// the kept Brownian vectors (GGEM_CHEB_KEEP) for nrhs operands
int ensure_keep(ggem_ctx* ctx, int nrhs) {
  const size_t len = size_t(nrhs) * 3 * size_t(ctx->n);
  if (ctx->keep_cap >= len && ctx->keep_y) return GGEM_OK;
  size_t c1 = ctx->keep_cap, c2 = ctx->keep_cap;
  int rc = ensure(ctx, &ctx->keep_y, &c1, len);
  if (!rc) rc = ensure(ctx, &ctx->keep_d, &c2, len);
  ctx->keep_cap = rc ? 0 : len;
  ctx->keep_nrhs = 0;
  ctx->keep_d_valid = false;
  return rc;
}

int refresh(ggem_ctx* ctx, int nrhs, const SlotVec& v) {
  hipError_t e = launch_make_gather(ctx, nrhs, v.soa, v.g32, v.g64);
  return e == hipSuccess ? GGEM_OK : hip_fail(ctx, e, "gather records");
}

// v = M x (+ epilogue), on the device or through a host operator.
struct Applier {
  ggem_ctx* ctx;
  const ggem_op* op;
  int precision;
  std::vector<double> hf, hu;
  int operator()(int nrhs, const SlotVec& x, const SlotVec& v, const ChebEpi& epi) {
    if (!op) {
      hipError_t e = launch_apply(ctx, precision, nrhs, x, v, epi);
      return e == hipSuccess ? GGEM_OK : hip_fail(ctx, e, "apply");
    }
    const int64_t n = ctx->n, len = 3 * n;
    double* io = pools(ctx).io;  // 6n: particle-order staging and M x by slot
    hf.resize(len);
    hu.resize(len);
    for (int r = 0; r < nrhs; ++r) {
      const SlotVec xr = offset(x, r, n), vr = offset(v, r, n);
      hipError_t e = launch_permute_out(ctx, 1, xr.soa, io);
      if (e != hipSuccess) return hip_fail(ctx, e, "permute");
      int rc = d2h(ctx, hf.data(), io, sizeof(double) * len);
      if (rc) return rc;
      if (op->apply(op->user, hf.data(), hu.data()) != 0)
        return fail(ctx, GGEM_ERR_INVALID_ARG, "operator callback failed");
      if ((rc = h2d(ctx, io, hu.data(), sizeof(double) * len))) return rc;
      double* mx = io + len;
      e = launch_permute_in(ctx, 1, io, mx);
      // v may alias x_{k-1}: it is written only after x_{k-1} is consumed
      if (e == hipSuccess && epi.kind == 1) {
        e = launch_axpby(ctx, len, epi.db, xr.soa, epi.da, mx);  // db x + da M x
      } else if (e == hipSuccess && epi.kind == 2) {
        e = launch_axpby(ctx, len, 2. * epi.db, xr.soa, 2. * epi.da, mx);  // 2db x + 2da M x
        if (e == hipSuccess) e = launch_axpby(ctx, len, -1.0, epi.xm1.soa + r * len, 1.0, mx);
        if (e == hipSuccess && (epi.ymask & 1))
          e = launch_axpby(ctx, len, epi.cm1, epi.xm1.soa + r * len, 1.0, epi.y + r * len);
        if (e == hipSuccess && (epi.ymask & 2))
          e = launch_axpby(ctx, len, epi.c0, xr.soa, 1.0, epi.y + r * len);
        if (e == hipSuccess && (epi.ymask & 4)) e = launch_axpby(ctx, len, epi.c, mx, 1.0, epi.y + r * len);
      }
      if (e == hipSuccess)
        e = hipMemcpyAsync(vr.soa, mx, sizeof(double) * len, hipMemcpyDeviceToDevice, ctx->stream);
      if (e != hipSuccess) return hip_fail(ctx, e, "host operator epilogue");
      if ((rc = refresh(ctx, 1, vr))) return rc;
    }
    return GGEM_OK;
  }
};

// Operands of the fixed-order recurrence.  par = [da, db, c_0 .. c_N]; with
// dpar set the kernels read the same values from device memory instead.
struct FixedStep {
  int N, nrhs;
  const double* par;
  const double* dpar;
  const double* dw_dev;  // particle order, interleaved
  SlotVec x0, x1, yv, dwm;
  bool disp;
  const int32_t* sys = nullptr;  // batched: system of each slot, dpar holds N + 3 values per system
};

// Enqueue the whole fixed-order step: x_0 = dw, x_1 = db x_0 + da M x_0,
// y = c_0 x_0 + c_1 x_1, x_{k+1} = 2da M x_k + 2db x_k - x_{k-1} with
// y += c_{k+1} x_{k+1} (brownian_system.C:1047-1100 without the estimators),
// then dwm = M y when disp is asked for.  Nothing here synchronises, so the
// sequence can be captured into a graph.
int enqueue_fixed(ggem_ctx* ctx, Applier& apply, const FixedStep& s) {
  const SlotVec none{nullptr, nullptr, nullptr};
  hipError_t e = launch_permute_in(ctx, s.nrhs, s.dw_dev, s.x0.soa);
  if (e != hipSuccess) return hip_fail(ctx, e, "permute");
  int rc = refresh(ctx, s.nrhs, s.x0);
  if (rc) return rc;
  ChebEpi init{1, s.par[0], s.par[1], 0.0, none, nullptr};
  init.dpar = s.dpar;
  init.pstride = s.N + 3;
  init.sys = s.sys;
  if ((rc = apply(s.nrhs, s.x0, s.x1, init))) return rc;
  e = launch_cheb_init_y(ctx, s.nrhs, s.x0.soa, s.x1.soa, s.yv.soa, s.par[2], s.par[3], s.dpar, s.N + 3,
                         s.sys);
  if (e != hipSuccess) return hip_fail(ctx, e, "init y");
  SlotVec a = s.x0, b = s.x1;
  int added = 1;  // y holds c_0 x_0 + c_1 x_1
  for (int i = 1; i < s.N; ++i) {
    // x_{k+1} into x_{k-1}'s storage; y gains the terms up to x_{k+1} every
    // third step and at the last one (x_{k-1}, x_k and x_{k+1} are all at
    // hand then), so two steps in three neither read nor write y
    ChebEpi st{2, s.par[0], s.par[1], s.par[2 + i + 1], a, s.yv.soa};
    st.dpar = s.dpar;
    st.pstride = s.N + 3;
    st.term = i + 1;
    st.sys = s.sys;
    st.ymask = 0;
    if (i % 3 == 0 || i == s.N - 1) {
      for (int j = added + 1; j <= i + 1; ++j) st.ymask |= 1 << (j - (i - 1));
      st.cm1 = s.par[2 + i - 1];
      st.c0 = s.par[2 + i];
      added = i + 1;
    }
    if ((rc = apply(s.nrhs, b, a, st))) return rc;
    std::swap(a, b);
  }
  if (s.disp) {
    if ((rc = refresh(ctx, s.nrhs, s.yv))) return rc;
    const ChebEpi plain{0, 0.0, 0.0, 0.0, none, nullptr};
    if ((rc = apply(s.nrhs, s.yv, s.dwm, plain))) return rc;
  }
  return GGEM_OK;
}

// Write the recurrence parameters of every system into ctx->dpar, in
// stream order ahead of the kernels that read them.
int upload_params(ggem_ctx* ctx, const double* vals, size_t cnt) {
  int rc = ensure(ctx, &ctx->dpar, &ctx->dpar_cap, cnt);
  if (rc) return rc;
  if (cnt <= 256) {  // small sets ride in kernel arguments
    hipError_t e = launch_set_params(ctx, ctx->dpar, 0, vals, int(cnt));
    return e == hipSuccess ? GGEM_OK : hip_fail(ctx, e, "recurrence parameters");
  }
  // pinned staging, reused once the previous copy out of it has completed
  if (!ctx->par_ev) GGEM_HIP(ctx, hipEventCreateWithFlags(&ctx->par_ev, hipEventDisableTiming));
  else GGEM_HIP(ctx, hipEventSynchronize(ctx->par_ev));
  if (ctx->par_cap < cnt) {
    if (ctx->par_h) GGEM_HIP(ctx, hipHostFree(ctx->par_h));
    ctx->par_h = nullptr;
    ctx->par_cap = 0;
    GGEM_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->par_h), sizeof(double) * cnt));
    ctx->par_cap = cnt;
  }
  std::copy(vals, vals + cnt, ctx->par_h);
  GGEM_HIP(ctx, hipMemcpyAsync(ctx->dpar, ctx->par_h, sizeof(double) * cnt, hipMemcpyHostToDevice, ctx->stream));
  GGEM_HIP(ctx, hipEventRecord(ctx->par_ev, ctx->stream));
  return GGEM_OK;
}

// The fixed-order step through a cached graph exec (device operator only);
// the parameters are already in ctx->dpar.
int run_fixed_graph(ggem_ctx* ctx, Applier& apply, FixedStep s) {
  int rc = GGEM_OK;
  s.dpar = ctx->dpar;
  const Pools& P = pools(ctx);
  const std::vector<uintptr_t> key = {
      uintptr_t(s.N), uintptr_t(s.nrhs), uintptr_t(ctx->p.precision), uintptr_t(s.disp), uintptr_t(ctx->n),
      uintptr_t(s.dw_dev), uintptr_t(s.dpar), uintptr_t(P.vec), uintptr_t(P.g32), uintptr_t(P.g64),
      uintptr_t(ctx->x4), uintptr_t(ctx->nbr_cnt), uintptr_t(ctx->slice_off), uintptr_t(ctx->ell),
      uintptr_t(ctx->perm), uintptr_t(ctx->fs), uintptr_t(ctx->tab32), uintptr_t(ctx->coef64),
      uintptr_t(ctx->box_valid), uintptr_t(ctx->box_max), uintptr_t(ctx->box_cnt), uintptr_t(ctx->box_slot),
      uintptr_t(ctx->box_pos), uintptr_t(ctx->box_idx), uintptr_t(ctx->cell_start),
      uintptr_t(ctx->box_tgt), uintptr_t(ctx->box_base),
      uintptr_t(s.sys), uintptr_t(ctx->nsys)};
  if (!ctx->cheb_exec || key != ctx->cheb_key) {
    if (ctx->cheb_exec) {
      GGEM_HIP(ctx, hipGraphExecDestroy(ctx->cheb_exec));
      ctx->cheb_exec = nullptr;
    }
    // launch shapes cannot be timed inside a capture: one eager pass first
    // lets the autotuner pick them (its results are overwritten by the replay)
    if (ctx->tune.on && (rc = enqueue_fixed(ctx, apply, s))) return rc;
    // event brackets cannot be recorded into a capture
    const bool prof = ctx->prof.on;
    ctx->prof.on = false;
    GGEM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_fixed(ctx, apply, s);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(ctx->stream, &g);
    ctx->prof.on = prof;
    if (rc || ee != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : hip_fail(ctx, ee, "hipStreamEndCapture");
    }
    const hipError_t ei = hipGraphInstantiate(&ctx->cheb_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      ctx->cheb_exec = nullptr;
      return hip_fail(ctx, ei, "hipGraphInstantiate");
    }
    ctx->cheb_key = key;
  }
  ProfScope ps(ctx, "cheb_graph");
  GGEM_HIP(ctx, hipGraphLaunch(ctx->cheb_exec, ctx->stream));
  return GGEM_OK;
}

int dot_host(ggem_ctx* ctx, int nrhs, const double* a, const double* b, double* out) {
  Pools& P = pools(ctx);
  int rc = ensure(ctx, &P.red, &P.red_cap, size_t(nrhs) * 513);
  if (rc) return rc;
  hipError_t e = launch_dot(ctx, nrhs, a, b, P.red + size_t(nrhs) * 512, P.red);
  if (e != hipSuccess) return hip_fail(ctx, e, "dot");
  return d2h(ctx, out, P.red + size_t(nrhs) * 512, sizeof(double) * nrhs);
}

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

// k-th smallest eigenvalue (0-based) by bisection on the Gershgorin interval
double tridiag_eig(const std::vector<double>& a, const std::vector<double>& b, int k) {
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

// one dot product into a device scalar (no readback)
int dot_dev(ggem_ctx* ctx, const double* a, const double* b, double* out_d) {
  Pools& P = pools(ctx);
  int rc = ensure(ctx, &P.red, &P.red_cap, size_t(513));
  if (rc) return rc;
  hipError_t e = launch_dot(ctx, 1, a, b, out_d, P.red);
  return e == hipSuccess ? GGEM_OK : hip_fail(ctx, e, "dot");
}

}  // namespace

extern "C" int ggem_cheb_brownian(ggem_ctx* ctx, const ggem_op* op, int32_t N, double eig_min,
                                  double eig_max, double tol, int32_t mode, int32_t nrhs,
                                  const double* dw, double* y, double* disp, ggem_cheb_info* info) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_cheb_brownian (use ggem_batch_cheb_brownian)"))) return rc;
  const bool keep = (mode & GGEM_CHEB_KEEP) != 0;
  mode &= ~GGEM_CHEB_KEEP;
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

  if ((rc = reserve(ctx, 4, nrhs))) return rc;
  if (!op && (rc = ensure_tensors(ctx, nrhs))) return rc;
  Pools& P = pools(ctx);
  const int prec = ctx->p.precision;
  SlotVec x0 = slot_vec(P, 0, nrhs, n, prec);
  SlotVec x1 = slot_vec(P, 1, nrhs, n, prec);
  SlotVec yv = slot_vec(P, 2, nrhs, n, prec);
  SlotVec dwm = slot_vec(P, 3, nrhs, n, prec);
  const size_t blk = size_t(nrhs) * len;
  const double* dw_dev = ctx->noise;
  if (dw) {
    if ((rc = h2d(ctx, P.io, dw, sizeof(double) * blk))) return rc;
    dw_dev = P.io;
  }
  Applier apply{ctx, op, prec, {}, {}};
  const SlotVec none{nullptr, nullptr, nullptr};
  const ChebEpi plain{0, 0.0, 0.0, 0.0, none, nullptr};
  std::vector<ggem_cheb_info> inf(nrhs);
  for (auto& i : inf) i = ggem_cheb_info{0, 1, 0, 0, 0.0, 0.0};
  hipError_t e = hipSuccess;
  bool disp_done = false;

  if (mode == GGEM_CHEB_FIXED) {
    std::vector<double> par(N + 3);
    par[0] = da;
    par[1] = db;
    std::copy(X.begin(), X.end(), par.begin() + 2);
    const FixedStep st{N, nrhs, par.data(), nullptr, dw_dev, x0, x1, yv, dwm, disp != nullptr || keep};
    if (!op && ctx->graph_on) {
      if ((rc = upload_params(ctx, par.data(), par.size()))) return rc;
      rc = run_fixed_graph(ctx, apply, st);
    } else {
      rc = enqueue_fixed(ctx, apply, st);
    }
    if (rc) return rc;
    disp_done = true;
    for (auto& i : inf) {
      i.converged = 1;
      i.terms = N;
      i.applies = N + (disp || keep ? 1 : 0);
    }
  } else {
    e = launch_permute_in(ctx, nrhs, dw_dev, x0.soa);
    if (e != hipSuccess) return hip_fail(ctx, e, "permute");
    if ((rc = refresh(ctx, nrhs, x0))) return rc;
    // reference-compatible (brownian_system.C:1030-1155), one right-hand side
    // at a time; the device's own operator decides where to stop on the
    // device, a host operator on the host
    if (!op) {
      if ((rc = ensure(ctx, &P.cref, &P.cref_cap, size_t(8 * nrhs)))) return rc;
      if ((rc = ensure(ctx, &P.red, &P.red_cap, size_t(513 + 3 * 512)))) return rc;
      if ((rc = ensure(ctx, &P.cstop, &P.cstop_cap, size_t(nrhs)))) return rc;
    }
    for (int r = 0; r < nrhs && !op; ++r) {
      SlotVec a0 = offset(x0, r, n), a1 = offset(x1, r, n);
      const SlotVec ay = offset(yv, r, n), am = offset(dwm, r, n);
      double* st = P.cref + 8 * r;
      int32_t* stop = P.cstop + r;
      double* px = P.red + 513;  // partials: x . x, then (M y).(M y) and y . y
      double* pmy = px + 512;
      if ((e = launch_cheb_ref_init(ctx, st, stop)) != hipSuccess) return hip_fail(ctx, e, "chebyshev");
      if ((rc = apply(1, a0, a1, plain))) return rc;  // x1 = M x0
      if ((rc = dot_dev(ctx, a1.soa, a0.soa, st + 0))) return rc;
      if ((e = launch_axpby(ctx, len, db, a0.soa, da, a1.soa)) != hipSuccess) return hip_fail(ctx, e, "axpby");
      if ((rc = refresh(ctx, 1, a1))) return rc;
      if ((e = launch_fill(ctx, len, 0.0, ay.soa)) != hipSuccess) return hip_fail(ctx, e, "fill");
      if ((e = launch_axpby(ctx, len, X[0], a0.soa, 1.0, ay.soa)) != hipSuccess) return hip_fail(ctx, e, "axpy");
      if ((rc = dot_dev(ctx, ay.soa, ay.soa, st + 1))) return rc;
      if ((e = launch_axpby(ctx, len, X[1], a1.soa, 1.0, ay.soa)) != hipSuccess) return hip_fail(ctx, e, "axpy");
      if ((e = launch_dot_partial(ctx, a1.soa, a1.soa, px)) != hipSuccess) return hip_fail(ctx, e, "dot");
      if ((rc = refresh(ctx, 1, ay))) return rc;
      if ((rc = apply(1, ay, am, plain))) return rc;
      if ((e = launch_dot2_partial(ctx, am.soa, am.soa, ay.soa, ay.soa, pmy)) != hipSuccess ||
          (e = launch_cheb_ref_term(ctx, st, stop, px, pmy, std::fabs(X[1]), tol, 1, 0)) != hipSuccess)
        return hip_fail(ctx, e, "chebyshev");
      ChebEpi skip_plain = plain;
      skip_plain.skip = stop;
      // terms are launched up to two past the last call's stop; one small
      // readback then tells whether the rest are needed (terms launched after
      // the stop return at entry, so the results do not depend on the guess)
      const int guess = std::min(N, std::max(2, ctx->cheb_terms + 2));
      for (int i = 1; i < N; ++i) {
        if (i == guess) {
          int64_t stopped = 0;
          if ((rc = peek(ctx, {{stop, 4}}, &stopped))) return rc;
          if (stopped) break;
        }
        // y . y before the term is in st[1] (from the previous term)
        ChebEpi stp{2, da, db, X[i + 1], a0, ay.soa};
        stp.skip = stop;
        if ((rc = apply(1, a1, a0, stp))) return rc;  // a0 <- x_{k+1}; y += c x_{k+1}
        if ((e = launch_dot_partial(ctx, a0.soa, a0.soa, px)) != hipSuccess) return hip_fail(ctx, e, "dot");
        if ((rc = refresh(ctx, 1, ay))) return rc;
        if ((rc = apply(1, ay, am, skip_plain))) return rc;
        if ((e = launch_dot2_partial(ctx, am.soa, am.soa, ay.soa, ay.soa, pmy)) != hipSuccess ||
            (e = launch_cheb_ref_term(ctx, st, stop, px, pmy, std::fabs(X[i + 1]), tol, i + 1, 1)) != hipSuccess)
          return hip_fail(ctx, e, "chebyshev");
        std::swap(a0, a1);
      }
    }
    if (!op && info) {  // one readback for every right-hand side's statistics
      std::vector<double> h(size_t(8 * nrhs));
      if ((rc = d2h(ctx, h.data(), P.cref, sizeof(double) * 8 * size_t(nrhs)))) return rc;
      for (int r = 0; r < nrhs; ++r) {
        const double* q = h.data() + 8 * r;
        const int terms = int(q[6]);
        ctx->cheb_terms = std::max(r ? ctx->cheb_terms : 0, terms);
        inf[r] = ggem_cheb_info{q[7] != 0.0 ? 1 : 0, terms, 2 * terms + (disp || keep ? 1 : 0), 0, q[4], q[5]};
      }
    }
    for (int r = 0; r < nrhs && op; ++r) {
      SlotVec a0 = offset(x0, r, n), a1 = offset(x1, r, n);
      const SlotVec ay = offset(yv, r, n), am = offset(dwm, r, n);
      int applies = 0;
      if ((rc = apply(1, a0, a1, plain))) return rc;  // x1 = M x0
      ++applies;
      double aux1, aux2, aux3, aux4;
      if ((rc = dot_host(ctx, 1, a1.soa, a0.soa, &aux2))) return rc;
      if ((e = launch_axpby(ctx, len, db, a0.soa, da, a1.soa)) != hipSuccess) return hip_fail(ctx, e, "axpby");
      if ((rc = refresh(ctx, 1, a1))) return rc;
      if ((e = launch_fill(ctx, len, 0.0, ay.soa)) != hipSuccess) return hip_fail(ctx, e, "fill");
      if ((e = launch_axpby(ctx, len, X[0], a0.soa, 1.0, ay.soa)) != hipSuccess) return hip_fail(ctx, e, "axpy");
      if ((rc = dot_host(ctx, 1, ay.soa, ay.soa, &aux4))) return rc;
      aux4 = std::sqrt(aux4);
      if ((e = launch_axpby(ctx, len, X[1], a1.soa, 1.0, ay.soa)) != hipSuccess) return hip_fail(ctx, e, "axpy");
      if ((rc = dot_host(ctx, 1, a1.soa, a1.soa, &aux3))) return rc;
      aux3 = std::sqrt(aux3) * std::fabs(X[1]);
      if ((rc = refresh(ctx, 1, ay))) return rc;
      if ((rc = apply(1, ay, am, plain))) return rc;
      ++applies;
      if ((rc = dot_host(ctx, 1, am.soa, am.soa, &aux1))) return rc;
      double err1 = 0, err2 = 0;
      if (aux2 != 0.0) err1 = std::sqrt(std::fabs(aux1 - aux2) / aux2);
      if (aux4 != 0.0) err2 = aux3 / aux4;
      bool conv = false;
      int terms = 1;
      for (int i = 1; i < N; ++i) {
        if ((rc = dot_host(ctx, 1, ay.soa, ay.soa, &aux4))) return rc;
        aux4 = std::sqrt(aux4);
        const ChebEpi st{2, da, db, X[i + 1], a0, ay.soa};
        if ((rc = apply(1, a1, a0, st))) return rc;  // a0 <- x_{k+1}; y += c x_{k+1}
        ++applies;
        terms = i + 1;
        if ((rc = dot_host(ctx, 1, a0.soa, a0.soa, &aux3))) return rc;
        aux3 = std::sqrt(aux3) * std::fabs(X[i + 1]);
        if ((rc = refresh(ctx, 1, ay))) return rc;
        if ((rc = apply(1, ay, am, plain))) return rc;
        ++applies;
        if ((rc = dot_host(ctx, 1, am.soa, am.soa, &aux1))) return rc;
        if (aux2 != 0.0) err1 = std::sqrt(std::fabs(aux1 - aux2) / aux2);
        if (aux4 != 0.0) err2 = aux3 / aux4;
        if (err1 <= tol) {
          conv = true;
          break;
        }
        if (err2 <= 1E-9) break;
        std::swap(a0, a1);
      }
      inf[r] = ggem_cheb_info{conv ? 1 : 0, terms, applies + (disp || keep ? 1 : 0), 0, err1, err2};
    }
  }
  if (y) {
    if ((e = launch_permute_out(ctx, nrhs, yv.soa, P.io)) != hipSuccess) return hip_fail(ctx, e, "permute");
    if ((rc = d2h(ctx, y, P.io, sizeof(double) * blk))) return rc;
  }
  if (disp || keep) {  // hi_ewald: disp = M y (copss.C:1686)
    if (!disp_done) {
      if ((rc = refresh(ctx, nrhs, yv))) return rc;
      if ((rc = apply(nrhs, yv, dwm, plain))) return rc;
    }
  }
  if (keep) {  // y and M y stay on the device, particle order
    if ((rc = ensure_keep(ctx, nrhs))) return rc;
    if ((e = launch_permute_out(ctx, nrhs, yv.soa, ctx->keep_y)) != hipSuccess ||
        (e = launch_permute_out(ctx, nrhs, dwm.soa, ctx->keep_d)) != hipSuccess)
      return hip_fail(ctx, e, "permute");
    ctx->keep_nrhs = nrhs;
    ctx->keep_d_valid = true;
  }
  if (disp) {
    if ((e = launch_permute_out(ctx, nrhs, dwm.soa, P.io)) != hipSuccess) return hip_fail(ctx, e, "permute");
    if ((rc = d2h(ctx, disp, P.io, sizeof(double) * blk))) return rc;
  }
  if (info)
    for (int r = 0; r < nrhs; ++r) info[r] = inf[r];
  return y || disp ? sync(ctx) : GGEM_OK;
}

extern "C" int ggem_get_brownian(ggem_ctx* ctx, double* y, double* disp) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (ctx->keep_nrhs < 1) return fail(ctx, GGEM_ERR_NOT_READY, "no kept Brownian step (GGEM_CHEB_KEEP)");
  const size_t bytes = sizeof(double) * size_t(ctx->keep_nrhs) * 3 * size_t(ctx->n);
  int rc = GGEM_OK;
  if (y && bytes && (rc = d2h(ctx, y, ctx->keep_y, bytes))) return rc;
  if (disp && bytes) {
    if (!ctx->keep_d_valid) return fail(ctx, GGEM_ERR_NOT_READY, "the kept displacement is not valid");
    rc = d2h(ctx, disp, ctx->keep_d, bytes);
  }
  return rc;
}

// hi_ewald(M, dw, dw_mid) at the mid-point (copss.C:1730): M at the current
// positions and lists applied to the kept y.
extern "C" int ggem_displacement_apply(ggem_ctx* ctx, double* disp) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_displacement_apply"))) return rc;
  const int nrhs = ctx->keep_nrhs;
  if (nrhs < 1) return fail(ctx, GGEM_ERR_NOT_READY, "no kept Brownian step (GGEM_CHEB_KEEP)");
  const int64_t n = ctx->n;
  if (n == 0) return GGEM_OK;
  if ((rc = reserve(ctx, 2, nrhs))) return rc;
  if ((rc = ensure_tensors(ctx, nrhs))) return rc;
  Pools& P = pools(ctx);
  const int prec = ctx->p.precision;
  const SlotVec yv = slot_vec(P, 0, nrhs, n, prec), dv = slot_vec(P, 1, nrhs, n, prec);
  hipError_t e = launch_permute_in(ctx, nrhs, ctx->keep_y, yv.soa);
  if (e != hipSuccess) return hip_fail(ctx, e, "permute");
  if ((rc = refresh(ctx, nrhs, yv))) return rc;
  Applier apply{ctx, nullptr, prec, {}, {}};
  const SlotVec none{nullptr, nullptr, nullptr};
  if ((rc = apply(nrhs, yv, dv, ChebEpi{0, 0.0, 0.0, 0.0, none, nullptr}))) return rc;
  if ((e = launch_permute_out(ctx, nrhs, dv.soa, ctx->keep_d)) != hipSuccess) return hip_fail(ctx, e, "permute");
  ctx->keep_d_valid = true;
  if (disp) return d2h(ctx, disp, ctx->keep_d, sizeof(double) * size_t(nrhs) * 3 * size_t(n));
  return GGEM_OK;
}

extern "C" int ggem_lanczos_bounds(ggem_ctx* ctx, const ggem_op* op, int32_t iters, uint64_t seed,
                                   double* lmin, double* lmax) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_lanczos_bounds (use ggem_batch_lanczos_bounds)"))) return rc;
  if (iters < 1 || !lmin || !lmax) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid Lanczos arguments");
  if (op && !op->apply) return fail(ctx, GGEM_ERR_INVALID_ARG, "operator without apply");
  const int64_t n = ctx->n, len = 3 * n;
  if (len == 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "no particles");
  if ((rc = reserve(ctx, 3, 1))) return rc;
  if (!op && (rc = ensure_tensors(ctx))) return rc;
  Pools& P = pools(ctx);
  const int prec = ctx->p.precision;
  SlotVec vprev = slot_vec(P, 0, 1, n, prec);
  SlotVec vcur = slot_vec(P, 1, 1, n, prec);
  const SlotVec w = slot_vec(P, 2, 1, n, prec);
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
  if ((rc = h2d(ctx, vcur.soa, h.data(), sizeof(double) * len))) return rc;
  if ((rc = refresh(ctx, 1, vcur))) return rc;
  hipError_t e = launch_fill(ctx, len, 0.0, vprev.soa);
  if (e != hipSuccess) return hip_fail(ctx, e, "fill");
  Applier apply{ctx, op, prec, {}, {}};
  const SlotVec none{nullptr, nullptr, nullptr};
  const ChebEpi plain{0, 0.0, 0.0, 0.0, none, nullptr};
  std::vector<double> alpha, beta;
  double bprev = 0.0;
  const int kmax = int(std::min<int64_t>(iters, len));
  for (int j = 0; j < kmax; ++j) {
    if ((rc = apply(1, vcur, w, plain))) return rc;
    double a;
    if ((rc = dot_host(ctx, 1, w.soa, vcur.soa, &a))) return rc;
    alpha.push_back(a);
    if ((e = launch_axpby(ctx, len, -a, vcur.soa, 1.0, w.soa)) != hipSuccess) return hip_fail(ctx, e, "axpby");
    if ((e = launch_axpby(ctx, len, -bprev, vprev.soa, 1.0, w.soa)) != hipSuccess) return hip_fail(ctx, e, "axpby");
    double b2;
    if ((rc = dot_host(ctx, 1, w.soa, w.soa, &b2))) return rc;
    const double b = std::sqrt(b2);
    if (j + 1 == kmax || b <= 1e-12 * std::fabs(a)) break;
    beta.push_back(b);
    std::swap(vprev, vcur);  // vprev <- v_j; vcur <- w / b
    if ((e = launch_fill(ctx, len, 0.0, vcur.soa)) != hipSuccess) return hip_fail(ctx, e, "fill");
    if ((e = launch_axpby(ctx, len, 1.0 / b, w.soa, 0.0, vcur.soa)) != hipSuccess) return hip_fail(ctx, e, "axpby");
    if ((rc = refresh(ctx, 1, vcur))) return rc;
    bprev = b;
  }
  beta.resize(alpha.size() > 0 ? alpha.size() - 1 : 0);
  *lmin = tridiag_eig(alpha, beta, 0);
  *lmax = tridiag_eig(alpha, beta, int(alpha.size()) - 1);
  return GGEM_OK;
}

extern "C" int ggem_batch_cheb_brownian(ggem_ctx* ctx, int32_t N, const double* eig_min, const double* eig_max,
                                        int32_t nrhs, const double* dw, double* y, double* disp) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  const int nsys = ctx->nsys;
  if (N < 1 || nrhs < 1 || !eig_min || !eig_max) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid Chebyshev arguments");
  for (int k = 0; k < nsys; ++k)
    if (!(eig_max[k] > eig_min[k]))
      return fail(ctx, GGEM_ERR_INVALID_ARG, "system " + std::to_string(k) + ": eig_max must exceed eig_min");
  if (!dw && (ctx->noise_nrhs != nrhs || !ctx->noise))
    return fail(ctx, GGEM_ERR_NOT_READY, "no device noise of this nrhs: call ggem_set_noise");
  const int64_t n = ctx->n, len = 3 * n;
  if (n == 0) return GGEM_OK;
  // per system [da, db, c_0 .. c_N] (brownian_system.C:1001-1003, chebyshev.C:57-151)
  const size_t stride = size_t(N) + 3;
  std::vector<double> par(stride * nsys), X;
  for (int k = 0; k < nsys; ++k) {
    const double d_eig = eig_max[k] - eig_min[k];
    const double da = 2.0 / d_eig, db = -(eig_max[k] + eig_min[k]) / d_eig;
    cheb_coefficients(N, da, db, &X);
    double* pk = par.data() + stride * k;
    pk[0] = da;
    pk[1] = db;
    std::copy(X.begin(), X.end(), pk + 2);
  }
  if ((rc = reserve(ctx, 4, nrhs))) return rc;
  if ((rc = ensure_tensors(ctx, nrhs))) return rc;
  Pools& P = pools(ctx);
  const int prec = ctx->p.precision;
  const SlotVec x0 = slot_vec(P, 0, nrhs, n, prec), x1 = slot_vec(P, 1, nrhs, n, prec);
  const SlotVec yv = slot_vec(P, 2, nrhs, n, prec), dwm = slot_vec(P, 3, nrhs, n, prec);
  const size_t blk = size_t(nrhs) * len;
  const double* dw_dev = ctx->noise;
  if (dw) {
    if ((rc = h2d(ctx, P.io, dw, sizeof(double) * blk))) return rc;
    dw_dev = P.io;
  }
  if ((rc = upload_params(ctx, par.data(), par.size()))) return rc;
  Applier apply{ctx, nullptr, prec, {}, {}};
  const FixedStep st{N, nrhs, par.data(), ctx->dpar, dw_dev, x0, x1, yv, dwm, disp != nullptr,
                     nsys > 1 ? ctx->sys_slot : nullptr};
  rc = ctx->graph_on ? run_fixed_graph(ctx, apply, st) : enqueue_fixed(ctx, apply, st);
  if (rc) return rc;
  hipError_t e;
  if (y) {
    if ((e = launch_permute_out(ctx, nrhs, yv.soa, P.io)) != hipSuccess) return hip_fail(ctx, e, "permute");
    if ((rc = d2h(ctx, y, P.io, sizeof(double) * blk))) return rc;
  }
  if (disp) {
    if ((e = launch_permute_out(ctx, nrhs, dwm.soa, P.io)) != hipSuccess) return hip_fail(ctx, e, "permute");
    if ((rc = d2h(ctx, disp, P.io, sizeof(double) * blk))) return rc;
  }
  return GGEM_OK;
}

// Lanczos on every system at once: one operator application per step for
// the whole batch, per-system dot products, and per-system breakdown.
extern "C" int ggem_batch_lanczos_bounds(ggem_ctx* ctx, int32_t iters, uint64_t seed, double* lmin,
                                         double* lmax) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if (iters < 1 || !lmin || !lmax) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid Lanczos arguments");
  const int nsys = ctx->nsys;
  const int64_t n = ctx->n, len = 3 * n;
  if (len == 0) return fail(ctx, GGEM_ERR_INVALID_ARG, "no particles");
  if (nsys == 1) return ggem_lanczos_bounds(ctx, nullptr, iters, seed, lmin, lmax);
  if ((rc = reserve(ctx, 3, 1))) return rc;
  if ((rc = ensure_tensors(ctx))) return rc;
  Pools& P = pools(ctx);
  const int prec = ctx->p.precision;
  SlotVec vprev = slot_vec(P, 0, 1, n, prec);
  SlotVec vcur = slot_vec(P, 1, 1, n, prec);
  const SlotVec w = slot_vec(P, 2, 1, n, prec);
  // start vectors: system k draws 3 n_k normals from mt19937_64(seed + k) in
  // the [3][n_k] slot layout its own context would use; the slots of system
  // k are its particle range
  std::vector<double> h(len, 0.0);
  std::vector<int> kmax(nsys);
  for (int k = 0; k < nsys; ++k) {
    const int64_t s0 = ctx->part_off[k], nk = ctx->part_off[k + 1] - s0;
    kmax[k] = int(std::min<int64_t>(iters, 3 * nk));
    std::mt19937_64 gen(seed + uint64_t(k));
    std::normal_distribution<double> nd(0.0, 1.0);
    double ss = 0;
    for (int64_t j = 0; j < 3 * nk; ++j) {
      const double v = nd(gen);
      h[(j / nk) * n + s0 + (j % nk)] = v;
      ss += v * v;
    }
    const double sc = ss > 0 ? 1.0 / std::sqrt(ss) : 0.0;
    for (int c = 0; c < 3; ++c)
      for (int64_t i = 0; i < nk; ++i) h[c * n + s0 + i] *= sc;
  }
  if ((rc = h2d(ctx, vcur.soa, h.data(), sizeof(double) * len))) return rc;
  if ((rc = refresh(ctx, 1, vcur))) return rc;
  hipError_t e = launch_fill(ctx, len, 0.0, vprev.soa);
  if (e != hipSuccess) return hip_fail(ctx, e, "fill");
  if ((rc = ensure(ctx, &P.red, &P.red_cap, size_t(4) * nsys))) return rc;
  double* d_a = P.red;             // a_k, then the scale 1/b_k
  double* d_b = P.red + nsys;      // b_{k-1}
  double* d_dot = P.red + 2 * nsys;
  Applier apply{ctx, nullptr, prec, {}, {}};
  const SlotVec none{nullptr, nullptr, nullptr};
  const ChebEpi plain{0, 0.0, 0.0, 0.0, none, nullptr};
  std::vector<std::vector<double>> alpha(nsys), beta(nsys);
  std::vector<double> a(nsys), b2(nsys), bprev(nsys, 0.0), sc(nsys);
  std::vector<char> active(nsys);
  int kall = 0;
  for (int k = 0; k < nsys; ++k) {
    active[k] = kmax[k] > 0;
    kall = std::max(kall, kmax[k]);
  }
  for (int j = 0; j < kall; ++j) {
    if ((rc = apply(1, vcur, w, plain))) return rc;
    if ((e = launch_seg_dot(ctx, w.soa, vcur.soa, d_dot)) != hipSuccess) return hip_fail(ctx, e, "dot");
    if ((rc = d2h(ctx, a.data(), d_dot, sizeof(double) * nsys))) return rc;
    if ((rc = h2d(ctx, d_a, a.data(), sizeof(double) * nsys))) return rc;
    if ((rc = h2d(ctx, d_b, bprev.data(), sizeof(double) * nsys))) return rc;
    if ((e = launch_seg_axpy2(ctx, d_a, vcur.soa, d_b, vprev.soa, w.soa)) != hipSuccess)
      return hip_fail(ctx, e, "update");
    if ((e = launch_seg_dot(ctx, w.soa, w.soa, d_dot)) != hipSuccess) return hip_fail(ctx, e, "dot");
    if ((rc = d2h(ctx, b2.data(), d_dot, sizeof(double) * nsys))) return rc;
    bool any = false;
    for (int k = 0; k < nsys; ++k) {
      sc[k] = 0.0;
      if (!active[k]) {
        bprev[k] = 0.0;
        continue;
      }
      alpha[k].push_back(a[k]);
      const double bk = std::sqrt(b2[k]);
      if (j + 1 == kmax[k] || bk <= 1e-12 * std::fabs(a[k])) {
        active[k] = 0;
        bprev[k] = 0.0;
        continue;
      }
      beta[k].push_back(bk);
      sc[k] = 1.0 / bk;
      bprev[k] = bk;
      any = true;
    }
    if (!any) break;
    std::swap(vprev, vcur);  // vprev <- v_j; vcur <- w / b
    if ((rc = h2d(ctx, d_a, sc.data(), sizeof(double) * nsys))) return rc;
    if ((e = launch_seg_scale(ctx, d_a, w.soa, vcur.soa)) != hipSuccess) return hip_fail(ctx, e, "scale");
    if ((rc = refresh(ctx, 1, vcur))) return rc;
  }
  for (int k = 0; k < nsys; ++k) {
    if (alpha[k].empty()) {
      lmin[k] = lmax[k] = 0.0;
      continue;
    }
    beta[k].resize(alpha[k].size() - 1);
    lmin[k] = tridiag_eig(alpha[k], beta[k], 0);
    lmax[k] = tridiag_eig(alpha[k], beta[k], int(alpha[k].size()) - 1);
  }
  return GGEM_OK;
}
