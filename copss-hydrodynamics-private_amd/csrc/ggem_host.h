// ggem_host.h -- host-side helpers shared by the C-ABI translation units:
// device buffer growth, scratch pools, synchronous copies.
#pragma once

#include <algorithm>
#include <string>

#include "ggem_internal.h"

namespace ggem {

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
  double* vec = nullptr;   // Chebyshev / Lanczos vectors (slot SoA)
  size_t vec_cap = 0;
  float4* g32 = nullptr;   // their FP32 gather records
  size_t g32_cap = 0;
  double4* g64 = nullptr;  // their FP64 gather records
  size_t g64_cap = 0;
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
  double* cref = nullptr;   // reference-mode Chebyshev state per right-hand side (8 doubles)
  size_t cref_cap = 0;
  int32_t* cstop = nullptr;  // ... and its stop flag
  size_t cstop_cap = 0;
};

inline Pools& pools(ggem_ctx* ctx) { return *reinterpret_cast<Pools*>(ctx->scratch); }

// Pair tensors of the current lists in the context's precision, evaluated
// on first use after a neighbour build, for applies of up to nrhs operands.
int ensure_tensors(ggem_ctx* ctx, int nrhs = 1);
// the sliced neighbour lists of the current build (filled on first use)
int ensure_ell(ggem_ctx* ctx);
// the element particle lists of the spreading (filled on first use)
int ensure_elist(ggem_ctx* ctx);
// what launch_spread reads at nqp_1d Gauss points per axis: the element
// lists unless the brick kernel serves it (and the autotuner does not time
// the other shapes)
inline int ensure_spread(ggem_ctx* ctx, int nqp_1d) {
  const int mode = spread_brick_mode(ctx, ctx->p.precision, nqp_1d);
  if (mode > 0) {
    const size_t nb = size_t(ctx->grid.n_cells >> (3 * kBrickBits));
    const int rc = ensure(ctx, &ctx->spread_unstaged, &ctx->spread_unstaged_cap, nb);
    if (rc || (mode == 2 && !ctx->tune.on)) return rc;
  }
  return ensure_elist(ctx);
}

inline int check_ready(ggem_ctx* ctx) {
  if (!ctx) return fail(nullptr, GGEM_ERR_INVALID_ARG, "null context");
  if (!ctx->built) return fail(ctx, GGEM_ERR_NOT_READY, "call ggem_build_neighbors after setting particles");
  if (!ctx->slots_current)
    return fail(ctx, GGEM_ERR_NOT_READY, "the particles moved: call ggem_build_neighbors or ggem_update_positions");
  return GGEM_OK;
}

// Calls that address the single mesh of a plain context.
inline int single_only(ggem_ctx* ctx, const char* what) {
  if (ctx->nsys > 1)
    return fail(ctx, GGEM_ERR_INVALID_ARG, std::string(what) + " is not available for a batched context");
  return GGEM_OK;
}

inline int sync(ggem_ctx* ctx) {
  GGEM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GGEM_OK;
}

inline int d2h(ggem_ctx* ctx, void* host, const void* dev, size_t bytes) {
  GGEM_HIP(ctx, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
  return sync(ctx);
}

inline int h2d(ggem_ctx* ctx, void* dev, const void* host, size_t bytes) {
  GGEM_HIP(ctx, hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  return GGEM_OK;
}

}  // namespace ggem
