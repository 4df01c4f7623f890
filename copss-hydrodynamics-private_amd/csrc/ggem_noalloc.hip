// ggem_noalloc.hip -- the field-point local velocity (GGEMStokes::
// local_solution_field for many points) without any device allocation in
// the call: its staging (points, element ids, results, an error flag) lives
// in one fixed device array per GPU instead of the context's growable pools.
//
// The array is file-scope device memory of kScratchBytes; its device address
// is looked up once per device, on the entry point's first call on that
// device (scratch_address), and that address is what host code passes to
// kernels and copies.  Every call on one GPU shares the array: calls that
// overlap on the same GPU (two threads, two contexts) race on it.
#include <mutex>

#include "copss_ggem_noalloc.h"
#include "ggem_host.h"

namespace ggem {

namespace {

constexpr size_t kScratchBytes = size_t(4) << 20;  // 4 MiB per GPU
constexpr size_t kFlagBytes = 256;                 // the error flag, padded to the array's alignment
constexpr int64_t kNoAllocMaxPoints = 65536;      // 256 + 52 B per point: 3.4 MB of the array
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

__global__ void k_clear_flag(int32_t* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

// flag = 1 when a located element is -1 (the point lies outside the mesh)
__global__ void k_flag_outside(int64_t m, const int32_t* __restrict__ elem, int32_t* __restrict__ flag) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q < m && elem[q] < 0) atomicOr(flag, 1);
}

// The array's address on `device`, which must be current; looked up on the
// first call per device and kept.
hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_addr[device] = p;
  }
  *out = g_addr[device];
  return hipSuccess;
}

}  // namespace

}  // namespace ggem

using namespace ggem;

extern "C" int ggem_local_velocity_points_noalloc(ggem_ctx* ctx, int64_t m, const double* pts,
                                                  const int32_t* elem, double* u) {
  int rc = check_ready(ctx);
  if (rc) return rc;
  if ((rc = single_only(ctx, "ggem_local_velocity_points_noalloc"))) return rc;
  if (m < 0 || (m > 0 && (!pts || !u))) return fail(ctx, GGEM_ERR_INVALID_ARG, "invalid point arrays");
  if (m > kNoAllocMaxPoints)
    return fail(ctx, GGEM_ERR_CAPACITY,
                "ggem_local_velocity_points_noalloc takes at most 65536 points per call (" + std::to_string(m) +
                    " given); split the points or use ggem_local_velocity_points");
  if (m == 0) return GGEM_OK;
  if (elem)
    for (int64_t q = 0; q < m; ++q)
      if (elem[q] < 0 || elem[q] >= ctx->grid.n_elem)
        return fail(ctx, GGEM_ERR_INVALID_ELEM, "invalid element id " + std::to_string(elem[q]) + " for point " +
                                                    std::to_string(q));  // ggem_stokes.C:369-380
  GGEM_HIP(ctx, hipSetDevice(ctx->device));
  void* base = nullptr;
  GGEM_HIP(ctx, scratch_address(ctx->device, &base));
  // [flag | points 3m | results 3m | elements m] inside the array
  int32_t* d_flag = static_cast<int32_t*>(base);
  double* d_pts = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + kFlagBytes);
  double* d_u = d_pts + 3 * m;
  int32_t* d_el = reinterpret_cast<int32_t*>(d_u + 3 * m);
  GGEM_HIP(ctx, hipMemcpyAsync(d_pts, pts, sizeof(double) * 3 * m, hipMemcpyHostToDevice, ctx->stream));
  if (elem) {
    GGEM_HIP(ctx, hipMemcpyAsync(d_el, elem, sizeof(int32_t) * m, hipMemcpyHostToDevice, ctx->stream));
  } else {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, ctx->stream, d_flag);
    hipError_t e = launch_locate(ctx, m, d_pts, d_el);
    if (e != hipSuccess) return hip_fail(ctx, e, "locate");
    hipLaunchKernelGGL(k_flag_outside, dim3(unsigned((m + 255) / 256)), dim3(256), 0, ctx->stream, m, d_el, d_flag);
    GGEM_HIP(ctx, hipGetLastError());
    int32_t outside = 0;
    if ((rc = d2h(ctx, &outside, d_flag, sizeof(int32_t)))) return rc;
    if (outside) return fail(ctx, GGEM_ERR_OUT_OF_DOMAIN, "a point lies outside the mesh");
  }
  const hipError_t e = launch_pair_points(ctx, ctx->p.precision, m, d_pts, d_el, d_u);
  if (e != hipSuccess) return hip_fail(ctx, e, "pair_points");
  return d2h(ctx, u, d_u, sizeof(double) * 3 * m);
}
