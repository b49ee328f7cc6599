// noalloc.hip -- idmx_max_abs_diff_f64_noalloc: the steady-state check
// max_i |a_i - b_i| for callers that cannot allocate device memory. Its
// per-block partial results live in a file-scope device array (one copy per
// GPU, held until the process ends) instead of a context's scratch buffer.
#include <cmath>
#include <mutex>

#include "internal.h"

// Partial results of one call: one per workgroup of the reduction grid.
#define IDMX_NOALLOC_PARTIALS 1024
static __device__ __align__(256) double scratch[IDMX_NOALLOC_PARTIALS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per GPU

// The address of this GPU's copy of `scratch`, looked up once per device
// with that device current. The array's name is a host-side stand-in and is
// never passed to a kernel or a copy.
int scratch_on_current_device(double** out) {
  int dev = 0;
  IDMX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    idmx::set_error("idmx_max_abs_diff_f64_noalloc: device index out of range");
    return IDMX_E_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return idmx::hip_fail(e == hipSuccess ? hipErrorInvalidSymbol : e,
                            "idmx_max_abs_diff_f64_noalloc: hipGetSymbolAddress");
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return IDMX_OK;
}

// part[block] = max |a - b| over the block's grid-stride share (+inf if any
// difference is NaN); every workgroup writes its slot, so nothing needs
// clearing between calls.
__global__ void __launch_bounds__(256)
k_max_abs_diff_partial(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                       double* __restrict__ part) {
  double v = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double d = fabs(a[i] - b[i]);
    v = d == d ? fmax(v, d) : INFINITY;
  }
  v = idmx::block_reduce<1>(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

}  // namespace

extern "C" int idmx_max_abs_diff_f64_noalloc(void* stream, const double* a, const double* b,
                                             int64_t n, double* out) {
  if (!a || !b || !out || n < 0) {
    idmx::set_error("idmx_max_abs_diff_f64_noalloc: invalid argument");
    return IDMX_E_ARG;
  }
  if (n == 0) {
    *out = -INFINITY;
    return IDMX_OK;
  }
  double* part = nullptr;
  const int rc = scratch_on_current_device(&part);
  if (rc) return rc;
  const int64_t blocks = (n + 255) / 256;
  const unsigned grid = (unsigned)(blocks < IDMX_NOALLOC_PARTIALS ? blocks : IDMX_NOALLOC_PARTIALS);
  if (grid > IDMX_NOALLOC_PARTIALS) {
    idmx::set_error("idmx_max_abs_diff_f64_noalloc: more partial results than the scratch holds");
    return IDMX_E_ARG;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_max_abs_diff_partial, dim3(grid), dim3(256), 0, st, a, b, n, part);
  IDMX_HIP(hipGetLastError());
  double h[IDMX_NOALLOC_PARTIALS];
  IDMX_HIP(hipMemcpyAsync(h, part, sizeof(double) * grid, hipMemcpyDeviceToHost, st));
  IDMX_HIP(hipStreamSynchronize(st));
  double r = h[0];
  for (unsigned k = 1; k < grid; k++) r = r > h[k] ? r : h[k];
  *out = r;
  return IDMX_OK;
}
