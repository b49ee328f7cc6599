// runtime.hip -- context, device memory, reductions and the Shan-Chen
// simulation object of libidmx.so (C ABI declared in include/idmx.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace idmx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return IDMX_E_HIP;
}

template <int OP>
__global__ void __launch_bounds__(256)
k_reduce_partial(const double* __restrict__ a, const double* __restrict__ b,
                 int64_t n, double* __restrict__ out) {
  // OP 0: sum a; 1: max a; 2: max |a - b|
  double v = OP == 0 ? 0.0 : -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    // max and max |a - b| report +inf for a NaN entry (fmax would skip it)
    if (OP == 0) v += a[i];
    else if (OP == 1) v = a[i] == a[i] ? fmax(v, a[i]) : INFINITY;
    else {
      const double d = fabs(a[i] - b[i]);
      v = d == d ? fmax(v, d) : INFINITY;
    }
  }
  v = block_reduce<OP == 0 ? 0 : 1>(v);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

template <int OP>
static int reduce(idmx_ctx_s* c, const double* a, const double* b, int64_t n,
                  double* out) {
  if (!c || !a || !out || n < 0 || (OP == 2 && !b)) {
    set_error("reduction: invalid argument");
    return IDMX_E_ARG;
  }
  if (n == 0) {
    *out = OP == 0 ? 0.0 : -INFINITY;
    return IDMX_OK;
  }
  const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL((k_reduce_partial<OP>), dim3(grid), dim3(256), 0, c->stream, a, b, n,
                     c->partials);
  IDMX_HIP(hipGetLastError());
  std::vector<double> h(grid);
  IDMX_HIP(hipMemcpyAsync(h.data(), c->partials, sizeof(double) * grid,
                          hipMemcpyDeviceToHost, c->stream));
  IDMX_HIP(hipStreamSynchronize(c->stream));
  double r = h[0];
  for (unsigned k = 1; k < grid; k++) r = OP == 0 ? r + h[k] : std::max(r, h[k]);
  *out = r;
  return IDMX_OK;
}

// K_CheckUMeta + IdpyMemory.Sum/Max (LBMKernelsMeta.py, LBM.py:1500-1527,
// MultiPhaseLoopChecks.py:96-163): per site delta = sum_d |u - old_u|,
// old_u <- u, u2 = sum_d u^2; per block partial sum(delta) and max(u2).
__global__ void __launch_bounds__(256)
k_check_u(const double* __restrict__ u, double* __restrict__ old_u, int64_t V, int D,
          double* __restrict__ out) {
  double sd = 0.0, mu = 0.0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x) {
    double dl = 0.0, u2 = 0.0;
    for (int d = 0; d < D; d++) {
      const double v = u[g + d * V];
      u2 += v * v;
      dl += fabs(v - old_u[g + d * V]);
      old_u[g + d * V] = v;
    }
    sd += dl;
    mu = fmax(mu, u2);
  }
  sd = block_reduce<0>(sd);
  mu = block_reduce<1>(mu);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = sd;
    out[2 * blockIdx.x + 1] = mu;
  }
}

// K_CenterOfMass + IdpyMemory.Sum (idpy/LBM/MultiPhase.py:108-137): per
// block partial sums of n and n * x_d, x_d the lattice coordinates of the
// lexicographic index (PosFromIndex, MultiPhase.py:100-106; no periodic
// unwrapping, as in the reference).
__global__ void __launch_bounds__(256)
k_com_partial(const double* __restrict__ n, int L0, int L1, int64_t V, double* __restrict__ out) {
  double m = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < V;
       g += (int64_t)gridDim.x * blockDim.x) {
    const double v = n[g];
    const int64_t r = g / L0;
    m += v;
    mx += v * (double)(g - r * L0);
    my += v * (double)(r % L1);
    mz += v * (double)(r / L1);
  }
  m = block_reduce<0>(m);
  mx = block_reduce<0>(mx);
  my = block_reduce<0>(my);
  mz = block_reduce<0>(mz);
  if (threadIdx.x == 0) {
    out[4 * blockIdx.x] = m;
    out[4 * blockIdx.x + 1] = mx;
    out[4 * blockIdx.x + 2] = my;
    out[4 * blockIdx.x + 3] = mz;
  }
}

// Plain streaming copy, one element per thread: the calibration kernel for
// achievable HBM bandwidth and for the FETCH_SIZE / WRITE_SIZE counters.
__global__ void __launch_bounds__(256)
k_stream_copy(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

}  // namespace idmx

using namespace idmx;

#define CHECK_ARG(cond, msg)          \
  do {                                \
    if (!(cond)) {                    \
      ::idmx::set_error(msg);         \
      return IDMX_E_ARG;              \
    }                                 \
  } while (0)

extern "C" {

int idmx_abi_version(void) { return IDMX_ABI_VERSION; }

int idmx_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return IDMX_E_ARG;
  snprintf(buf, len, "%s", g_err.c_str());
  return IDMX_OK;
}

int idmx_device_count(int* count) {
  CHECK_ARG(count, "idmx_device_count: null");
  IDMX_HIP(hipGetDeviceCount(count));
  return IDMX_OK;
}

int idmx_ctx_create(int device, idmx_ctx_t* out) {
  CHECK_ARG(out, "idmx_ctx_create: null out");
  int n = 0;
  IDMX_HIP(hipGetDeviceCount(&n));
  CHECK_ARG(device >= 0 && device < n, "idmx_ctx_create: no such device");
  IDMX_HIP(hipSetDevice(device));
  idmx_ctx_s* c = new idmx_ctx_s();
  c->device = device;
  hipError_t e = hipStreamCreate(&c->stream);
  if (e == hipSuccess) {
    c->partials_cap = 4 * 1024;
    e = hipMalloc(&c->partials, sizeof(double) * c->partials_cap);
  }
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    idmx_ctx_destroy(c);
    return hip_fail(e, "idmx_ctx_create");
  }
  *out = c;
  return IDMX_OK;
}

int idmx_ctx_destroy(idmx_ctx_t c) {
  if (!c) return IDMX_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->partials) hipFree(c->partials);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return IDMX_OK;
}

int idmx_sync(idmx_ctx_t c) {
  CHECK_ARG(c, "idmx_sync: null context");
  IDMX_HIP(hipStreamSynchronize(c->stream));
  return IDMX_OK;
}

int idmx_device_name(idmx_ctx_t c, char* buf, size_t len) {
  CHECK_ARG(c && buf && len, "idmx_device_name: invalid argument");
  hipDeviceProp_t p;
  IDMX_HIP(hipGetDeviceProperties(&p, c->device));
  snprintf(buf, len, "%s (%s)", p.name, p.gcnArchName);
  return IDMX_OK;
}

int idmx_alloc(idmx_ctx_t c, size_t bytes, void** dptr) {
  CHECK_ARG(c && dptr, "idmx_alloc: invalid argument");
  IDMX_HIP(hipSetDevice(c->device));
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e == hipErrorOutOfMemory) {
    set_error("idmx_alloc: out of device memory");
    return IDMX_E_NOMEM;
  }
  IDMX_HIP(e);
  *dptr = p;
  return IDMX_OK;
}

int idmx_free(idmx_ctx_t c, void* p) {
  CHECK_ARG(c, "idmx_free: null context");
  if (!p) return IDMX_OK;
  IDMX_HIP(hipStreamSynchronize(c->stream));
  IDMX_HIP(hipFree(p));
  return IDMX_OK;
}

int idmx_h2d(idmx_ctx_t c, void* dst, const void* src, size_t bytes) {
  CHECK_ARG(c && dst && src, "idmx_h2d: invalid argument");
  IDMX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  IDMX_HIP(hipStreamSynchronize(c->stream));
  return IDMX_OK;
}

int idmx_d2h(idmx_ctx_t c, void* dst, const void* src, size_t bytes) {
  CHECK_ARG(c && dst && src, "idmx_d2h: invalid argument");
  IDMX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  IDMX_HIP(hipStreamSynchronize(c->stream));
  return IDMX_OK;
}

int idmx_d2d(idmx_ctx_t c, void* dst, const void* src, size_t bytes) {
  CHECK_ARG(c && dst && src, "idmx_d2d: invalid argument");
  IDMX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return IDMX_OK;
}

int idmx_memset(idmx_ctx_t c, void* dst, int value, size_t bytes) {
  CHECK_ARG(c && dst, "idmx_memset: invalid argument");
  IDMX_HIP(hipMemsetAsync(dst, value, bytes, c->stream));
  return IDMX_OK;
}

int idmx_center_of_mass_f64(idmx_ctx_t c, const double* n, int L0, int L1, int L2,
                            double* out) {
  CHECK_ARG(c && n && out && L0 > 0 && L1 > 0 && L2 > 0,
            "idmx_center_of_mass_f64: invalid argument");
  const int64_t V = (int64_t)L0 * L1 * L2;
  const unsigned grid = (unsigned)std::min<int64_t>((V + 255) / 256, 1024);
  CHECK_ARG((size_t)grid * 4 <= c->partials_cap, "idmx_center_of_mass_f64: scratch too small");
  hipLaunchKernelGGL(k_com_partial, dim3(grid), dim3(256), 0, c->stream, n, L0, L1, V,
                     c->partials);
  IDMX_HIP(hipGetLastError());
  std::vector<double> h(4 * (size_t)grid);
  IDMX_HIP(hipMemcpyAsync(h.data(), c->partials, sizeof(double) * h.size(),
                          hipMemcpyDeviceToHost, c->stream));
  IDMX_HIP(hipStreamSynchronize(c->stream));
  double acc[4] = {0, 0, 0, 0};
  for (unsigned b = 0; b < grid; b++)
    for (int k = 0; k < 4; k++) acc[k] += h[4 * b + k];
  out[0] = acc[0];
  for (int k = 1; k < 4; k++) out[k] = acc[k] / acc[0];
  return IDMX_OK;
}

int idmx_sum_f64(idmx_ctx_t c, const double* x, int64_t n, double* out) {
  return reduce<0>(c, x, nullptr, n, out);
}
int idmx_max_f64(idmx_ctx_t c, const double* x, int64_t n, double* out) {
  return reduce<1>(c, x, nullptr, n, out);
}
int idmx_max_abs_diff_f64(idmx_ctx_t c, const double* a, const double* b, int64_t n,
                          double* out) {
  return reduce<2>(c, a, b, n, out);
}

int idmx_stream_copy_f64(idmx_ctx_t c, double* dst, const double* src, int64_t n,
                         double* ms) {
  CHECK_ARG(c && dst && src && n > 0 && ms, "idmx_stream_copy_f64: invalid argument");
  const int64_t blocks = (n + 255) / 256;
  CHECK_ARG(blocks < ((int64_t)1 << 31), "idmx_stream_copy_f64: too large");
  IDMX_HIP(hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0, c->stream, src, dst, n);
  IDMX_HIP(hipGetLastError());
  IDMX_HIP(hipEventRecord(c->ev1, c->stream));
  IDMX_HIP(hipEventSynchronize(c->ev1));
  float t = 0;
  IDMX_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
  *ms = t;
  return IDMX_OK;
}

int idmx_check_u(idmx_ctx_t c, const double* u, double* old_u, int64_t V, int D,
                 double* sum_delta, double* max_u2) {
  CHECK_ARG(c && u && old_u && sum_delta && max_u2 && V > 0 && D >= 1 && D <= 3,
            "idmx_check_u: invalid argument");
  const unsigned grid = (unsigned)std::min<int64_t>((V + 255) / 256, 1024);
  hipLaunchKernelGGL(k_check_u, dim3(grid), dim3(256), 0, c->stream, u, old_u, V, D,
                     c->partials);
  IDMX_HIP(hipGetLastError());
  std::vector<double> h(2 * grid);
  IDMX_HIP(hipMemcpyAsync(h.data(), c->partials, sizeof(double) * 2 * grid,
                          hipMemcpyDeviceToHost, c->stream));
  IDMX_HIP(hipStreamSynchronize(c->stream));
  double sd = 0.0, mu = 0.0;
  for (unsigned k = 0; k < grid; k++) {
    sd += h[2 * k];
    mu = std::max(mu, h[2 * k + 1]);
  }
  *sum_delta = sd;
  *max_u2 = mu;
  return IDMX_OK;
}

int idmx_sc_create(idmx_ctx_t c, const idmx_sc_desc* d, idmx_sc_t* out) {
  CHECK_ARG(c && d && out, "idmx_sc_create: invalid argument");
  CHECK_ARG(d->xi == IDMX_D2Q9 || d->xi == IDMX_D3Q19, "idmx_sc_create: unknown velocity set");
  CHECK_ARG(d->estencil == IDMX_E4 || d->estencil == IDMX_E6 || d->estencil == IDMX_E8,
            "idmx_sc_create: unknown E-stencil");
  CHECK_ARG(d->psi_kind >= 0 && d->psi_kind <= 3, "idmx_sc_create: unknown psi kind");
  CHECK_ARG(d->collision == IDMX_BGK || d->collision == IDMX_MRT ||
                d->collision == IDMX_BGK_SCF || d->collision == IDMX_BGK_ISO,
            "idmx_sc_create: unknown collision");
  CHECK_ARG(d->collision != IDMX_BGK_ISO ||
                (d->xi == IDMX_D2Q9 && d->precision == IDMX_PREC_F64),
            "idmx_sc_create: the isotropy filter needs D2Q9 and FP64");
  CHECK_ARG(d->precision == IDMX_PREC_F64 || d->precision == IDMX_PREC_F32S,
            "idmx_sc_create: unknown precision");
  CHECK_ARG(d->tau > 0, "idmx_sc_create: tau must be positive");
  const int D = d->xi == IDMX_D2Q9 ? 2 : 3;
  const int Q = d->xi == IDMX_D2Q9 ? 9 : 19;
  const int R = d->estencil == IDMX_E4 ? 1 : 2;
  for (int k = 0; k < D; k++)
    CHECK_ARG(d->dims[k] >= 2 * R + 1 && d->dims[k] <= (1 << 20),
              "idmx_sc_create: each lattice size must be >= 2*reach+1 and <= 2^20");
  if (D == 2) CHECK_ARG(d->dims[2] <= 1, "idmx_sc_create: dims[2] must be 1 for D2Q9");
  const int64_t V = (int64_t)d->dims[0] * d->dims[1] * (D == 3 ? d->dims[2] : 1);
  // one population stream (V values) must stay below 4 GiB: the kernels
  // address a site by a 32-bit byte offset from the stream's base
  CHECK_ARG(V * (d->precision == IDMX_PREC_F64 ? 8 : 4) < ((int64_t)1 << 32),
            "idmx_sc_create: lattice too large (V * sizeof(population) must be < 4 GiB)");
  if (d->collision == IDMX_MRT)
    CHECK_ARG(d->mrt_diag || (d->mrt_A && d->mrt_beta && d->mrt_gamma),
              "idmx_sc_create: MRT tables missing");

  idmx_sc_s* s = new idmx_sc_s();
  s->ctx = c;
  s->desc = *d;
  s->desc.mrt_A = s->desc.mrt_beta = s->desc.mrt_gamma = s->desc.mrt_diag = nullptr;
  s->coll = d->collision == IDMX_BGK       ? 0
            : d->collision == IDMX_BGK_SCF ? 3
            : d->collision == IDMX_BGK_ISO ? 4
                                           : (d->mrt_diag ? 2 : 1);
  if (D == 2) s->desc.dims[2] = 1;
  if (d->precision == IDMX_PREC_F64) s->desc.rho0 = 0.0;
  s->D = D;
  s->Q = Q;
  s->geo.L0 = d->dims[0];
  s->geo.L1 = d->dims[1];
  s->geo.L2 = D == 3 ? d->dims[2] : 1;
  s->geo.V = V;
  s->pop_elem = d->precision == IDMX_PREC_F64 ? 8 : 4;
  const bool f64 = d->precision == IDMX_PREC_F64;
  if (D == 2)
    s->ops = f64 ? get_ops_d2q9_f64(d->estencil, s->coll) : get_ops_d2q9_f32(d->estencil, s->coll);
  else if (d->estencil == IDMX_E4)
    s->ops = f64 ? get_ops_d3q19_e4_f64(s->coll) : get_ops_d3q19_e4_f32(s->coll);
  else if (d->estencil == IDMX_E6)
    s->ops = f64 ? get_ops_d3q19_e6_f64(s->coll) : get_ops_d3q19_e6_f32(s->coll);
  else
    s->ops = f64 ? get_ops_d3q19_e8_f64(s->coll) : get_ops_d3q19_e8_f32(s->coll);
  if (!s->ops) {
    delete s;
    set_error("idmx_sc_create: unsupported combination");
    return IDMX_E_UNSUPPORTED;
  }
  // Fused launch shape: rows of 256 (or 64) in 2D, 32 x 8 tiles in 3D; march
  // segments sized to give >= 2048 blocks where the lattice allows it, and
  // at least 8 planes long unless that leaves fewer than 256 blocks: on
  // small lattices a block's march is a chain of dependent plane steps, so
  // short segments run faster (measured per step: 128^2 15.2 us at 8 planes,
  // 5.6 us at 1; 64^3 41.4 us at 8, 24.7 us at 2; 96^3 and up best at 8).
  if (D == 2) {
    s->tx = d->dims[0] >= 256 ? 256 : 64;
    s->ty = 1;
  } else {
    s->tx = 32;
    s->ty = 8;
  }
  {
    const int ntx = (s->geo.L0 + s->tx - 1) / s->tx;
    const int nty = D == 3 ? (s->geo.L1 + s->ty - 1) / s->ty : 1;
    const int Lm = D == 3 ? s->geo.L2 : s->geo.L1;
    const int want_seg = std::max(1, 2048 / std::max(1, ntx * nty));
    int H = (Lm + want_seg - 1) / want_seg;
    const int blocks_at_8 = ntx * nty * ((Lm + 7) / 8);
    H = std::max(H, blocks_at_8 >= 256 ? 8 : (D == 2 ? 1 : 2));
    s->H = std::min(H, Lm);
  }
  if (d->collision == IDMX_MRT) {
    const int NB = D == 2 ? 6 : 10;
    std::vector<double> hd;
    if (s->coll == 2) {
      hd.assign(d->mrt_diag, d->mrt_diag + Q);
    } else {
      hd.resize((size_t)Q * Q + 2 * (size_t)NB * Q);
      memcpy(hd.data(), d->mrt_A, sizeof(double) * Q * Q);
      memcpy(hd.data() + Q * Q, d->mrt_beta, sizeof(double) * NB * Q);
      memcpy(hd.data() + Q * Q + NB * Q, d->mrt_gamma, sizeof(double) * NB * Q);
    }
    const size_t cnt = hd.size();
    hipError_t e = hipMalloc(&s->mrt_dev, cnt * s->pop_elem);
    if (e == hipSuccess) {
      if (s->pop_elem == 8) {
        e = hipMemcpy(s->mrt_dev, hd.data(), cnt * 8, hipMemcpyHostToDevice);
      } else {
        std::vector<float> hf(hd.begin(), hd.end());
        e = hipMemcpy(s->mrt_dev, hf.data(), cnt * 4, hipMemcpyHostToDevice);
      }
    }
    if (e != hipSuccess) {
      if (s->mrt_dev) hipFree(s->mrt_dev);
      delete s;
      return hip_fail(e, "idmx_sc_create: MRT tables");
    }
  }
  *out = s;
  return IDMX_OK;
}

int idmx_sc_destroy(idmx_sc_t s) {
  if (!s) return IDMX_OK;
  if (s->mrt_dev) {
    hipStreamSynchronize(s->ctx->stream);
    hipFree(s->mrt_dev);
  }
  for (hipEvent_t e : s->prof_ev) hipEventDestroy(e);
  if (s->psi_prog_dev || s->wmask_dev) {
    hipStreamSynchronize(s->ctx->stream);
    if (s->psi_prog_dev) hipFree(s->psi_prog_dev);
    if (s->wmask_dev) hipFree(s->wmask_dev);
  }
  delete s;
  return IDMX_OK;
}

int idmx_sc_pop_bytes(idmx_sc_t s, size_t* bytes) {
  CHECK_ARG(s && bytes, "idmx_sc_pop_bytes: invalid argument");
  *bytes = (size_t)s->Q * (size_t)s->geo.V * s->pop_elem;
  return IDMX_OK;
}

int idmx_sc_set_launch(idmx_sc_t s, int tx, int ty, int H, int hist) {
  CHECK_ARG(s, "idmx_sc_set_launch: null simulation");
  if (s->D == 2)
    CHECK_ARG((tx == 64 || tx == 128 || tx == 256) && ty == 1,
              "idmx_sc_set_launch: 2D tiles are 64, 128 or 256 sites wide, ty = 1");
  else
    CHECK_ARG((tx == 32 && ty == 8) || (tx == 64 && ty == 4) || (tx == 128 && ty == 2) ||
                  (tx == 256 && ty == 2),
              "idmx_sc_set_launch: 3D tiles are 32x8, 64x4, 128x2 or 256x2");
  CHECK_ARG(H >= 1, "idmx_sc_set_launch: H must be >= 1");
  CHECK_ARG(hist >= 0 && hist <= 10, "idmx_sc_set_launch: variant is 0..10");
  CHECK_ARG(s->ops->launch_ok(tx, ty, hist),
            "idmx_sc_set_launch: the LDS staging of variants 9 and 10 does not fit this tile");
  s->tx = tx;
  s->ty = ty;
  s->H = H;
  s->hist = hist;
  return IDMX_OK;
}

int idmx_sc_set_psi_program(idmx_sc_t s, const unsigned char* ops, int nops,
                            const double* consts, int nconsts) {
  CHECK_ARG(s && ops && nops > 0 && nops <= kPsiOps && nconsts >= 0 && nconsts <= kPsiConsts &&
                (nconsts == 0 || consts),
            "idmx_sc_set_psi_program: invalid argument");
  CHECK_ARG(s->desc.psi_kind == IDMX_PSI_PROGRAM,
            "idmx_sc_set_psi_program: the simulation was not created with IDMX_PSI_PROGRAM");
  // validate: known ops, constants used exactly, stack within bounds, one result
  PsiProg p;
  p.n = nops;
  int depth = 0, used = 0;
  for (int i = 0; i < nops; i++) {
    const int op = ops[i];
    CHECK_ARG(op < PSI_NOPS, "idmx_sc_set_psi_program: unknown op");
    if (op == PSI_N || op == PSI_C) {
      depth++;
      if (op == PSI_C) used++;
    } else if (op == PSI_ADD || op == PSI_SUB || op == PSI_MUL || op == PSI_DIV ||
               op == PSI_POW || op == PSI_FMAX || op == PSI_FMIN) {
      depth--;
    }
    CHECK_ARG(depth >= 1 && depth <= kPsiStack && used <= nconsts,
              "idmx_sc_set_psi_program: malformed program (stack or constants)");
    p.op[i] = (unsigned char)op;
  }
  CHECK_ARG(depth == 1 && used == nconsts, "idmx_sc_set_psi_program: malformed program");
  for (int i = 0; i < nconsts; i++) p.c[i] = consts[i];
  if (!s->psi_prog_dev) IDMX_HIP(hipMalloc(&s->psi_prog_dev, sizeof(PsiProg)));
  IDMX_HIP(hipMemcpyAsync(s->psi_prog_dev, &p, sizeof(PsiProg), hipMemcpyHostToDevice,
                          s->ctx->stream));
  IDMX_HIP(hipStreamSynchronize(s->ctx->stream));
  return IDMX_OK;
}

int idmx_sc_set_walls(idmx_sc_t s, const void* walls) {
  CHECK_ARG(s, "idmx_sc_set_walls: null simulation");
  // the per-site masks of the fused step with walls (FP64 BGK), allocated
  // here so that no run allocates
  if (walls && !s->wmask_dev && s->pop_elem == 8 && s->coll == 0)
    IDMX_HIP(hipMalloc(&s->wmask_dev, sizeof(uint64_t) * (size_t)s->geo.V));
  s->walls = (const uint8_t*)walls;
  return IDMX_OK;
}

int idmx_sc_set_body_force(idmx_sc_t s, double fx, double fy, double fz) {
  CHECK_ARG(s, "idmx_sc_set_body_force: null simulation");
  CHECK_ARG(std::isfinite(fx) && std::isfinite(fy) && std::isfinite(fz),
            "idmx_sc_set_body_force: non-finite force");
  s->bforce[0] = fx;
  s->bforce[1] = fy;
  s->bforce[2] = s->D == 3 ? fz : 0.0;
  return IDMX_OK;
}

int idmx_sc_get_launch(idmx_sc_t s, int* tx, int* ty, int* H, int* hist) {
  CHECK_ARG(s && tx && ty && H && hist, "idmx_sc_get_launch: invalid argument");
  *tx = s->tx;
  *ty = s->ty;
  *H = s->H;
  *hist = s->hist;
  return IDMX_OK;
}

int idmx_sc_init_pop(idmx_sc_t s, void* pop, const double* n, const double* u) {
  CHECK_ARG(s && pop && n && u, "idmx_sc_init_pop: invalid argument");
  return s->ops->init_pop(s, pop, n, u);
}

int idmx_sc_init_pop_coeffs(idmx_sc_t s, void* pop, const double* n, const double* u,
                            const double* coeffs) {
  CHECK_ARG(s && pop && n && u && coeffs, "idmx_sc_init_pop_coeffs: invalid argument");
  for (int i = 0; i < (s->D == 2 ? 6 : 10) * s->Q; i++)
    CHECK_ARG(std::isfinite(coeffs[i]), "idmx_sc_init_pop_coeffs: non-finite coefficient");
  return s->ops->init_pop_tab(s, pop, n, u, coeffs);
}

int idmx_sc_density_psi(idmx_sc_t s, double* n, double* psi, const void* pop) {
  CHECK_ARG(s && n && psi && pop, "idmx_sc_density_psi: invalid argument");
  CHECK_ARG(s->desc.psi_kind != IDMX_PSI_PROGRAM || s->psi_prog_dev,
            "idmx_sc_density_psi: IDMX_PSI_PROGRAM without a program");
  return s->ops->density_psi(s, n, psi, pop);
}

int idmx_sc_force_collide_stream(idmx_sc_t s, void* swap, const void* pop,
                                 const double* n, const double* psi) {
  CHECK_ARG(s && swap && pop && n && psi && swap != pop,
            "idmx_sc_force_collide_stream: invalid argument");
  return s->ops->fcs(s, swap, pop, n, psi, nullptr);
}

int idmx_sc_run(idmx_sc_t s, void** pop, void** swap, double* n, double* psi, double* u,
                int nsteps, int flags) {
  CHECK_ARG(s && pop && swap && *pop && *swap && *pop != *swap,
            "idmx_sc_run: invalid population buffers");
  CHECK_ARG(nsteps >= 0, "idmx_sc_run: nsteps must be >= 0");
  CHECK_ARG(s->desc.psi_kind != IDMX_PSI_PROGRAM || s->psi_prog_dev,
            "idmx_sc_run: IDMX_PSI_PROGRAM without a program (idmx_sc_set_psi_program)");
  if (nsteps == 0) {
    s->last_ms = 0;
    return IDMX_OK;
  }
  return s->ops->run(s, pop, swap, n, psi, u, nsteps, flags);
}

int idmx_sc_run_profiled(idmx_sc_t s, void** pop, void** swap, double* n, double* psi,
                         double* u, int nsteps, int flags, float* step_ms) {
  CHECK_ARG(s && step_ms, "idmx_sc_run_profiled: invalid argument");
  CHECK_ARG(nsteps >= 0 && nsteps <= (1 << 24), "idmx_sc_run_profiled: nsteps out of range");
  s->prof_ms = step_ms;
  const int rc = idmx_sc_run(s, pop, swap, n, psi, u, nsteps, flags);
  s->prof_ms = nullptr;
  return rc;
}

int idmx_sc_moments(idmx_sc_t s, const void* pop, double* n, double* psi, double* u) {
  CHECK_ARG(s && pop && n && psi && u, "idmx_sc_moments: invalid argument");
  CHECK_ARG(s->desc.psi_kind != IDMX_PSI_PROGRAM || s->psi_prog_dev,
            "idmx_sc_moments: IDMX_PSI_PROGRAM without a program");
  return s->ops->moments(s, pop, n, psi, u);
}

int idmx_sc_density_velocity(idmx_sc_t s, const void* pop, double* n, double* u) {
  CHECK_ARG(s && pop && n && u, "idmx_sc_density_velocity: invalid argument");
  return s->ops->density_velocity(s, pop, n, u);
}

int idmx_sc_mass_momentum(idmx_sc_t s, const void* pop, double* mass, double* mom) {
  CHECK_ARG(s && pop && mass && mom, "idmx_sc_mass_momentum: invalid argument");
  return s->ops->mass_mom(s, pop, mass, mom);
}

int idmx_sc_pop_to_f64(idmx_sc_t s, const void* pop, double* dst) {
  CHECK_ARG(s && pop && dst, "idmx_sc_pop_to_f64: invalid argument");
  return s->ops->to_f64(s, pop, dst);
}

int idmx_sc_pop_from_f64(idmx_sc_t s, void* pop, const double* src) {
  CHECK_ARG(s && pop && src, "idmx_sc_pop_from_f64: invalid argument");
  return s->ops->from_f64(s, pop, src);
}

int idmx_sc_last_run_ms(idmx_sc_t s, double* ms) {
  CHECK_ARG(s && ms, "idmx_sc_last_run_ms: invalid argument");
  *ms = s->last_ms;
  return IDMX_OK;
}

}  // extern "C"
