// batch_impl.h -- state of one batched sweep (idmx_batch_t) and the step
// launchers shared by batch.hip (host API, FP64 step) and batch_f32.hip
// (FP32-shifted step). The FP64 unit is compiled without floating-point
// contraction and the FP32 one with it, matching the single-domain kernels
// of the same precision so that a batched domain follows the single-domain
// trajectory.
#pragma once

#include <algorithm>
#include <vector>

#include "batch_kernels.h"
#include "internal.h"

struct idmx_batch_s {
  idmx_ctx_s* ctx = nullptr;
  int prec = 0, estencil = 4;
  int kind = 1;  // 0: marching items (tx x H rows), 1: tiles (tx x H sites), 2/3: 0/1 on the
                 // pull form (converted before and after each run), 4: tiles advanced two
                 // steps per pass (E4; single steps with the kind-1 kernel)
  int tx = 32, H = 8;
  int ndom = 0;
  std::vector<idmx::BatchDom> doms;
  std::vector<int> active;
  size_t elem = 8;
  int64_t pop_elems = 0, sites = 0;
  void* pop[2] = {nullptr, nullptr};
  int cur = 0;  // index of the buffer holding the state
  double* n_prev = nullptr;
  double* stage = nullptr;  // per-domain FP64 staging (max V)
  int64_t max_v = 0;
  idmx::BatchDom* d_doms = nullptr;
  idmx::BatchItem* d_items = nullptr;
  int* d_nitems = nullptr;
  unsigned long long* d_dmax = nullptr;
  int max_items = 0, n_items = 0;
  int grid = 0;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  int graph_grid = 0;  // step grid the graph was captured with
  int graph_cur = 0;   // buffer holding the state when the graph starts
  double last_ms = 0;
};

namespace idmx {

// Blocks of a step launch: the tuned grid, but never more than one per live
// work item. The step kernels order their items per XCD (xcd_block), which
// spreads the items over all XCDs only when no block is surplus.
inline int step_grid(const idmx_batch_s* b) {
  return std::max(1, std::min(b->grid, b->n_items));
}

template <class ES, class T, bool SHIFT, int TX, bool PULL>
hipError_t launch_march(idmx_batch_s* b, const void* in, void* out) {
  hipLaunchKernelGGL((k_batched<ES, T, SHIFT, TX, PULL>), dim3(step_grid(b)), dim3(TX), 0,
                     b->ctx->stream, (const T*)in, (T*)out, b->d_doms, b->d_items, b->d_nitems);
  return hipGetLastError();
}

template <class ES, class T, bool SHIFT, int TX>
hipError_t launch_step(idmx_batch_s* b, const void* in, void* out) {
  return b->kind == 2 ? launch_march<ES, T, SHIFT, TX, true>(b, in, out)
                      : launch_march<ES, T, SHIFT, TX, false>(b, in, out);
}

template <class ES, class T, bool SHIFT, int BX, int BY>
hipError_t launch_tile(idmx_batch_s* b, const void* in, void* out) {
  if (b->kind == 3)
    hipLaunchKernelGGL((k_batched_tile<ES, T, SHIFT, BX, BY, true>), dim3(step_grid(b)),
                       dim3(BX * BY), 0, b->ctx->stream, (const T*)in, (T*)out, b->d_doms,
                       b->d_items, b->d_nitems);
  else
    hipLaunchKernelGGL((k_batched_tile<ES, T, SHIFT, BX, BY, false>), dim3(step_grid(b)),
                       dim3(BX * BY), 0, b->ctx->stream, (const T*)in, (T*)out, b->d_doms,
                       b->d_items, b->d_nitems);
  return hipGetLastError();
}

template <class ES, class T, bool SHIFT>
hipError_t step_tx(idmx_batch_s* b, const void* in, void* out) {
  if (b->kind == 1 || b->kind == 3 || b->kind == 4 || b->kind == 5) {
    if (b->tx == 64) return launch_tile<ES, T, SHIFT, 64, 4>(b, in, out);
    if (b->tx == 16) return launch_tile<ES, T, SHIFT, 16, 16>(b, in, out);
    if (b->H == 4) return launch_tile<ES, T, SHIFT, 32, 4>(b, in, out);
    if (b->H == 16) return launch_tile<ES, T, SHIFT, 32, 16>(b, in, out);
    return launch_tile<ES, T, SHIFT, 32, 8>(b, in, out);
  }
  switch (b->tx) {
    case 64: return launch_step<ES, T, SHIFT, 64>(b, in, out);
    case 256: return launch_step<ES, T, SHIFT, 256>(b, in, out);
    default: return launch_step<ES, T, SHIFT, 128>(b, in, out);
  }
}

template <class T, bool SHIFT>
hipError_t step_es(idmx_batch_s* b, const void* in, void* out) {
  switch (b->estencil) {
    case IDMX_E6: return step_tx<D2E6, T, SHIFT>(b, in, out);
    case IDMX_E8: return step_tx<D2E8, T, SHIFT>(b, in, out);
    default: return step_tx<D2E4, T, SHIFT>(b, in, out);
  }
}

// Two steps of every live domain in one launch (kinds 4 and 5, E4 tiles).
template <class T, bool SHIFT, int BX, int BY>
hipError_t launch_tile2(idmx_batch_s* b, const void* in, void* out, int xsh) {
  hipLaunchKernelGGL((k_batched_tile2<T, SHIFT, BX, BY>), dim3(step_grid(b)), dim3(BX * BY), 0,
                     b->ctx->stream, (const T*)in, (T*)out, b->d_doms, b->d_items, b->d_nitems,
                     xsh);
  return hipGetLastError();
}

template <class T, bool SHIFT>
hipError_t two_steps_tx(idmx_batch_s* b, const void* in, void* out, int xsh) {
  if (b->tx == 16) return launch_tile2<T, SHIFT, 16, 16>(b, in, out, xsh);
  if (b->H == 16) return launch_tile2<T, SHIFT, 32, 16>(b, in, out, xsh);
  return launch_tile2<T, SHIFT, 32, 8>(b, in, out, xsh);
}

// One step of every live domain, in the FP64 and the FP32-shifted precision,
// and two steps in one launch (kinds 4 and 5; xsh: the x-shifted storage
// flags of k_batched_tile2).
hipError_t batch_step_f64(idmx_batch_s* b, const void* in, void* out);
hipError_t batch_step_f32(idmx_batch_s* b, const void* in, void* out);
hipError_t batch_two_steps_f64(idmx_batch_s* b, const void* in, void* out, int xsh);
hipError_t batch_two_steps_f32(idmx_batch_s* b, const void* in, void* out, int xsh);
// FP32-shifted populations of one domain at rest equilibrium of density n.
hipError_t batch_init_f32(idmx_batch_s* b, float* pop, const double* n, int64_t V, double rho0,
                          unsigned grid);

}  // namespace idmx
