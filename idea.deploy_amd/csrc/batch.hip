// batch.hip -- batched Shan-Chen sweep over many D2Q9 domains (C ABI in
// include/idmx.h, idmx_batch_*). The library owns the packed population
// buffers; a step is one launch of k_batched (march items) or
// k_batched_tile (tiles) over the live work items, in the push or the pull
// form, and idmx_batch_run replays a captured graph of an even number of
// such launches (pull form: between a conversion to it and one back).
#include <algorithm>
#include <cstring>
#include <vector>

#include "batch_impl.h"

using namespace idmx;

hipError_t idmx::batch_step_f64(idmx_batch_s* b, const void* in, void* out) {
  return step_es<double, false>(b, in, out);
}

hipError_t idmx::batch_two_steps_f64(idmx_batch_s* b, const void* in, void* out, int xsh) {
  return two_steps_tx<double, false>(b, in, out, xsh);
}

namespace {

// Streamed <-> pull form of every live domain (kinds 2 and 3).
hipError_t restream(idmx_batch_s* b, int dir) {
  const void* in = b->pop[b->cur];
  void* out = b->pop[b->cur ^ 1];
  if (b->prec == IDMX_PREC_F64)
    hipLaunchKernelGGL(k_batched_restream<double>, dim3(b->grid), dim3(256), 0, b->ctx->stream,
                       (const double*)in, (double*)out, b->d_doms, b->d_items, b->d_nitems,
                       b->tx, dir);
  else
    hipLaunchKernelGGL(k_batched_restream<float>, dim3(b->grid), dim3(256), 0, b->ctx->stream,
                       (const float*)in, (float*)out, b->d_doms, b->d_items, b->d_nitems, b->tx,
                       dir);
  return hipGetLastError();
}

// Tile shapes (tx x H) accepted by idmx_batch_set_launch for kinds 1 and 3,
// and for kinds 4 and 5 (two steps per pass).
bool tile_shape_ok(int tx, int H) {
  return (tx == 64 && H == 4) || (tx == 32 && H == 8) || (tx == 16 && H == 16) ||
         (tx == 32 && H == 4);
}
bool tile2_shape_ok(int tx, int H) {
  return (tx == 32 && H == 8) || (tx == 16 && H == 16) || (tx == 32 && H == 16);
}

hipError_t one_step(idmx_batch_s* b, const void* in, void* out) {
  return b->prec == IDMX_PREC_F64 ? batch_step_f64(b, in, out) : batch_step_f32(b, in, out);
}

hipError_t two_steps(idmx_batch_s* b, const void* in, void* out, int xsh = 0) {
  return b->prec == IDMX_PREC_F64 ? batch_two_steps_f64(b, in, out, xsh)
                                  : batch_two_steps_f32(b, in, out, xsh);
}

template <class T, bool SHIFT>
hipError_t launch_dn(idmx_batch_s* b) {
  hipLaunchKernelGGL((k_batched_dn<T, SHIFT>), dim3(b->grid), dim3(256), 0, b->ctx->stream,
                     (const T*)b->pop[b->cur], b->d_doms, b->d_items, b->d_nitems, b->tx,
                     b->n_prev, b->d_dmax);
  return hipGetLastError();
}

void drop_graph(idmx_batch_s* b) {
  if (b->graph) hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_steps = 0;
}

// Rebuilds the work-item list from the active flags: for every live domain,
// x-chunks of tx sites and runs of H rows (kind 0), or tx x H tiles (kind 1;
// the last tile of a row or column may reach past the domain edge).
int rebuild_items(idmx_batch_s* b) {
  std::vector<BatchItem> items;
  for (int d = 0; d < b->ndom; d++) {
    if (!b->active[d]) continue;
    const BatchDom& D = b->doms[d];
    for (int m0 = 0; m0 < D.L1; m0 += b->H)
      for (int x0 = 0; x0 < D.L0; x0 += b->tx)
        items.push_back({d, x0, m0, std::min(b->H, D.L1 - m0)});
  }
  if ((int)items.size() > b->max_items) {
    if (b->d_items) hipFree(b->d_items);
    b->max_items = std::max<int>((int)items.size(), 1);
    IDMX_HIP(hipMalloc(&b->d_items, sizeof(BatchItem) * b->max_items));
    drop_graph(b);  // the graph holds the old item pointer
  }
  b->n_items = (int)items.size();
  // the graph's step launches hold the grid they were captured with
  if (b->graph && step_grid(b) != b->graph_grid) drop_graph(b);
  if (!items.empty())
    IDMX_HIP(hipMemcpyAsync(b->d_items, items.data(), sizeof(BatchItem) * items.size(),
                            hipMemcpyHostToDevice, b->ctx->stream));
  IDMX_HIP(hipMemcpyAsync(b->d_nitems, &b->n_items, sizeof(int), hipMemcpyHostToDevice,
                          b->ctx->stream));
  IDMX_HIP(hipStreamSynchronize(b->ctx->stream));
  return IDMX_OK;
}

}  // namespace

#define BCHECK(cond, msg)     \
  do {                        \
    if (!(cond)) {            \
      idmx::set_error(msg);   \
      return IDMX_E_ARG;      \
    }                         \
  } while (0)

namespace {
// k_batched_force_rest for one domain in the batch's storage.
template <class ES>
void launch_force_rest(idmx_batch_s* b, const BatchDom& D, void* pop, const double* n,
                       unsigned grid, hipStream_t st) {
  if (b->prec == IDMX_PREC_F64)
    hipLaunchKernelGGL((k_batched_force_rest<ES, double, false>), dim3(grid), dim3(256), 0, st,
                       n, (double*)pop, D.L0, D.L1, D.G, D.psi_kind, 0.0);
  else
    hipLaunchKernelGGL((k_batched_force_rest<ES, float, true>), dim3(grid), dim3(256), 0, st,
                       n, (float*)pop, D.L0, D.L1, D.G, D.psi_kind, D.rho0);
}
}  // namespace

extern "C" {

int idmx_batch_create(idmx_ctx_t ctx, const idmx_batch_domain* doms, int ndom, int precision,
                      int estencil, int psi_kind, idmx_batch_t* out) {
  BCHECK(ctx && doms && ndom > 0 && out, "idmx_batch_create: invalid argument");
  BCHECK(precision == IDMX_PREC_F64 || precision == IDMX_PREC_F32S,
         "idmx_batch_create: unknown precision");
  BCHECK(estencil == IDMX_E4 || estencil == IDMX_E6 || estencil == IDMX_E8,
         "idmx_batch_create: unknown E-stencil");
  BCHECK(psi_kind >= 0 && psi_kind <= 2, "idmx_batch_create: unknown psi kind");
  const int R = estencil == IDMX_E4 ? 1 : 2;
  idmx_batch_s* b = new idmx_batch_s();
  b->ctx = ctx;
  b->prec = precision;
  b->estencil = estencil;
  b->ndom = ndom;
  b->elem = precision == IDMX_PREC_F64 ? 8 : 4;
  int64_t po = 0, so = 0;
  for (int d = 0; d < ndom; d++) {
    const idmx_batch_domain& s = doms[d];
    if (s.L0 < 2 * R + 1 || s.L1 < 2 * R + 1 || s.L0 > 65536 || s.L1 > 65536 || s.tau <= 0 ||
        (int64_t)s.L0 * s.L1 * (int64_t)b->elem >= ((int64_t)1 << 32)) {
      delete b;
      set_error("idmx_batch_create: domain sizes must be in [2R+1, 65536] with "
                "L0 L1 sizeof(population) < 4 GiB, tau > 0");
      return IDMX_E_ARG;
    }
    BatchDom D{};
    D.L0 = s.L0;
    D.L1 = s.L1;
    D.pop_off = po;
    D.site_off = so;
    D.G = s.SC_G;
    D.omega = 1.0 / s.tau;
    D.rho0 = precision == IDMX_PREC_F32S ? s.rho0 : 0.0;
    D.psi_kind = psi_kind;
    const int64_t V = (int64_t)s.L0 * s.L1;
    po += 9 * V;
    so += V;
    b->max_v = std::max(b->max_v, V);
    b->doms.push_back(D);
  }
  b->pop_elems = po;
  b->sites = so;
  b->active.assign(ndom, 1);
  hipError_t e = hipMalloc(&b->pop[0], b->elem * po);
  if (e == hipSuccess) e = hipMalloc(&b->pop[1], b->elem * po);
  if (e == hipSuccess) e = hipMalloc(&b->n_prev, sizeof(double) * so);
  if (e == hipSuccess) e = hipMalloc(&b->stage, sizeof(double) * b->max_v);
  if (e == hipSuccess) e = hipMalloc(&b->d_doms, sizeof(BatchDom) * ndom);
  if (e == hipSuccess) e = hipMalloc(&b->d_nitems, sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&b->d_dmax, sizeof(unsigned long long) * ndom);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_doms, b->doms.data(), sizeof(BatchDom) * ndom, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(b->n_prev, 0, sizeof(double) * so);
  if (e != hipSuccess) {
    idmx_batch_destroy(b);
    return hip_fail(e, "idmx_batch_create");
  }
  hipDeviceProp_t p;
  hipGetDeviceProperties(&p, ctx->device);
  b->grid = p.multiProcessorCount * 8;
  int rc = rebuild_items(b);
  if (rc) {
    idmx_batch_destroy(b);
    return rc;
  }
  *out = b;
  return IDMX_OK;
}

int idmx_batch_destroy(idmx_batch_t b) {
  if (!b) return IDMX_OK;
  if (b->ctx && b->ctx->stream) hipStreamSynchronize(b->ctx->stream);
  drop_graph(b);
  for (void* p : {b->pop[0], b->pop[1], (void*)b->n_prev, (void*)b->stage, (void*)b->d_doms,
                  (void*)b->d_items, (void*)b->d_nitems, (void*)b->d_dmax})
    if (p) hipFree(p);
  delete b;
  return IDMX_OK;
}

int idmx_batch_set_launch(idmx_batch_t b, int kind, int tx, int H, int grid) {
  BCHECK(b, "idmx_batch_set_launch: null batch");
  BCHECK(kind >= 0 && kind <= 5,
         "idmx_batch_set_launch: kind is 0 (march), 1 (tile), 2 / 3 (pull-form march / tile), "
         "4 (tile, two steps per pass), 5 (4 on the x-shifted storage)");
  if (kind == 0 || kind == 2) {
    BCHECK(tx == 64 || tx == 128 || tx == 256, "idmx_batch_set_launch: march tx is 64, 128 or 256");
    BCHECK(H >= 1, "idmx_batch_set_launch: H >= 1");
  } else if (kind == 4 || kind == 5) {
    BCHECK(b->estencil == IDMX_E4, "idmx_batch_set_launch: kinds 4 and 5 need the E4 force");
    BCHECK(tile2_shape_ok(tx, H),
           "idmx_batch_set_launch: kind 4 and 5 tiles are 32x8, 16x16, 32x16");
  } else {
    BCHECK(tile_shape_ok(tx, H),
           "idmx_batch_set_launch: tile is 64x4, 32x8, 16x16 or 32x4");
  }
  BCHECK(grid >= 0, "idmx_batch_set_launch: grid >= 0");
  b->kind = kind;
  b->tx = tx;
  b->H = H;
  if (grid > 0) b->grid = grid;
  drop_graph(b);
  return rebuild_items(b);
}

int idmx_batch_get_launch(idmx_batch_t b, int* kind, int* tx, int* H, int* grid,
                          int* n_items) {
  BCHECK(b && kind && tx && H && grid && n_items, "idmx_batch_get_launch: invalid argument");
  *kind = b->kind;
  *tx = b->tx;
  *H = b->H;
  *grid = b->grid;
  *n_items = b->n_items;
  return IDMX_OK;
}

int idmx_batch_set_active(idmx_batch_t b, const int* active) {
  BCHECK(b && active, "idmx_batch_set_active: invalid argument");
  for (int d = 0; d < b->ndom; d++) b->active[d] = active[d] ? 1 : 0;
  return rebuild_items(b);
}

int idmx_batch_set_density(idmx_batch_t b, int dom, const double* n_host) {
  BCHECK(b && n_host && dom >= 0 && dom < b->ndom, "idmx_batch_set_density: invalid argument");
  const BatchDom& D = b->doms[dom];
  const int64_t V = (int64_t)D.L0 * D.L1;
  IDMX_HIP(hipMemcpyAsync(b->stage, n_host, sizeof(double) * V, hipMemcpyHostToDevice,
                          b->ctx->stream));
  const unsigned grid = (unsigned)std::min<int64_t>((V + 255) / 256, 4096);
  if (b->prec == IDMX_PREC_F64)
    hipLaunchKernelGGL((k_batched_init<double, false>), dim3(grid), dim3(256), 0, b->ctx->stream,
                       (double*)b->pop[b->cur] + D.pop_off, b->stage, V, 0.0);
  else
    IDMX_HIP(batch_init_f32(b, (float*)b->pop[b->cur] + D.pop_off, b->stage, V, D.rho0, grid));
  IDMX_HIP(hipGetLastError());
  IDMX_HIP(hipMemcpyAsync(b->n_prev + D.site_off, b->stage, sizeof(double) * V,
                          hipMemcpyDeviceToDevice, b->ctx->stream));
  IDMX_HIP(hipStreamSynchronize(b->ctx->stream));
  return IDMX_OK;
}

int idmx_batch_get_density(idmx_batch_t b, int dom, double* n_host) {
  BCHECK(b && n_host && dom >= 0 && dom < b->ndom, "idmx_batch_get_density: invalid argument");
  const BatchDom& D = b->doms[dom];
  const int64_t V = (int64_t)D.L0 * D.L1;
  const unsigned grid = (unsigned)std::min<int64_t>((V + 255) / 256, 4096);
  if (b->prec == IDMX_PREC_F64)
    hipLaunchKernelGGL((k_batched_density<double, false>), dim3(grid), dim3(256), 0,
                       b->ctx->stream, (const double*)b->pop[b->cur] + D.pop_off, b->stage, V,
                       0.0);
  else
    hipLaunchKernelGGL((k_batched_density<float, true>), dim3(grid), dim3(256), 0,
                       b->ctx->stream, (const float*)b->pop[b->cur] + D.pop_off, b->stage, V,
                       D.rho0);
  IDMX_HIP(hipGetLastError());
  IDMX_HIP(hipMemcpyAsync(n_host, b->stage, sizeof(double) * V, hipMemcpyDeviceToHost,
                          b->ctx->stream));
  IDMX_HIP(hipStreamSynchronize(b->ctx->stream));
  return IDMX_OK;
}

int idmx_batch_run(idmx_batch_t b, int nsteps) {
  BCHECK(b && nsteps >= 0, "idmx_batch_run: invalid argument");
  hipStream_t st = b->ctx->stream;
  const bool pull = (b->kind == 2 || b->kind == 3) && b->n_items > 0 && nsteps > 0;
  // Inactive domains are never written by the step or conversion launches:
  // their state stays in the buffer that was current on entry.
  const int cur_in = b->cur;
  // The graph (an even number of steps: A->B, B->A, ...; replays keep the
  // state in the same buffer) is captured before the timed region, for the
  // buffer parity the steps will see: flipped by the conversion to the
  // pull form.
  // kinds 4 and 5 advance two steps per launch, so their graph holds an even
  // number of launches (a multiple of 4 steps) to end in the buffer it
  // started in. Kind 5 keeps the populations in the x-shifted storage between
  // the graph's launches: the first reads the streamed storage, the last
  // writes it, so every replay starts and ends streamed.
  const bool pairs = b->kind == 4 || b->kind == 5;
  const int G = pairs ? 4 * std::max(1, std::min(nsteps / 4, 25))
                      : 2 * std::max(1, std::min(nsteps / 2, 50));
  const int cur_steps = b->cur ^ (pull ? 1 : 0);
  if (b->n_items > 0 && nsteps >= G && (b->graph_steps != G || b->graph_cur != cur_steps)) {
    drop_graph(b);
    hipGraph_t g;
    IDMX_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int nl = pairs ? G / 2 : G;
    for (int k = 0; k < nl; k++) {
      const int src = (cur_steps + k) & 1;
      const int xsh = b->kind == 5 ? (k > 0 ? 1 : 0) | (k < nl - 1 ? 2 : 0) : 0;
      hipError_t e = pairs ? two_steps(b, b->pop[src], b->pop[src ^ 1], xsh)
                           : one_step(b, b->pop[src], b->pop[src ^ 1]);
      if (e != hipSuccess) {
        hipStreamEndCapture(st, &g);
        return hip_fail(e, "idmx_batch_run capture");
      }
    }
    IDMX_HIP(hipStreamEndCapture(st, &g));
    IDMX_HIP(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
    hipGraphDestroy(g);
    b->graph_steps = G;
    b->graph_grid = step_grid(b);
    b->graph_cur = cur_steps;
  }
  IDMX_HIP(hipEventRecord(b->ctx->ev0, st));
  if (pull) {
    // to the pull form (the buffer flips; the conversion back flips it again)
    hipError_t e = restream(b, 1);
    if (e != hipSuccess) return hip_fail(e, "idmx_batch_run restream");
    b->cur ^= 1;
  }
  if (b->n_items > 0 && nsteps > 0) {
    int done = 0;
    // a graph replays the buffers it was captured with; cur is unchanged
    // after every replay
    while (b->graph && b->graph_cur == b->cur && nsteps - done >= b->graph_steps) {
      IDMX_HIP(hipGraphLaunch(b->graph, st));
      done += b->graph_steps;
    }
    while (done < nsteps) {
      const bool two = pairs && nsteps - done >= 2;
      hipError_t e = two ? two_steps(b, b->pop[b->cur], b->pop[b->cur ^ 1])
                         : one_step(b, b->pop[b->cur], b->pop[b->cur ^ 1]);
      if (e != hipSuccess) return hip_fail(e, "idmx_batch_run launch");
      b->cur ^= 1;
      done += two ? 2 : 1;
    }
  }
  if (pull) {
    hipError_t e = restream(b, -1);
    if (e != hipSuccess) return hip_fail(e, "idmx_batch_run restream");
    b->cur ^= 1;
  }
  if (b->cur != cur_in) {
    // odd number of single steps: carry the inactive domains' state over
    // into the buffer that is now current
    for (int d = 0; d < b->ndom; d++) {
      if (b->active[d]) continue;
      const BatchDom& D = b->doms[d];
      const size_t off = b->elem * (size_t)D.pop_off;
      IDMX_HIP(hipMemcpyAsync((char*)b->pop[b->cur] + off, (const char*)b->pop[cur_in] + off,
                              b->elem * 9 * (size_t)D.L0 * D.L1, hipMemcpyDeviceToDevice, st));
    }
  }
  IDMX_HIP(hipEventRecord(b->ctx->ev1, st));
  IDMX_HIP(hipEventSynchronize(b->ctx->ev1));
  float ms = 0;
  IDMX_HIP(hipEventElapsedTime(&ms, b->ctx->ev0, b->ctx->ev1));
  b->last_ms = ms;
  return IDMX_OK;
}

int idmx_batch_check(idmx_batch_t b, double* dmax) {
  BCHECK(b && dmax, "idmx_batch_check: invalid argument");
  hipStream_t st = b->ctx->stream;
  IDMX_HIP(hipMemsetAsync(b->d_dmax, 0, sizeof(unsigned long long) * b->ndom, st));
  if (b->n_items > 0) {
    hipError_t e =
        b->prec == IDMX_PREC_F64 ? launch_dn<double, false>(b) : launch_dn<float, true>(b);
    if (e != hipSuccess) return hip_fail(e, "idmx_batch_check");
  }
  std::vector<unsigned long long> h(b->ndom);
  IDMX_HIP(hipMemcpyAsync(h.data(), b->d_dmax, sizeof(unsigned long long) * b->ndom,
                          hipMemcpyDeviceToHost, st));
  IDMX_HIP(hipStreamSynchronize(st));
  for (int d = 0; d < b->ndom; d++) {
    double v;
    memcpy(&v, &h[d], sizeof(double));
    dmax[d] = b->active[d] ? v : 0.0;
  }
  return IDMX_OK;
}

int idmx_batch_copy_state(idmx_batch_t dst, idmx_batch_t src, int flags) {
  BCHECK(dst && src && dst->ndom == src->ndom, "idmx_batch_copy_state: batches differ");
  BCHECK(flags == 0 || flags == IDMX_BATCH_COPY_REST, "idmx_batch_copy_state: unknown flags");
  for (int d = 0; d < dst->ndom; d++)
    BCHECK(dst->doms[d].L0 == src->doms[d].L0 && dst->doms[d].L1 == src->doms[d].L1,
           "idmx_batch_copy_state: domain sizes differ");
  IDMX_HIP(hipStreamSynchronize(src->ctx->stream));
  hipStream_t st = dst->ctx->stream;
  const bool s32 = src->prec == IDMX_PREC_F32S, d32 = dst->prec == IDMX_PREC_F32S;
  for (int d = 0; d < dst->ndom; d++) {
    const BatchDom& S = src->doms[d];
    const BatchDom& D = dst->doms[d];
    const int64_t V = (int64_t)S.L0 * S.L1;
    const unsigned grid = (unsigned)std::min<int64_t>((V + 255) / 256, 4096);
    const void* in = (const char*)src->pop[src->cur] + src->elem * S.pop_off;
    void* out = (char*)dst->pop[dst->cur] + dst->elem * D.pop_off;
    double* nref = dst->n_prev + D.site_off;
    if (flags == IDMX_BATCH_COPY_REST) {
      if (s32 && !d32)
        hipLaunchKernelGGL((k_batched_rest<float, true, double, false>), dim3(grid), dim3(256),
                           0, st, (const float*)in, (double*)out, V, S.rho0, 0.0, nref);
      else if (!s32 && d32)
        hipLaunchKernelGGL((k_batched_rest<double, false, float, true>), dim3(grid), dim3(256),
                           0, st, (const double*)in, (float*)out, V, 0.0, D.rho0, nref);
      else if (!s32)
        hipLaunchKernelGGL((k_batched_rest<double, false, double, false>), dim3(grid),
                           dim3(256), 0, st, (const double*)in, (double*)out, V, 0.0, 0.0, nref);
      else
        hipLaunchKernelGGL((k_batched_rest<float, true, float, true>), dim3(grid), dim3(256), 0,
                           st, (const float*)in, (float*)out, V, S.rho0, D.rho0, nref);
      IDMX_HIP(hipGetLastError());
      continue;
    }
    if (s32 && !d32)
      hipLaunchKernelGGL((k_batched_convert<float, true, double, false>), dim3(grid), dim3(256),
                         0, st, (const float*)in, (double*)out, V, S.rho0, 0.0);
    else if (!s32 && d32)
      hipLaunchKernelGGL((k_batched_convert<double, false, float, true>), dim3(grid), dim3(256),
                         0, st, (const double*)in, (float*)out, V, 0.0, D.rho0);
    else if (!s32)
      hipLaunchKernelGGL((k_batched_convert<double, false, double, false>), dim3(grid),
                         dim3(256), 0, st, (const double*)in, (double*)out, V, 0.0, 0.0);
    else
      hipLaunchKernelGGL((k_batched_convert<float, true, float, true>), dim3(grid), dim3(256), 0,
                         st, (const float*)in, (float*)out, V, S.rho0, D.rho0);
    IDMX_HIP(hipGetLastError());
    IDMX_HIP(hipMemcpyAsync(nref, src->n_prev + S.site_off, sizeof(double) * V,
                            hipMemcpyDeviceToDevice, st));
  }
  IDMX_HIP(hipStreamSynchronize(st));
  return IDMX_OK;
}

int idmx_batch_symmetrize(idmx_batch_t b, int dom, const double* target_mass) {
  BCHECK(b && dom >= -1 && dom < b->ndom, "idmx_batch_symmetrize: invalid argument");
  hipStream_t st = b->ctx->stream;
  idmx_ctx_s* c = b->ctx;
  for (int d = dom < 0 ? 0 : dom; d < (dom < 0 ? b->ndom : dom + 1); d++) {
    if (dom < 0 && !b->active[d]) continue;
    const BatchDom& D = b->doms[d];
    const int64_t V = (int64_t)D.L0 * D.L1;
    const unsigned grid = (unsigned)std::min<int64_t>((V + 255) / 256, 4096);
    void* pop = (char*)b->pop[b->cur] + b->elem * D.pop_off;
    double* nref = b->n_prev + D.site_off;
    if (b->prec == IDMX_PREC_F64)
      hipLaunchKernelGGL((k_batched_density<double, false>), dim3(grid), dim3(256), 0, st,
                         (const double*)pop, b->stage, V, 0.0);
    else
      hipLaunchKernelGGL((k_batched_density<float, true>), dim3(grid), dim3(256), 0, st,
                         (const float*)pop, b->stage, V, D.rho0);
    IDMX_HIP(hipGetLastError());
    double shift = 0.0;
    if (target_mass) {
      const unsigned rg = (unsigned)std::min<int64_t>((V + 255) / 256, c->partials_cap);
      hipLaunchKernelGGL(k_batched_sum_partial, dim3(rg), dim3(256), 0, st, b->stage, V,
                         c->partials);
      IDMX_HIP(hipGetLastError());
      std::vector<double> h(rg);
      IDMX_HIP(hipMemcpyAsync(h.data(), c->partials, sizeof(double) * rg,
                              hipMemcpyDeviceToHost, st));
      IDMX_HIP(hipStreamSynchronize(st));
      double m = 0.0;
      for (double v : h) m += v;
      shift = (target_mass[d] - m) / (double)V;
    }
    if (b->prec == IDMX_PREC_F64)
      hipLaunchKernelGGL((k_batched_symmetrize<double, false>), dim3(grid), dim3(256), 0, st,
                         b->stage, (double*)pop, D.L0, D.L1, 0.0, shift, nref);
    else
      hipLaunchKernelGGL((k_batched_symmetrize<float, true>), dim3(grid), dim3(256), 0, st,
                         b->stage, (float*)pop, D.L0, D.L1, D.rho0, shift, nref);
    IDMX_HIP(hipGetLastError());
  }
  IDMX_HIP(hipStreamSynchronize(st));
  return IDMX_OK;
}


int idmx_batch_force_rest(idmx_batch_t b, int dom) {
  BCHECK(b && dom >= -1 && dom < b->ndom, "idmx_batch_force_rest: invalid argument");
  hipStream_t st = b->ctx->stream;
  for (int d = dom < 0 ? 0 : dom; d < (dom < 0 ? b->ndom : dom + 1); d++) {
    if (dom < 0 && !b->active[d]) continue;
    const BatchDom& D = b->doms[d];
    const int64_t V = (int64_t)D.L0 * D.L1;
    const unsigned grid = (unsigned)std::min<int64_t>((V + 255) / 256, 4096);
    void* pop = (char*)b->pop[b->cur] + b->elem * D.pop_off;
    const double* nref = b->n_prev + D.site_off;
    switch (b->estencil) {
      case IDMX_E6: launch_force_rest<D2E6>(b, D, pop, nref, grid, st); break;
      case IDMX_E8: launch_force_rest<D2E8>(b, D, pop, nref, grid, st); break;
      default: launch_force_rest<D2E4>(b, D, pop, nref, grid, st); break;
    }
    IDMX_HIP(hipGetLastError());
  }
  IDMX_HIP(hipStreamSynchronize(st));
  return IDMX_OK;
}

int idmx_batch_get_populations(idmx_batch_t b, int dom, double* f_host) {
  BCHECK(b && f_host && dom >= 0 && dom < b->ndom,
         "idmx_batch_get_populations: invalid argument");
  const BatchDom& D = b->doms[dom];
  const int64_t V = (int64_t)D.L0 * D.L1;
  hipStream_t st = b->ctx->stream;
  const void* pop = (const char*)b->pop[b->cur] + b->elem * D.pop_off;
  if (b->prec == IDMX_PREC_F64) {
    IDMX_HIP(hipMemcpyAsync(f_host, pop, sizeof(double) * 9 * V, hipMemcpyDeviceToHost, st));
    IDMX_HIP(hipStreamSynchronize(st));
    return IDMX_OK;
  }
  std::vector<float> h(9 * V);
  IDMX_HIP(hipMemcpyAsync(h.data(), pop, sizeof(float) * 9 * V, hipMemcpyDeviceToHost, st));
  IDMX_HIP(hipStreamSynchronize(st));
  for (int q = 0; q < 9; q++)
    for (int64_t g = 0; g < V; g++)
      f_host[q * V + g] = (double)h[q * V + g] + D2Q9::w[q] * D.rho0;
  return IDMX_OK;
}

int idmx_batch_last_run_ms(idmx_batch_t b, double* ms) {
  BCHECK(b && ms, "idmx_batch_last_run_ms: invalid argument");
  *ms = b->last_ms;
  return IDMX_OK;
}

}  // extern "C"
