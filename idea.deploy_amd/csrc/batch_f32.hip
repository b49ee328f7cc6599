// batch_f32.hip -- FP32-shifted step and initialisation of the batched sweep,
// in a unit of their own so that they are compiled with floating-point
// contraction like the single-domain FP32-shifted kernels.
#include "batch_impl.h"

namespace idmx {

hipError_t batch_step_f32(idmx_batch_s* b, const void* in, void* out) {
  return step_es<float, true>(b, in, out);
}

hipError_t batch_two_steps_f32(idmx_batch_s* b, const void* in, void* out, int xsh) {
  return two_steps_tx<float, true>(b, in, out, xsh);
}

hipError_t batch_init_f32(idmx_batch_s* b, float* pop, const double* n, int64_t V, double rho0,
                          unsigned grid) {
  hipLaunchKernelGGL((k_batched_init<float, true>), dim3(grid), dim3(256), 0, b->ctx->stream,
                     pop, n, V, rho0);
  return hipGetLastError();
}

}  // namespace idmx
