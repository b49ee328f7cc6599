// Instantiates the FP32-shifted D3Q19 kernels for the E6 stencil.
#include "sc_impl.h"

const SimOps* get_ops_d3q19_e6_f32(int coll) {
  return idmx::ops_for_f32<D3Q19, D3E6>(coll);
}
