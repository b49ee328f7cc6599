// Instantiates the FP32-shifted D3Q19 kernels for the E8 stencil.
#include "sc_impl.h"

const SimOps* get_ops_d3q19_e8_f32(int coll) {
  return idmx::ops_for_f32<D3Q19, D3E8>(coll);
}
