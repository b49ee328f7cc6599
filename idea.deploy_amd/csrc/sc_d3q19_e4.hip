// Instantiates the FP64 D3Q19 kernels for the E4 stencil (compiled without
// floating-point contraction).
#include "sc_impl.h"

const SimOps* get_ops_d3q19_e4_f64(int coll) {
  return idmx::ops_for_f64<D3Q19, D3E4>(coll);
}
