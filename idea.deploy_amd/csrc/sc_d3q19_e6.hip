// Instantiates the D3Q19 kernels for the E6 stencil.
#include "sc_impl.h"

const SimOps* get_ops_d3q19_e6(int precision, int coll) {
  return idmx::ops_for<D3Q19, D3E6>(precision, coll);
}
