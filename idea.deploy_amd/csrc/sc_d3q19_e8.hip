// Instantiates the D3Q19 kernels for the E8 stencil.
#include "sc_impl.h"

const SimOps* get_ops_d3q19_e8(int precision, int coll) {
  return idmx::ops_for<D3Q19, D3E8>(precision, coll);
}
