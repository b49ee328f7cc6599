// Instantiates the D3Q19 kernels for the E4 stencil.
#include "sc_impl.h"

const SimOps* get_ops_d3q19_e4(int precision, int coll) {
  return idmx::ops_for<D3Q19, D3E4>(precision, coll);
}
