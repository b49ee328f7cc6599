// Instantiates the D3Q19 kernels for the E4, E6 and E8 stencils.
#include "sc_impl.h"

const SimOps* get_ops_d3q19(int estencil, int precision, int collision) {
  using namespace idmx;
  switch (estencil) {
    case IDMX_E4: return ops_for<D3Q19, D3E4>(precision, collision);
    case IDMX_E6: return ops_for<D3Q19, D3E6>(precision, collision);
    case IDMX_E8: return ops_for<D3Q19, D3E8>(precision, collision);
  }
  return nullptr;
}
