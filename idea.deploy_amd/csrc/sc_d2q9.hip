// Instantiates the D2Q9 kernels for the E4, E6 and E8 stencils.
#include "sc_impl.h"

const SimOps* get_ops_d2q9(int estencil, int precision, int coll) {
  using namespace idmx;
  switch (estencil) {
    case IDMX_E4: return ops_for<D2Q9, D2E4>(precision, coll);
    case IDMX_E6: return ops_for<D2Q9, D2E6>(precision, coll);
    case IDMX_E8: return ops_for<D2Q9, D2E8>(precision, coll);
  }
  return nullptr;
}
