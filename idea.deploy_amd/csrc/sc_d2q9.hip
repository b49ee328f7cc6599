// Instantiates the FP64 D2Q9 kernels for the E4, E6 and E8 stencils
// (compiled without floating-point contraction).
#include "sc_impl.h"

const SimOps* get_ops_d2q9_f64(int estencil, int coll) {
  using namespace idmx;
  switch (estencil) {
    case IDMX_E4: return ops_for_f64<D2Q9, D2E4>(coll);
    case IDMX_E6: return ops_for_f64<D2Q9, D2E6>(coll);
    case IDMX_E8: return ops_for_f64<D2Q9, D2E8>(coll);
  }
  return nullptr;
}
