// internal.h -- context/simulation objects shared by the translation units of
// libidmx.so (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/idmx.h"
#include "sc_kernels.h"

namespace idmx {

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define IDMX_HIP(call)                                        \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return ::idmx::hip_fail(e_, #call); \
  } while (0)

}  // namespace idmx

struct idmx_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  double* partials = nullptr;      // device scratch for reductions
  size_t partials_cap = 0;         // in doubles
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

struct SimOps;

struct idmx_sc_s {
  idmx_ctx_s* ctx = nullptr;
  idmx_sc_desc desc{};
  int D = 0, Q = 0;
  idmx::Geom geo{};
  size_t pop_elem = 8;       // bytes per population value
  void* mrt_dev = nullptr;   // MRT tables in the working precision
  idmx::PsiProg* psi_prog_dev = nullptr;  // psi_kind 3: the compiled psi_code
  uint64_t* wmask_dev = nullptr;  // V wall masks of the fused step with walls
  double last_ms = 0;
  const SimOps* ops = nullptr;
  // fused launch shape (chosen at create time, autotuner may override)
  int tx = 0, ty = 0, H = 0, hist = 1;
  int xsh = 0;  // variant 10: the running step's x-shifted storage flags
  int coll = 0;  // 0 BGK, 1 dense MRT, 2 diagonal MRT, 3 BGK with Shan-Chen forcing,
                 // 4 BGK followed by the isotropy filter
  const uint8_t* walls = nullptr;  // device flags, 1 fluid / 0 solid (borrowed)
  // per-step timing of idmx_sc_run_profiled: 2 events per step + 1, reused
  std::vector<hipEvent_t> prof_ev;
  float* prof_ms = nullptr;  // 2 * nsteps host floats while a profiled run is active
  double bforce[3] = {0.0, 0.0, 0.0};  // body force added to the Shan-Chen force
};

// Per-(velocity set, E-stencil, precision, collision) entry points.
struct SimOps {
  int (*run)(idmx_sc_s*, void** pop, void** swap, double* n, double* psi,
             double* u, int nsteps, int flags);
  int (*density_psi)(idmx_sc_s*, double* n, double* psi, const void* pop);
  int (*fcs)(idmx_sc_s*, void* swap, const void* pop, const double* n,
             const double* psi, double* u);
  int (*moments)(idmx_sc_s*, const void* pop, double* n, double* psi, double* u);
  int (*density_velocity)(idmx_sc_s*, const void* pop, double* n, double* u);
  int (*init_pop)(idmx_sc_s*, void* pop, const double* n, const double* u);
  int (*mass_mom)(idmx_sc_s*, const void* pop, double* mass, double* mom);
  int (*to_f64)(idmx_sc_s*, const void* pop, double* dst);
  int (*from_f64)(idmx_sc_s*, void* pop, const double* src);
  int (*launch_ok)(int tx, int ty, int variant);
  int (*init_pop_tab)(idmx_sc_s*, void* pop, const double* n, const double* u,
                      const double* coeffs);
};

const SimOps* get_ops_d2q9_f64(int estencil, int coll);
const SimOps* get_ops_d2q9_f32(int estencil, int coll);
const SimOps* get_ops_d3q19_e4_f64(int coll);
const SimOps* get_ops_d3q19_e4_f32(int coll);
const SimOps* get_ops_d3q19_e6_f64(int coll);
const SimOps* get_ops_d3q19_e6_f32(int coll);
const SimOps* get_ops_d3q19_e8_f64(int coll);
const SimOps* get_ops_d3q19_e8_f32(int coll);
