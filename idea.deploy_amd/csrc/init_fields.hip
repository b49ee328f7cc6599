// init_fields.hip -- the initial-interface kernels of the simulation classes
// on the device (C ABI: idmx_init_* in include/idmx.h). Each writes the
// density n of every site from its lattice position and zeroes the velocity
// u, as the reference kernels do; InitPopulations then sets f = f_eq(n, u).
// Integer coordinates and the reference's integer arithmetic (SType), double
// for lengths and densities (LengthType, NType).
#include "internal.h"

namespace {

struct Box {
  int D;
  int L[3];
  int c[3];  // dim_center (L_d / 2, idpy/LBM/LBM.py:422-423)
  int64_t V;
};

__device__ inline void pos_from_index(const Box& b, int64_t g, int* p) {
  p[0] = (int)(g % b.L[0]);
  const int64_t r = g / b.L[0];
  p[1] = b.D > 1 ? (int)(r % b.L[1]) : 0;
  p[2] = b.D > 2 ? (int)(r / b.L[1]) : 0;
}

// delta_n = full_flag (n_l - n_g) + (1 - full_flag) (n_g - n_l)
__device__ inline double delta_n(double n_g, double n_l, int full) {
  return full * (n_l - n_g) + (1 - full) * (n_g - n_l);
}

// K_InitRadialInterface: n = (n_g + n_l)/2 - delta_n/2 tanh((r - R) / W),
// r the distance to dim_center from integer coordinate differences
// (F_PointDistanceCenterFirst, idpy/LBM/MultiPhaseKernels.py:99-112; W = 1
// for the Meta class, :686-727; legacy class idpy/LBM/LBM.py:1786-1826).
__global__ void __launch_bounds__(256)
k_init_radial(double* __restrict__ n, double* __restrict__ u, Box b, double n_g, double n_l,
              double R, double W, int full) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < b.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    int p[3];
    pos_from_index(b, g, p);
    double dist = 0;
    for (int d = 0; d < b.D; d++) dist += (b.c[d] - p[d]) * (b.c[d] - p[d]);
    const double r = sqrt(dist);
    n[g] = 0.5 * (n_g + n_l) - 0.5 * delta_n(n_g, n_l, full) * tanh((r - R) / W);
    for (int d = 0; d < b.D; d++) u[g + d * b.V] = 0.;
  }
}

// K_InitFlatInterface (idpy/LBM/LBM.py:1745-1784) with the periodic profile
// F_NFlatProfilePeriodic (idpy/LBM/MultiPhaseKernels.py:55-74):
// n = (n_g + n_l)/2 + delta_n/2 (prof - 1), a slab of width `width` centred
// on dim_center along `direction`.
__global__ void __launch_bounds__(256)
k_init_flat(double* __restrict__ n, double* __restrict__ u, Box b, double n_g, double n_l,
            double width, int dir, int full) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < b.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    int p[3];
    pos_from_index(b, g, p);
    const int x = p[dir], x0 = b.c[dir], size = b.L[dir];
    const int xm = size / 2;
    int xp;
    if (x0 - xm >= 0) xp = x < x0 - xm ? x - (x0 - xm) + size : x - (x0 - xm);
    else xp = x < x0 + xm ? x - (x0 + xm) + size : x - (x0 + xm);
    const double prof = tanh((double)(xp - (xm - 0.5 * width))) -
                        tanh((double)(xp - (xm + 0.5 * width)));
    n[g] = 0.5 * (n_g + n_l) + 0.5 * delta_n(n_g, n_l, full) * (prof - 1.);
    for (int d = 0; d < b.D; d++) u[g + d * b.V] = 0.;
  }
}

// K_InitCylinderInterface (idpy/LBM/LBM.py:1828-1874): the radial profile
// with W = 1 and the distance taken in the plane orthogonal to `direction`.
__global__ void __launch_bounds__(256)
k_init_cylinder(double* __restrict__ n, double* __restrict__ u, Box b, double n_g, double n_l,
                double R, int dir, int full) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < b.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    int p[3];
    pos_from_index(b, g, p);
    double dist = 0;
    for (int d = 0; d < b.D; d++)
      if (d != dir) dist += (b.c[d] - p[d]) * (b.c[d] - p[d]);
    const double r = sqrt(dist);
    n[g] = 0.5 * (n_g + n_l) - 0.5 * delta_n(n_g, n_l, full) * tanh(r - R);
    for (int d = 0; d < b.D; d++) u[g + d * b.V] = 0.;
  }
}

// K_InitSingleFlatInterface (idpy/LBM/MultiPhaseKernels.py:650-684) with
// F_NSingleFlatProfile (:45-53): n = n_g + delta_n (tanh(x - x0) + 1)/2.
__global__ void __launch_bounds__(256)
k_init_single_flat(double* __restrict__ n, double* __restrict__ u, Box b, double n_g,
                   double n_l, int dir, int full) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < b.V;
       g += (int64_t)gridDim.x * blockDim.x) {
    int p[3];
    pos_from_index(b, g, p);
    n[g] = n_g + delta_n(n_g, n_l, full) * ((tanh((double)(p[dir] - b.c[dir])) + 1.) * 0.5);
    for (int d = 0; d < b.D; d++) u[g + d * b.V] = 0.;
  }
}

int make_box(int D, const int* dims, Box* b, const char* who) {
  if (D < 2 || D > 3 || !dims) {
    idmx::set_error(std::string(who) + ": D must be 2 or 3");
    return IDMX_E_ARG;
  }
  b->D = D;
  b->V = 1;
  for (int d = 0; d < 3; d++) {
    b->L[d] = d < D ? dims[d] : 1;
    if (b->L[d] < 1) {
      idmx::set_error(std::string(who) + ": lattice sizes must be positive");
      return IDMX_E_ARG;
    }
    b->c[d] = d < D ? b->L[d] / 2 : 0;
    b->V *= b->L[d];
  }
  return IDMX_OK;
}

unsigned grid_of(int64_t V) {
  const int64_t g = (V + 255) / 256;
  return (unsigned)(g < 65536 ? g : 65536);
}

}  // namespace

extern "C" {

int idmx_init_radial(idmx_ctx_t c, double* n, double* u, int D, const int* dims, double n_g,
                     double n_l, double R, double W, int full_flag) {
  Box b;
  if (!c || !n || !u || !(W > 0)) {
    idmx::set_error("idmx_init_radial: invalid argument");
    return IDMX_E_ARG;
  }
  if (int rc = make_box(D, dims, &b, "idmx_init_radial")) return rc;
  hipLaunchKernelGGL(k_init_radial, dim3(grid_of(b.V)), dim3(256), 0, c->stream, n, u, b, n_g,
                     n_l, R, W, full_flag ? 1 : 0);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_init_flat(idmx_ctx_t c, double* n, double* u, int D, const int* dims, double n_g,
                   double n_l, double width, int direction, int full_flag) {
  Box b;
  if (!c || !n || !u || direction < 0 || direction >= D) {
    idmx::set_error("idmx_init_flat: invalid argument");
    return IDMX_E_ARG;
  }
  if (int rc = make_box(D, dims, &b, "idmx_init_flat")) return rc;
  hipLaunchKernelGGL(k_init_flat, dim3(grid_of(b.V)), dim3(256), 0, c->stream, n, u, b, n_g, n_l,
                     width, direction, full_flag ? 1 : 0);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_init_cylinder(idmx_ctx_t c, double* n, double* u, int D, const int* dims, double n_g,
                       double n_l, double R, int direction, int full_flag) {
  Box b;
  if (!c || !n || !u || direction < 0 || direction >= D) {
    idmx::set_error("idmx_init_cylinder: invalid argument");
    return IDMX_E_ARG;
  }
  if (int rc = make_box(D, dims, &b, "idmx_init_cylinder")) return rc;
  hipLaunchKernelGGL(k_init_cylinder, dim3(grid_of(b.V)), dim3(256), 0, c->stream, n, u, b, n_g,
                     n_l, R, direction, full_flag ? 1 : 0);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

int idmx_init_single_flat(idmx_ctx_t c, double* n, double* u, int D, const int* dims, double n_g,
                          double n_l, int direction, int full_flag) {
  Box b;
  if (!c || !n || !u || direction < 0 || direction >= D) {
    idmx::set_error("idmx_init_single_flat: invalid argument");
    return IDMX_E_ARG;
  }
  if (int rc = make_box(D, dims, &b, "idmx_init_single_flat")) return rc;
  hipLaunchKernelGGL(k_init_single_flat, dim3(grid_of(b.V)), dim3(256), 0, c->stream, n, u, b,
                     n_g, n_l, direction, full_flag ? 1 : 0);
  IDMX_HIP(hipGetLastError());
  return IDMX_OK;
}

}  // extern "C"
